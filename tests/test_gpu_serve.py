"""Serving engine (apsm_server_*) on the GPU.

* A frame's result is the same alone and in the mix: links served together (frames arriving
  while others are on the GPU, sharing batched rounds) give bitwise the estimates each link
  gets when served alone, and those of a plain handle doing detect then train per frame (same
  kernels, same order).
* Against the FP64 oracle (detect, then incremental train, per frame): the parity tolerance of
  test_gpu_parity (soft outputs within 2e-3, decisions equal away from boundaries).
* Admission: links by HBM budget (APSM_ERR_BUSY), frames by arena room (they wait, results
  unchanged); the submit/poll interface gives the replay's results.
* Frames in pinned host memory (moved by the server's upload kernel, estimates written by the
  detect into host memory) give bitwise the results of the same frames in pageable memory
  (engine copies), with 16-byte aligned and unaligned buffers.
"""
import numpy as np
import pytest

from helpers import to_ri
from test_gpu_parity import compare

pytestmark = pytest.mark.gpu

MB = 1 << 20


def trace(n_links, n_frames, rate_hz, seed, max_rb=12):
    from libapsm_amd import datagen
    fr = datagen.serving_frames(n_frames, seed=seed, links=n_links, rate_hz=rate_hz, max_rb=max_rb)
    for f in fr:
        f.update(datagen.serving_payload(f))
    return fr, datagen.serving_links(n_links, seed)


def make_server(links, cfg_of, budget=2048 * MB, arena=256 * MB, max_pilots=512):
    from libapsm_amd import ApsmServer
    srv = ApsmServer(budget, arena)
    for lp in links:
        srv.add_link(cfg_of(lp), lp["N"], 1, max_pilots)
    return srv


def cfg_of(lp):
    from libapsm_amd import datagen
    return datagen.serving_link_config(lp, window=16, dict_size=300)


def frames_of(fr, keep=None, relink=None, t0=0.0):
    from libapsm_amd import Frame
    out = []
    for f in fr:
        if keep is not None and f["link"] not in keep:
            continue
        link = f["link"] if relink is None else relink[f["link"]]
        out.append(Frame(link, f["rx_data"], f["rx_pilots"], f["tx_pilots"], arrival_s=f["arrival_s"] - t0))
    return out


def plain_run(lp, frames):
    from libapsm_amd import ApsmNomaDetectorFactory
    det = ApsmNomaDetectorFactory.build(dict(cfg_of(lp), otype="APSM"), lp["N"])
    out = []
    for f in frames:
        out.append(det.detect(f["rx_data"]).copy())
        det.train(f["rx_pilots"], f["tx_pilots"][0])
    return out


def test_result_is_the_same_alone_and_in_the_mix(require_gpu):
    fr, links = trace(6, 48, rate_hz=20000.0, seed=11)
    print("frames per link", np.bincount([f["link"] for f in fr], minlength=6), flush=True)
    srv = make_server(links, cfg_of)
    mix = frames_of(fr)
    res = srv.replay(mix)
    assert all(r["status"] == 0 for r in res)
    st = srv.stats()
    print("mix stats", st, "p50 latency ms", 1e3 * np.median([r["latency_s"] for r in res]), flush=True)
    assert st["max_frames_per_round"] > 1  # frames really shared rounds
    for link in (0, 3, 5):
        own = [f for f in fr if f["link"] == link]
        if not own:
            continue
        alone_srv = make_server([links[link]], cfg_of)
        alone = frames_of(fr, keep={link}, relink={link: 0})
        alone_srv.replay(alone)
        mixed = [m for m in mix if m.link == link]
        plain = plain_run(links[link], own)
        for k, (a, m, p) in enumerate(zip(alone, mixed, plain)):
            np.testing.assert_array_equal(m.est[0], a.est[0], err_msg=f"link {link} frame {k}: mix vs alone")
            np.testing.assert_array_equal(m.est[0], p, err_msg=f"link {link} frame {k}: served vs plain handle")


def test_served_link_matches_oracle(require_gpu, oracle):
    fr, links = trace(3, 30, rate_hz=5000.0, seed=12, max_rb=20)
    srv = make_server(links, cfg_of)
    frames = frames_of(fr)
    srv.replay(frames)
    link = max(range(3), key=lambda l: sum(f["link"] == l for f in fr))
    o = oracle.Apsm(cfg_of(links[link]), links[link]["N"])
    n = 0
    for f, fs in zip(fr, frames):
        if f["link"] != link:
            continue
        ref = o.detect(f["rx_data"]) if n > 0 else np.zeros((f["samples"], 2))
        compare(fs.est[0], ref, f["bps"])
        o.train(f["rx_pilots"], f["tx_pilots"][0], incremental=True)
        n += 1
    print(f"link {link}: {n} frames within tolerance", flush=True)
    assert n >= 5
    st = srv.link_state(link)
    assert st["used_rows"] == 2 * sum(f["pilots"] for f in fr if f["link"] == link)


def test_links_are_admitted_by_hbm_budget(require_gpu):
    from libapsm_amd import ApsmError, ApsmServer
    from libapsm_amd.serve import APSM_ERR_BUSY
    links = [dict(N=16, modulation="QPSK")] * 64
    srv = ApsmServer(24 * MB, 8 * MB)
    admitted = 0
    with pytest.raises(ApsmError) as e:
        for lp in links:
            srv.add_link(cfg_of(lp), lp["N"], 1, 512)
            admitted += 1
    assert e.value.code == APSM_ERR_BUSY
    per_link = srv.stats()["link_bytes"] / admitted
    print(f"{admitted} links admitted, {per_link / 1024:.0f} KiB each", flush=True)
    assert admitted >= 1 and srv.stats()["link_bytes"] + 8 * MB <= 24 * MB
    assert srv.stats()["link_bytes"] + per_link + 8 * MB > 24 * MB


def test_frames_wait_for_arena_room(require_gpu):
    fr, links = trace(4, 24, rate_hz=50000.0, seed=13)
    big = make_server(links, cfg_of)
    a = frames_of(fr)
    big.replay(a)
    largest = max(8 * (f["N"] * (f["samples"] + f["pilots"]) + f["samples"] + f["pilots"]) for f in fr)
    small = make_server(links, cfg_of, arena=largest + 4096)  # about one frame at a time
    b = frames_of(fr)
    res = small.replay(b)
    waited = sum(r["admit_s"] > r["arrival_s"] + 1e-4 for r in res)
    print(f"frames admitted after waiting: {waited} of {len(res)}", flush=True)
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x.est[0], y.est[0])


def test_submit_poll_gives_the_replay_results(require_gpu):
    fr, links = trace(4, 20, rate_hz=10000.0, seed=14)
    a = frames_of(fr)
    make_server(links, cfg_of).replay(a)
    srv = make_server(links, cfg_of)
    b = frames_of(fr)
    for f in b:
        srv.submit(f)
    done = srv.poll(wait=True)
    assert len(done) == len(b) and all(f.result["status"] == 0 for f in b)
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x.est[0], y.est[0])
    assert all(f.result["output_s"] >= f.result["arrival_s"] for f in b)


def test_frame_without_pilots_and_errors(require_gpu):
    from libapsm_amd import ApsmError, Frame
    fr, links = trace(2, 6, rate_hz=1000.0, seed=15)
    srv = make_server(links, cfg_of, max_pilots=64)
    f0 = fr[0]
    with pytest.raises(ApsmError):
        srv.submit(Frame(f0["link"], f0["rx_data"], np.zeros((f0["N"], 100), np.complex64), np.zeros((1, 100), np.complex64)))
    with pytest.raises(ValueError):
        srv.submit(Frame(7, f0["rx_data"]))
    frames = [Frame(f["link"], f["rx_data"], arrival_s=f["arrival_s"]) for f in fr]
    res = srv.replay(frames)
    assert all(r["status"] == 0 for r in res)
    assert all(np.all(f.est == 0) for f in frames)  # no train yet: zero weights, empty dictionary


def pinned_copy(a, offset_elems=0):
    """A pinned copy of a complex64 array; offset_elems > 0 places it 8*offset bytes into the
    allocation (so the buffer is not 16-byte aligned)."""
    from libapsm_amd import PinnedArray
    a = np.ascontiguousarray(a, dtype=np.complex64)
    pa = PinnedArray((a.size + offset_elems,), np.complex64)
    view = pa.array[offset_elems:].reshape(a.shape)
    view[...] = a
    return pa, view


def test_pinned_frames_give_the_pageable_results(require_gpu):
    from libapsm_amd import Frame
    fr, links = trace(4, 24, rate_hz=20000.0, seed=16)
    a = frames_of(fr)
    make_server(links, cfg_of).replay(a)
    keep = []
    b = []
    for k, f in enumerate(fr):
        off = k % 2  # odd frames: every buffer 8 bytes off 16-byte alignment
        rxd, rxd_v = pinned_copy(f["rx_data"], off)
        rxp, rxp_v = pinned_copy(f["rx_pilots"], off)
        txp, txp_v = pinned_copy(f["tx_pilots"], off)
        est, est_v = pinned_copy(np.zeros((1, f["samples"]), np.complex64), off)
        keep += [rxd, rxp, txp, est]
        b.append(Frame(f["link"], rxd_v, rxp_v, txp_v, arrival_s=f["arrival_s"], est=est_v))
    res = make_server(links, cfg_of).replay(b)
    assert all(r["status"] == 0 for r in res)
    for k, (x, y) in enumerate(zip(a, b)):
        np.testing.assert_array_equal(y.est[0], x.est[0], err_msg=f"frame {k}: pinned vs pageable")
    assert any(np.any(y.est != 0) for y in b)


def test_bad_late_frame_fails_before_any_work(require_gpu):
    """A replay whose last frame cannot be served fails before anything is enqueued; the server
    stays idle and serves the good frames afterwards with the results of a fresh server."""
    from libapsm_amd import ApsmError, Frame
    fr, links = trace(3, 12, rate_hz=20000.0, seed=17)
    srv = make_server(links, cfg_of, max_pilots=512)
    good = frames_of(fr)
    f0 = fr[0]
    bad = Frame(f0["link"], f0["rx_data"], np.zeros((f0["N"], 600), np.complex64), np.zeros((1, 600), np.complex64),
                arrival_s=fr[-1]["arrival_s"] + 1e-3)
    with pytest.raises(ApsmError):
        srv.replay(frames_of(fr) + [bad])
    assert srv.stats()["rounds"] == 0  # nothing was launched
    res = srv.replay(good)
    assert all(r["status"] == 0 for r in res)
    ref = frames_of(fr)
    make_server(links, cfg_of, max_pilots=512).replay(ref)
    for k, (x, y) in enumerate(zip(ref, good)):
        np.testing.assert_array_equal(y.est[0], x.est[0], err_msg=f"frame {k}")


def test_poll_without_wait_gives_the_replay_results(require_gpu):
    """Frames in pageable buffers reported by non-waiting polls hold their final estimates."""
    import time
    fr, links = trace(4, 20, rate_hz=10000.0, seed=18)
    a = frames_of(fr)
    make_server(links, cfg_of).replay(a)
    srv = make_server(links, cfg_of)
    b = frames_of(fr)
    for f in b:
        srv.submit(f)
    done, seen = [], {}
    t_end = time.time() + 60.0
    while len(done) < len(b) and time.time() < t_end:
        for f in srv.poll(wait=False):
            seen[id(f)] = f.est[0].copy()  # read right after the poll reported it
            done.append(f)
    assert len(done) == len(b)
    for k, (x, y) in enumerate(zip(a, b)):
        np.testing.assert_array_equal(seen[id(y)], x.est[0], err_msg=f"frame {k}")
