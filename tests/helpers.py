"""Shared test helpers: small seeded problems and the bundled reference data."""
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
OFFLINE = os.path.join(GOLDEN, "offline")


def small_problem(N=4, U=3, T=150, M=120, seed=0, mod="QPSK", snr_db=15.0):
    from libapsm_amd import datagen
    return datagen.generate(N, U, T, M, modulation=mod, snr_db=snr_db, seed=seed)


def to_ri(z):
    z = np.asarray(z)
    return np.stack([z.real, z.imag], axis=-1).astype(np.float32)


def load_offline(mod="QPSK", domain="time", antennas=None):
    """Bundled reference data (cpp/data/offline): rx training/data [N][S], tx training/data [U][S]
    scaled to unit energy as the reference CLI does (argparse_helper.cpp:77-105)."""
    from libapsm_amd import binfile, datagen
    bps = datagen.BITS_PER_SYMBOL[mod]
    rx = binfile.read(os.path.join(OFFLINE, f"rxData_{mod}_{domain}.bin"))
    tx = binfile.read(os.path.join(OFFLINE, f"NOMA_signals_{mod.lower()}_complex.bin"))
    s = datagen.mod_scale(bps)
    out = dict(rx_train=binfile.to_complex(rx["rxSigTraining"]), rx_data=binfile.to_complex(rx["rxSigData"]),
               tx_train=binfile.to_complex(tx["mod_training"], s), tx_data=binfile.to_complex(tx["mod_data"], s), bps=bps)
    if antennas is not None:
        out["rx_train"] = np.ascontiguousarray(out["rx_train"][antennas])
        out["rx_data"] = np.ascontiguousarray(out["rx_data"][antennas])
    return out


def equidistant(n, max_antennas=16):
    """Antenna indices of the reference's equidistant pattern (util.cpp:262-272)."""
    return [i * max_antennas // n for i in range(n)]
