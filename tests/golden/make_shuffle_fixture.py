"""Writes tests/golden/shuffle_perm.json: the head of the permutation that the reference's
training shuffle applies (std::shuffle with std::mt19937(seed_seq{1..8}),
cpp/lib/noma/noma_detector.cu:59-74), as produced by this image's libstdc++ through the
oracle. Pins the permutation against changes of the standard library."""
import json
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from oracle import pyoracle  # noqa: E402

out = {str(T): pyoracle.shuffle_perm(T)[:16].tolist() for T in (10, 685, 20000)}
with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "shuffle_perm.json"), "w") as f:
    json.dump(out, f, indent=1)
print(out)
