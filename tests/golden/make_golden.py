"""Writes tests/golden/c1_uniform_f32_2p20.json: the config-1 known answers.

Input: n = 2^20 float32 U[0,1) on the 24-bit grid from SplitMix64 seed 0
(ggks-multiselect_amd/workloads.py:uniform_f32); k = round(p (n-1)) + 1 for
p = 0.00 .. 1.00 (k-th largest; p = 0 is the maximum).  The expected values
come from numpy's sort (independent of the C oracle), indexed at n - k
(multiselectTimingFunctions.cu:41-48).  The input's SHA-256 pins the
generator.
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
from __graft_entry__ import load_package  # noqa: E402

W = load_package().workloads
n = 1 << 20
x = W.uniform_f32(0, n)
k = W.k_percentiles(n)
s = np.sort(x)  # no zeros of either sign matter here: U[0,1) has +0 only
vals = s[n - k.astype(np.int64)]
doc = {
    "generator": "workloads.uniform_f32(seed=0, n=2^20); k = workloads.k_percentiles(n)",
    "n": n,
    "input_sha256": hashlib.sha256(x.tobytes()).hexdigest(),
    "k": k.tolist(),
    "expected_hex": [f"{int(v):08x}" for v in vals.view(np.uint32)],
}
with open(os.path.join(HERE, "c1_uniform_f32_2p20.json"), "w") as f:
    json.dump(doc, f, indent=0)
print("wrote", len(vals), "values; max", vals[0], "min", vals[-1])
