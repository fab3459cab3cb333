#!/usr/bin/env python3
"""Writes tests/golden/soup_10k_64.npz: the oracle's results on the seeded random-soup
workload (10,000 triangles, 64^3 grid, 6 levels, 256x256 G-buffer), to pin regressions
of both the oracle and the HIP path against committed data.

The reference ships no fixtures for this path (it has no tests), so these vectors come
from this repository's oracle; the oracle itself is pinned by closed forms and
independent methods in tests/test_oracle_*.py.

Usage: python tests/golden/make_golden.py
"""
import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from final184_amd import scenegen as sg  # noqa: E402
from oracle import oracle as orc  # noqa: E402

OUT = os.path.join(HERE, "soup_10k_64.npz")
PIXELS = 256


def sha(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def compute():
    w = sg.random_soup()
    acc, st = orc.voxelize(w.positions, w.grid, albedo=w.albedo)
    out = {
        "acc_sha256": sha(acc),
        "acc_word_sums": acc.reshape(-1, 8).astype(np.uint64).sum(0),
        "occupied": np.int64((acc[..., 3] > 0).sum()),
        "voxel_hits": np.int64(st.voxel_hits),
    }
    for prec in ("fp32", "fp16", "bf16"):
        lv = orc.inject_mips(acc, w.grid, w.lights, prec)
        out[f"{prec}_level_sha256"] = np.array([sha(l) for l in lv])
        out[f"{prec}_level_sums"] = np.array([l.astype(np.float64).sum(axis=(0, 1, 2)) for l in lv])
    px = np.random.default_rng(2024).choice(w.num_pixels, PIXELS, replace=False)
    l64 = orc.inject_mips(acc, w.grid, w.lights, f64=True)
    d64, s64, _ = orc.trace(l64, w.grid, w.gb_pos, w.gb_nrm, w.cones, pixels=px, f64=True)
    out["pixels"] = px.astype(np.int64)
    out["diffuse_f64"] = d64
    out["specular_f64"] = s64
    return out


if __name__ == "__main__":
    np.savez_compressed(OUT, **compute())
    print("wrote", OUT, os.path.getsize(OUT), "bytes")
