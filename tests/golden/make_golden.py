"""Writes the committed golden vectors (tests/golden/golden_small.npz) from
the oracle on seeded small systems. Run from the repository root:

    python tests/golden/make_golden.py

Inputs are generated here from seeds; outputs are the oracle's FP64 results
(the oracle itself is pinned by tests/test_oracle_*.py).
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from oracle import oracle as O  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden_small.npz")


def rsa(n, box, seed, dmin=2.0):
    rng = np.random.default_rng(seed)
    pts = []
    while len(pts) < n:
        c = rng.uniform(0, box, 3)
        if all(np.linalg.norm(np.remainder(c - q + box / 2, box) - box / 2) >= dmin for q in pts):
            pts.append(c)
    return np.concatenate(pts)


def main():
    g = {}
    rng = np.random.default_rng(2026)
    # mono SPME: N=64, L=12, K=24, p=6
    n, box = 64, 12.0
    pos = rng.uniform(0, box, 3 * n)
    rdi = np.ones(n)
    f = rng.uniform(0, 1, (3 * n, 3))
    v0 = rng.uniform(0, 1, (3 * n, 3))
    m = O.MobSpme(n, rdi, box, 0.9, 5.0, 24, 6)
    assert m.init()
    m.update(pos, rdi)
    g["mono_pos"], g["mono_rdi"], g["mono_f"], g["mono_v0"] = pos, rdi, f, v0
    g["mono_params"] = np.array([box, 0.9, 5.0, 24, 6])
    for mode in ("full", "real", "recip"):
        g["mono_" + mode] = m.mul_vector(f, 2.1, 1.1, v0, mode=mode)
    # poly SPME with odd order: N=48, L=14, K=32, p=5
    n, box = 48, 14.0
    pos = rng.uniform(0, box, 3 * n)
    rdi = rng.uniform(0.7, 1.3, n)
    f = rng.standard_normal((3 * n, 2))
    m = O.MobSpme(n, rdi, box, 1.1, 4.5, 32, 5)
    assert m.init()
    m.update(pos, rdi)
    g["poly_pos"], g["poly_rdi"], g["poly_f"] = pos, rdi, f
    g["poly_params"] = np.array([box, 1.1, 4.5, 32, 5])
    for mode in ("full", "real", "recip"):
        g["poly_" + mode] = m.mul_vector(f, 1.0, 0.0, None, mode=mode)
    # pair lists: N=300, L=20
    n, box = 300, 20.0
    pos = rng.uniform(0, box, 3 * n)
    rdi = rng.uniform(0.6, 1.4, n)
    g["pl_pos"], g["pl_rdi"] = pos, rdi
    for mode, cut in (("R", 4.0), ("S", 2.2)):
        ia, ja = O.pairlist(pos, rdi, box, cut, mode)
        g["pl_%s_ia" % mode], g["pl_%s_ja" % mode] = ia, ja
    g["pl_cutoffs"] = np.array([4.0, 2.2])
    # steric on the same system
    g["steric_f"] = O.steric(pos, rdi, box, 2.2, 125.0)
    # Lanczos on the mono SPME operator (5 columns, block 4 then single)
    n, box = 64, 12.0
    pos, rdi = g["mono_pos"], g["mono_rdi"]
    m = O.MobSpme(n, rdi, box, 0.9, 5.0, 24, 6)
    assert m.init()
    m.update(pos, rdi)
    z = O.gaussian(77, 0, 3 * n * 5).reshape(5, 3 * n).T.copy()
    y, it = O.lanczos(m, z, 100, 4, 1e-10)
    g["lz_z"], g["lz_y"], g["lz_iters"] = z, y, it
    # BD: N=40 RSA at L=15, 6 steps, mob interval 4
    n, box = 40, 15.0
    pos = rsa(n, box, 31)
    rdi = np.ones(n)
    bd = O.BDSim(pos, rdi, box, 0.9, 5.0, 24, 6, 100, 4, 1e-10, 2.0, 125.0, 1e-3, 4, 555)
    bd.advance(6)
    g["bd_pos0"], g["bd_pos6"] = pos, bd.positions()
    g["bd_params"] = np.array([box, 0.9, 5.0, 24, 6, 100, 4, 1e-10, 2.0, 125.0, 1e-3, 4, 555])
    g["gauss_1234_5"] = O.gaussian(1234, 5, 16)
    np.savez_compressed(OUT, **g)
    print("wrote", OUT, os.path.getsize(OUT), "bytes")


if __name__ == "__main__":
    main()
