"""The oracle reproduces the committed golden vectors (tests/golden/
golden_small.npz, written by tests/golden/make_golden.py)."""
import os

import numpy as np

G = np.load(os.path.join(os.path.dirname(__file__), "golden", "golden_small.npz"))


def test_golden_mono_and_poly(orc):
    for tag in ("mono", "poly"):
        box, xi, rmax, K, p = G[tag + "_params"]
        pos, rdi, f = G[tag + "_pos"], G[tag + "_rdi"], G[tag + "_f"]
        m = orc.MobSpme(rdi.size, rdi, box, xi, rmax, int(K), int(p))
        assert m.init()
        m.update(pos, rdi)
        for mode in ("full", "real", "recip"):
            if tag == "mono":
                out = m.mul_vector(f, 2.1, 1.1, G["mono_v0"], mode=mode)
            else:
                out = m.mul_vector(f, mode=mode)
            assert orc.frob_rel_err(G[tag + "_" + mode], out) < 1e-13


def test_golden_pairlists_and_steric(orc):
    pos, rdi = G["pl_pos"], G["pl_rdi"]
    for mode, cut in zip("RS", G["pl_cutoffs"]):
        ia, ja = orc.pairlist(pos, rdi, 20.0, cut, mode)
        assert np.array_equal(ia, G["pl_%s_ia" % mode])
        assert np.array_equal(ja, G["pl_%s_ja" % mode])
    f = orc.steric(pos, rdi, 20.0, 2.2, 125.0)
    assert np.abs(f - G["steric_f"]).max() <= 1e-12 * np.abs(G["steric_f"]).max()


def test_golden_gaussian(orc):
    assert np.abs(orc.gaussian(1234, 5, 16) - G["gauss_1234_5"]).max() == 0.0
