"""Config 4 on the GPU: N = 10^6 spheres (a = 1) at volume fraction 0.2
(L = 275.65), seeded random-sequential-addition positions, SPME mesh 256^3,
p = 8, rc = 5, xi = 0.896, block Lanczos b = 16 (tol 1e-6) on 16 seeded
Gaussian right-hand sides. FP32-mixed within 1e-5 of FP64 for M f and the
displacements, operator symmetry, the iteration count; FP64 against the
oracle at K = 256, p = 8 on a smaller system (the oracle's FP64 CPU mesh
transforms at N = 10^6 take minutes).
"""
import numpy as np
import pytest

from conftest import box_for

pytestmark = pytest.mark.gpu

N4, PHI4, K4, P4, XI, RC, B4 = 1_000_000, 0.2, 256, 8, 0.896, 5.0, 16


@pytest.fixture(scope="module")
def sys4(sdt):
    L = box_for(N4, PHI4)
    print("config 4: generating %d positions, L = %.3f" % (N4, L), flush=True)
    pos = sdt.rsa_positions(N4, L, 1.0, 4)
    z = sdt.gaussian(4, 0, 3 * N4 * B4).reshape(B4, 3 * N4).T.copy()
    return pos, np.ones(N4), L, z


def test_config4_fp32_vs_fp64(sdt, orc, sys4):
    pos, rdi, L, z = sys4
    assert abs(L - 275.65) < 0.01
    res = {}
    for prec in (sdt.FP64, sdt.FP32):
        print("config 4: precision %d" % prec, flush=True)
        m = sdt.MobSpme(N4, rdi, L, XI, RC, K4, P4, prec)
        assert m.init(), getattr(m, "init_error", "")
        m.update(pos)
        mf = m.mul_vector(z)
        bl = sdt.BrwnLanczos(3 * N4, 100, B4, 1e-6)
        assert bl.init()
        y, ok = bl.compute(m, z)
        assert ok
        it = list(bl.iters())
        print("config 4: precision %d Lanczos iterations %s" % (prec, it), flush=True)
        assert 0 < it[0] < 100
        if prec == sdt.FP64:
            x = z[:, 0].copy()
            w = z[:, 1].copy()
            mx = m.mul_vector(x)[:, 0]
            mw = m.mul_vector(w)[:, 0]
            asym = abs(x @ mw - mx @ w) / abs(x @ mw)
            print("config 4: symmetry %.3e" % asym, flush=True)
            assert asym < 1e-12
        res[prec] = (mf, y)
        del m, bl
    e_mf = orc.frob_rel_err(res[sdt.FP64][0], res[sdt.FP32][0])
    e_d = orc.frob_rel_err(res[sdt.FP64][1], res[sdt.FP32][1])
    print("config 4: FP32 vs FP64: M f %.3e, displacements %.3e" % (e_mf, e_d), flush=True)
    assert e_mf < 1e-5 and e_d < 1e-5


def test_config4_parameters_fp64_vs_oracle(sdt, orc):
    # K = 256, p = 8 on 20 000 particles at the same volume fraction
    n = 20_000
    L = box_for(n, PHI4)
    pos = sdt.rsa_positions(n, L, 1.0, 44)
    rdi = np.ones(n)
    f = np.random.default_rng(4).standard_normal((3 * n, 2))
    g = sdt.MobSpme(n, rdi, L, XI, RC, K4, P4, sdt.FP64)
    assert g.init()
    g.update(pos)
    o = orc.MobSpme(n, rdi, L, XI, RC, K4, P4)
    assert o.init()
    o.update(pos, rdi)
    assert orc.frob_rel_err(o.mul_vector(f), g.mul_vector(f)) < 1e-10
