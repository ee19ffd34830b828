"""Config 2 on the GPU at its full size: N = 10^5 spheres (a = 1) at volume
fraction 0.3 (L = 111.7693), seeded random-sequential-addition positions,
SPME mesh 128^3, p = 6, rc = 5, xi = 0.896, block Lanczos b = 8 at tol 1e-6,
mobility interval 100, dt = 1e-3, steric r0 = 2, k0 = 125.

- Block Lanczos (BrwnLanczos::Compute, src/brwn/brwn_lanczos.cc:16-304) on
  one group of 8 seeded Gaussian columns: FP64 against the oracle's block
  Lanczos on the same operator (iteration counts within 3, iterates within
  the 1e-5 gate: after 50 iterations without reorthogonalisation rounding
  differences move the iterate by a fraction of tol 1e-6), FP32-mixed within
  1e-5 of FP64, and per column |y|^2 = z^T M z (y = M^{1/2} z) within 1e-4,
  with M z from the FP64 mobility.
- BD steps (BDSimBox::Advance, src/simbox/simbox_bd.cc:296-366) from the
  same start in FP64 and FP32-mixed: the first step runs the mobility
  update and the Lanczos solve of the interval's 100 columns (13 groups);
  the displacements after 3 steps agree within 1e-5 and every group
  converges.
"""
import numpy as np
import pytest

from conftest import box_for

pytestmark = pytest.mark.gpu

N2, PHI2, K2, P2, XI, RC, B2, TOL = 100_000, 0.3, 128, 6, 0.896, 5.0, 8, 1e-6


@pytest.fixture(scope="module")
def sys2(sdt):
    L = box_for(N2, PHI2)
    print("config 2: generating %d positions, L = %.4f" % (N2, L), flush=True)
    pos = sdt.rsa_positions(N2, L, 1.0, 2)
    return pos, np.ones(N2), L


def test_config2_lanczos_group(sdt, orc, sys2):
    pos, rdi, L = sys2
    assert abs(L - 111.7693) < 1e-3
    z = sdt.gaussian(2, 0, 3 * N2 * B2).reshape(B2, 3 * N2).T.copy()
    ys, its = {}, {}
    for prec in (sdt.FP64, sdt.FP32):
        m = sdt.MobSpme(N2, rdi, L, XI, RC, K2, P2, prec)
        assert m.init(), getattr(m, "init_error", "")
        m.update(pos)
        bl = sdt.BrwnLanczos(3 * N2, 100, B2, TOL)
        assert bl.init()
        y, ok = bl.compute(m, z)
        assert ok
        ys[prec], its[prec] = y, list(bl.iters())
        print("config 2: precision %d Lanczos iterations %s" % (prec, its[prec]), flush=True)
        assert 0 < its[prec][0] < 100
        if prec == sdt.FP64:
            mz = m.mul_vector(z)
            q = np.einsum("ij,ij->j", z, mz)
            yy = np.einsum("ij,ij->j", y, y)
            e_q = np.abs(yy - q).max() / np.abs(q).max()
            print("config 2: |y|^2 against z^T M z %.3e" % e_q, flush=True)
            assert e_q < 1e-4
        del m, bl
    e32 = orc.frob_rel_err(ys[sdt.FP64], ys[sdt.FP32])
    print("config 2: FP32-mixed against FP64 %.3e" % e32, flush=True)
    assert e32 < 1e-5

    print("config 2: oracle block Lanczos (host)", flush=True)
    o = orc.MobSpme(N2, rdi, L, XI, RC, K2, P2)
    assert o.init()
    o.update(pos, rdi)
    oy, oit = orc.lanczos(o, z, 100, B2, TOL)
    e_o = orc.frob_rel_err(oy, ys[sdt.FP64])
    print("config 2: oracle iterations %s, FP64 against the oracle %.3e" % (list(oit), e_o),
          flush=True)
    assert abs(int(oit[0]) - its[sdt.FP64][0]) <= 3
    # Without reorthogonalisation the Krylov basis loses orthogonality long
    # before iteration 50, after which rounding differences (CholQR here,
    # Householder QR in the reference, other summation orders) move the
    # returned iterate by a fraction of the tolerance: the oracle's own
    # iterate moves by 3e-7 to 9e-7 at tol 1e-6 (N = 2000) when z is
    # perturbed by 1e-15, or when it is run again with its threaded sums
    # (test_oracle_physics.py::test_lanczos_iterate_rounding_sensitivity). The
    # same iterate is pinned to 1e-10 where no orthogonality is lost
    # (test_gpu_lanczos.py::test_lanczos_early_stop_same_iterate_as_oracle).
    assert e_o < 1e-5


def test_config2_bd_steps_fp32_vs_fp64(sdt, orc, sys2):
    pos, rdi, L = sys2
    disp, iters = {}, {}
    for prec in (sdt.FP64, sdt.FP32):
        bd = sdt.BDSim(pos, rdi, L, XI, RC, K2, P2, prec, 100, B2, TOL, 2.0, 125.0, 1e-3, 100, 7)
        bd.init()
        bd.advance(3)
        d = bd.positions() - pos
        d -= L * np.round(d / L)
        disp[prec], iters[prec] = d, bd.last_iters()
        print("config 2: BD precision %d iterations %s" % (prec, list(iters[prec])), flush=True)
        del bd
    for prec in (sdt.FP64, sdt.FP32):
        assert np.all(np.isfinite(disp[prec]))
        assert len(iters[prec]) == 13 and np.all((iters[prec] > 0) & (iters[prec] < 100))
    assert np.linalg.norm(disp[sdt.FP64]) > 0.0
    e = orc.frob_rel_err(disp[sdt.FP64], disp[sdt.FP32])
    print("config 2: BD displacements after 3 steps, FP32-mixed against FP64 %.3e" % e,
          flush=True)
    assert e < 1e-5
