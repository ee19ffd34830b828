"""Config 1 at its exact workload on the GPU: N = 1000 monodisperse spheres
(a = 1) at volume fraction 0.1 (L = 34.73, the box of the reference's
N1000_Phi0.1.xyz), seeded random-sequential-addition positions, SPME mesh
32^3, p = 6, rc = 5, xi = 0.896; 16 seeded N(0, 1) right-hand sides; block
Lanczos b = 16, tol 1e-6; the 3000 x 3000 dense Ewald matrix and its
Cholesky factor; steric k0 = 125, r0 = 2; 10 BD steps with dt = 1e-3.

Gates: FP64 GPU = oracle to 1e-10 (matvec) and 1e-6 (trajectory); FP32 =
FP64 to 1e-5; Lanczos = eigen-sqrt of the same operator to 1e-5 (tol 1e-6);
L L^T = M.

The SPME operator of these parameters is far from the dense Ewald sum: xi =
0.896 (forced by rc = 5 at 1e-6 truncation) puts the reciprocal Gaussian at
exp(-k^2 / 4 xi^2) = 0.07 at the Nyquist wavenumber of the 32^3 mesh, and the
modes beyond it are lost. The oracle shows the same: relative error 1.2 at
K = 32, 3.5e-2 at K = 64, 1.6e-4 at K = 128. That difference is reported, not
gated; the comparison with the dense Ewald sum is gated with a splitting
balanced for the 32^3 mesh (xi = 0.4, rc = 10: 4.6e-4 on the oracle).
"""
import numpy as np
import pytest

from conftest import box_for

pytestmark = pytest.mark.gpu

N, PHI, K, P, XI, RC, B, TOL = 1000, 0.1, 32, 6, 0.896, 5.0, 16, 1e-6
L1 = box_for(N, PHI)


def _eig_sqrt(M):
    w, V = np.linalg.eigh(0.5 * (M + M.T))
    return (V * np.sqrt(np.maximum(w, 0.0))) @ V.T


@pytest.fixture(scope="module")
def sys1(sdt):
    pos = sdt.rsa_positions(N, L1, 1.0, 101)
    rdi = np.ones(N)
    z = sdt.gaussian(2024, 0, 3 * N * 16).reshape(16, 3 * N).T.copy()
    return pos, rdi, z


def _spme(sdt, pos, rdi, prec):
    m = sdt.MobSpme(N, rdi, L1, XI, RC, K, P, prec)
    assert m.init(), getattr(m, "init_error", "")
    m.update(pos)
    return m


def test_config1_matvec(sdt, orc, sys1):
    pos, rdi, z = sys1
    assert abs(L1 - 34.729) < 0.01
    o = orc.MobSpme(N, rdi, L1, XI, RC, K, P)
    assert o.init()
    o.update(pos, rdi)
    ref = o.mul_vector(z)
    g64 = _spme(sdt, pos, rdi, sdt.FP64).mul_vector(z)
    g32 = _spme(sdt, pos, rdi, sdt.FP32).mul_vector(z)
    assert orc.frob_rel_err(ref, g64) < 1e-10
    assert orc.frob_rel_err(g64, g32) < 1e-5
    me = sdt.MobEwald(N, rdi, L1, 1e-12)
    assert me.init()
    me.update(pos)
    mz = me.mul_vector(z)
    e = orc.frob_rel_err(mz, g64)
    print("config 1: SPME (K=32, xi=0.896) vs dense Ewald, M z: %.3e (reported)" % e, flush=True)
    bal = sdt.MobSpme(N, rdi, L1, 0.4, 10.0, K, P, sdt.FP64)
    assert bal.init()
    bal.update(pos)
    eb = orc.frob_rel_err(mz, bal.mul_vector(z))
    print("config 1: SPME (K=32, xi=0.4, rc=10) vs dense Ewald: %.3e" % eb, flush=True)
    assert eb < 1e-3


def test_config1_lanczos_and_cholesky(sdt, orc, sys1):
    pos, rdi, z = sys1
    out = {}
    for prec in (sdt.FP64, sdt.FP32):
        m = _spme(sdt, pos, rdi, prec)
        bl = sdt.BrwnLanczos(3 * N, 100, B, TOL)
        assert bl.init()
        y, ok = bl.compute(m, z)
        assert ok
        out[prec] = (y, list(bl.iters()), m)
    y64, it64, m64 = out[sdt.FP64]
    print("config 1: Lanczos iterations FP64 %s FP32 %s" % (it64, out[sdt.FP32][1]), flush=True)
    Ms = m64.mul_vector(np.eye(3 * N))
    assert orc.frob_rel_err(_eig_sqrt(Ms) @ z, y64) < 1e-5
    assert orc.frob_rel_err(y64, out[sdt.FP32][0]) < 1e-5
    # the dense Ewald matrix: Cholesky displacements on the same vectors
    me = sdt.MobEwald(N, rdi, L1, 1e-12)
    assert me.init()
    me.update(pos)
    Me = me.get_matrix()
    ch = sdt.BrwnChol(3 * N)
    assert ch.init()
    # covariance of y = L z: for an orthogonal sample Q (Q Q^T = I),
    # (L Q)(L Q)^T = M exactly
    Q, _ = np.linalg.qr(np.random.default_rng(8).standard_normal((3 * N, 3 * N)))
    LQ = ch.compute(me, Q)
    assert np.abs(LQ @ LQ.T - Me).max() < 1e-11 * np.abs(Me).max()
    yc = ch.compute(me, z)
    # Cholesky and Lanczos square roots differ pointwise (L != M^{1/2}); the
    # Lanczos one reproduces z^T M z of its (SPME) operator
    qs = np.einsum("ij,ij->j", z, Ms @ z)
    assert np.allclose((y64 * y64).sum(0), qs, rtol=1e-5)
    e = orc.frob_rel_err(_eig_sqrt(Me) @ z, y64)
    print("config 1: Lanczos on SPME vs eigen-sqrt of dense Ewald: %.3e (reported); "
          "|L z| = %.6e" % (e, np.linalg.norm(yc)), flush=True)


def test_config1_bd_10_steps(sdt, orc, sys1):
    pos, rdi, _ = sys1
    args = (L1, XI, RC, K, P)
    res = {}
    for prec in (sdt.FP64, sdt.FP32):
        bd = sdt.BDSim(pos, rdi, *args, prec, 100, B, TOL, 2.0, 125.0, 1e-3, 100, 77)
        bd.init()
        bd.advance(10)
        res[prec] = bd.positions() - pos
    ob = orc.BDSim(pos, rdi, *args, 100, B, TOL, 2.0, 125.0, 1e-3, 100, 77)
    ob.advance(10)
    d_ref = ob.positions() - pos
    assert orc.frob_rel_err(d_ref, res[sdt.FP64]) < 1e-6
    assert orc.frob_rel_err(res[sdt.FP64], res[sdt.FP32]) < 1e-5
    pb = sdt.PlainBDSim(pos, rdi, *args, 100, B, TOL, 2.0, 125.0, 1e-3, 100, 77)
    pb.init()
    pb.advance(10)
    assert orc.frob_rel_err(d_ref, pb.positions() - pos) < 1e-6
