"""GPU block Lanczos (BrwnLanczos, src/brwn/brwn_lanczos.cc:16-304), the
BD step (BDSimBox::Advance, src/simbox/simbox_bd.cc:296-366) and the
Philox Gaussian stream, against the oracle.

Lanczos is checked the way tests/lanczos/testlanczos.cc:81-144 does it,
against the eigen-sqrt of the same dense operator; Cholesky displacements
differ from M^{1/2} z by design (M = L L^T but L != M^{1/2}), so they are not
compared pointwise.
"""
import os

import numpy as np
import pytest

from conftest import uniform_system

pytestmark = pytest.mark.gpu
G = np.load(os.path.join(os.path.dirname(__file__), "golden", "golden_small.npz"))


def _mob(sdt, orc, n, box, K, p, prec, seed, xi=0.9, rmax=5.0):
    pos, rdi = uniform_system(n, box, seed)
    g = sdt.MobSpme(n, rdi, box, xi, rmax, K, p, prec)
    assert g.init()
    g.update(pos)
    o = orc.MobSpme(n, rdi, box, xi, rmax, K, p)
    assert o.init()
    o.update(pos, rdi)
    return g, o


def _eig_sqrt(M):
    w, V = np.linalg.eigh(0.5 * (M + M.T))
    return (V * np.sqrt(w)) @ V.T


def test_gram_kernel_matches_numpy(sdt):
    rng = np.random.default_rng(1)
    for n, b in [(1000, 8), (4097, 16), (333, 3), (50, 1), (70000, 4)]:
        A = rng.standard_normal((n, b))
        B = rng.standard_normal((n, b))
        ref = A.T @ B
        assert np.abs(sdt.debug_gram(A, B) - ref).max() < 1e-12 * np.abs(ref).max()


def test_grouped_gram_matches_numpy(sdt):
    rng = np.random.default_rng(4)
    for n, btot, bs in [(5000, 100, 12), (3001, 20, 16), (700, 7, 3), (4096, 5, 1)]:
        A = rng.standard_normal((n, btot))
        B = rng.standard_normal((n, btot))
        G = sdt.debug_gram_groups(A, B, bs)
        for g in range((btot + bs - 1) // bs):
            c0, c1 = g * bs, min(btot, (g + 1) * bs)
            ref = A[:, c0:c1].T @ B[:, c0:c1]
            w = c1 - c0
            assert np.abs(G[g, :w, :w] - ref).max() < 1e-12 * np.abs(ref).max()
            assert not G[g, w:, :].any() and not G[g, :, w:].any()


def test_gaussian_stream_matches_oracle(sdt, orc):
    g = sdt.gaussian(1234, 3, 100001)
    assert np.abs(g - orc.gaussian(1234, 3, 100001)).max() < 1e-14
    assert np.abs(sdt.gaussian(1234, 5, 16) - G["gauss_1234_5"]).max() < 1e-14


def test_lanczos_matches_eigen_sqrt(sdt, orc):
    # dense operator of the GPU SPME mobility; tol 1e-7 -> error < 1e-6
    n, box = 200, 16.0
    g, _ = _mob(sdt, orc, n, box, 32, 6, sdt.FP64, 3)
    M = g.mul_vector(np.eye(3 * n))
    M2 = _eig_sqrt(M)
    z = np.random.default_rng(2).uniform(0, 1, (3 * n, 3))
    for nrhs in (1, 3):
        bl = sdt.BrwnLanczos(3 * n, 100, 16, 1e-7)
        assert bl.init()
        y, ok = bl.compute(g, z[:, :nrhs])
        assert ok
        assert orc.frob_rel_err(M2 @ z[:, :nrhs], y) < 1e-6


@pytest.mark.parametrize("maxnrhs", [1, 4, 16])
def test_lanczos_matches_oracle(sdt, orc, maxnrhs):
    n, box = 300, 20.0
    g, o = _mob(sdt, orc, n, box, 32, 6, sdt.FP64, 4)
    z = orc.gaussian(7, 0, 3 * n * 20).reshape(20, 3 * n).T.copy()
    bl = sdt.BrwnLanczos(3 * n, 100, maxnrhs, 1e-10)
    assert bl.init()
    y, ok = bl.compute(g, z)
    oy, oit = orc.lanczos(o, z, 100, maxnrhs, 1e-10)
    assert ok
    assert orc.frob_rel_err(oy, y) < 1e-8
    # the device checks ||D_i - D_{i-1}||_F (equal to ||Y_i - Y_{i-1}||_F
    # while the Krylov basis is orthonormal); without reorthogonalisation
    # the two criteria can stop an iteration or two apart after ~60 steps
    assert np.abs(bl.iters() - oit).max() <= 3


@pytest.mark.parametrize("tol", [1e-2, 1e-3])
@pytest.mark.parametrize("maxnrhs", [4, 8])
def test_lanczos_early_stop_same_iterate_as_oracle(sdt, orc, tol, maxnrhs):
    # Stopped within 5-10 iterations the basis is still orthonormal and the
    # oracle's iterate is stable to rounding (1e-14 under a 1e-15
    # perturbation of z), while consecutive iterates differ by 8e-4 to 9e-3:
    # the same iteration count and 1e-10 agreement pin that the returned Y_i
    # is the iterate the convergence test accepted (the first i with
    # ||Y_i - Y_{i-1}||_F / ||Z||_F < tol, src/brwn/brwn_lanczos.cc:198-206).
    n, box = 300, 20.0
    g, o = _mob(sdt, orc, n, box, 32, 6, sdt.FP64, 4)
    z = orc.gaussian(7, 0, 3 * n * 20).reshape(20, 3 * n).T.copy()
    bl = sdt.BrwnLanczos(3 * n, 100, maxnrhs, tol)
    assert bl.init()
    y, ok = bl.compute(g, z)
    oy, oit = orc.lanczos(o, z, 100, maxnrhs, tol)
    assert ok
    assert list(bl.iters()) == list(oit)
    assert orc.frob_rel_err(oy, y) < 1e-10


def test_lanczos_many_groups_and_batches(sdt, orc):
    # 130 right-hand sides in groups of 8: two lockstep batches (128 + 2
    # columns), a narrower last group in each
    n, box = 300, 20.0
    g, o = _mob(sdt, orc, n, box, 32, 6, sdt.FP64, 8)
    z = orc.gaussian(12, 0, 3 * n * 130).reshape(130, 3 * n).T.copy()
    bl = sdt.BrwnLanczos(3 * n, 100, 8, 1e-8)
    assert bl.init()
    y, ok = bl.compute(g, z)
    oy, oit = orc.lanczos(o, z, 100, 8, 1e-8)
    assert ok
    assert orc.frob_rel_err(oy, y) < 1e-6
    assert len(bl.iters()) == len(oit) == 17
    assert np.abs(np.asarray(bl.iters()) - oit).max() <= 3


def test_lanczos_repeat_call_is_history_independent(sdt, orc):
    # a handle keeps no convergence state between calls: a second call on
    # the same input stops at the same iteration with the same
    # (deterministic) result, and a fresh handle agrees bit for bit
    n, box = 300, 20.0
    g, _ = _mob(sdt, orc, n, box, 32, 6, sdt.FP64, 9)
    M2 = _eig_sqrt(g.mul_vector(np.eye(3 * n)))
    z = orc.gaussian(13, 0, 3 * n * 12).reshape(12, 3 * n).T.copy()
    bl = sdt.BrwnLanczos(3 * n, 100, 4, 1e-7)
    assert bl.init()
    y1, ok1 = bl.compute(g, z)
    it1 = list(bl.iters())
    y2, ok2 = bl.compute(g, z)
    it2 = list(bl.iters())
    assert ok1 and ok2
    assert it1 == it2
    assert np.array_equal(y1, y2)
    fresh = sdt.BrwnLanczos(3 * n, 100, 4, 1e-7)
    assert fresh.init()
    y3, ok3 = fresh.compute(g, z)
    assert ok3 and list(fresh.iters()) == it1 and np.array_equal(y1, y3)
    assert orc.frob_rel_err(M2 @ z, y1) < 1e-6


def test_lanczos_golden(sdt):
    box, xi, rmax, K, p = G["mono_params"]
    pos, rdi = G["mono_pos"], G["mono_rdi"]
    g = sdt.MobSpme(rdi.size, rdi, box, xi, rmax, int(K), int(p), sdt.FP64)
    assert g.init()
    g.update(pos)
    bl = sdt.BrwnLanczos(3 * rdi.size, 100, 4, 1e-10)
    assert bl.init()
    y, ok = bl.compute(g, G["lz_z"])
    assert ok
    assert np.linalg.norm(y - G["lz_y"]) / np.linalg.norm(G["lz_y"]) < 1e-8


def test_lanczos_fp32_within_1e5(sdt, orc):
    n, box = 2000, 30.0
    z = orc.gaussian(8, 1, 3 * n * 8).reshape(8, 3 * n).T.copy()
    ys = []
    for prec in (sdt.FP64, sdt.FP32):
        g, _ = _mob(sdt, orc, n, box, 64, 6, prec, 5)
        bl = sdt.BrwnLanczos(3 * n, 100, 8, 1e-7)
        assert bl.init()
        y, ok = bl.compute(g, z)
        assert ok
        ys.append(y)
    assert orc.frob_rel_err(ys[0], ys[1]) < 1e-5


def test_lanczos_reorth_agrees(sdt, orc):
    n, box = 300, 20.0
    g, _ = _mob(sdt, orc, n, box, 32, 6, sdt.FP64, 6)
    z = orc.gaussian(9, 0, 3 * n * 6).reshape(6, 3 * n).T.copy()
    out = []
    for reorth in (False, True):
        bl = sdt.BrwnLanczos(3 * n, 100, 6, 1e-9)
        assert bl.init()
        bl.set_reorth(reorth)
        y, ok = bl.compute(g, z)
        assert ok
        out.append(y)
    assert orc.frob_rel_err(out[0], out[1]) < 1e-7


def test_lanczos_exhausted_krylov_space(sdt, orc):
    # 4 particles, 12 unknowns, blocks of 4: the Krylov space is the whole
    # space after three iterations, later blocks are rounding noise (or
    # exactly rank deficient); the iterate stays the eigen-sqrt product
    n, box = 4, 16.0
    g, o = _mob(sdt, orc, n, box, 32, 6, sdt.FP64, 11)
    M2 = _eig_sqrt(g.mul_vector(np.eye(3 * n)))
    z = np.random.default_rng(5).standard_normal((3 * n, 4))
    for maxnrhs in (4, 2):
        bl = sdt.BrwnLanczos(3 * n, 40, maxnrhs, 1e-9)
        assert bl.init()
        y, ok = bl.compute(g, z)
        _, oit = orc.lanczos(o, z, 40, maxnrhs, 1e-9)
        assert ok
        assert np.all(np.isfinite(y))
        assert orc.frob_rel_err(M2 @ z, y) < 1e-8
        assert np.abs(np.asarray(bl.iters()) - oit).max() <= 3


def test_lanczos_non_convergence_reported(sdt, orc):
    n, box = 200, 16.0
    g, _ = _mob(sdt, orc, n, box, 32, 6, sdt.FP64, 7)
    z = np.random.default_rng(3).standard_normal((3 * n, 4))
    bl = sdt.BrwnLanczos(3 * n, 2, 4, 1e-14)
    assert bl.init()
    y, ok = bl.compute(g, z)
    assert not ok
    assert list(bl.iters()) == [-1]
    assert np.linalg.norm(y) > 0   # the result is still written


def test_bd_matches_oracle(sdt, orc):
    box, xi, rmax, K, p, mi_, mn, tol, r0, k0, dt, mi, seed = G["bd_params"]
    pos0 = G["bd_pos0"]
    rdi = np.ones(pos0.size // 3)
    bd = sdt.BDSim(pos0, rdi, box, xi, rmax, int(K), int(p), sdt.FP64, int(mi_), int(mn), tol,
                   r0, k0, dt, int(mi), int(seed))
    bd.init()
    bd.advance(6)
    ref = G["bd_pos6"]
    assert np.abs(bd.positions() - ref).max() < 1e-8 * max(1.0, np.abs(ref - pos0).max())


def test_bd_overlapping_start_fp32_vs_fp64(sdt, orc):
    # overlapping uniform start: steric forces act every step
    n, box = 500, 22.0
    pos, rdi = uniform_system(n, box, 31)
    res = []
    for prec in (sdt.FP64, sdt.FP32):
        bd = sdt.BDSim(pos, rdi, box, 0.9, 5.0, 32, 6, prec, 100, 8, 1e-7, 2.0, 125.0, 1e-3, 10,
                       77)
        bd.init()
        bd.advance(12)
        res.append(bd.positions() - pos)
    ob = orc.BDSim(pos, rdi, box, 0.9, 5.0, 32, 6, 100, 8, 1e-7, 2.0, 125.0, 1e-3, 10, 77)
    ob.advance(12)
    disp_ref = ob.positions() - pos
    assert orc.frob_rel_err(disp_ref, res[0]) < 1e-6
    assert orc.frob_rel_err(res[0], res[1]) < 1e-5
