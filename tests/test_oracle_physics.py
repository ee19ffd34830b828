"""Pins the oracle's mobility, pair lists, steric force, Lanczos and BD step
with the relations the reference's own tests assert and with properties the
mathematics fixes.

Reference tests restated here: tests/spme/testspme.cc (SPME vs dense Ewald
< 1e-6 at the parameters of tests/spme/test.sh:2-9), tests/steric/
teststeric.cc (cell list vs O(N^2) DirectSteric < 1e-6, :34-79,155),
tests/lanczos/testlanczos.cc (single and block Lanczos vs dense eigen-sqrt
< 1e-6 at tol 1e-7, :81-144). Where a relation is stated on the
reference's own inputs (tests/data/N1000_Phi0.1.xyz, N412.xyz with mono.mod
/ poly.mod), the committed copies under tests/golden/reference_data are
read (conftest.reference_system); other systems are generated from seeds.

The overlap correction (r < a_i + a_j) and the unequal-radius a^2 terms of
the real-space RPY scalars are pinned against the free-space
Rotne-Prager-Yamakawa tensor (Rotne & Prager 1969; Yamakawa 1970;
overlapping equal spheres as in Wajnryb et al. 2013), which the periodic
sums approach as 1/L in growing boxes.
"""
import numpy as np
import pytest

from conftest import direct_steric, poly_system, reference_system, uniform_system

L1000 = 34.729314   # box of the reference's N1000 phi=0.1 system


def _spme_vs_ewald(orc, pos, rdi, box, xi, rmax, K, p, nrhs=3, seed=7):
    n = rdi.size
    M = orc.ewald_matrix(pos, rdi, box, tol=1e-12)
    m = orc.MobSpme(n, rdi, box, xi, rmax, K, p)
    assert m.init()
    m.update(pos, rdi)
    f = np.random.default_rng(seed).uniform(0, 1, (3 * n, nrhs))
    v = m.mul_vector(f, alpha=2.1, beta=1.1, v=f)
    ve = 2.1 * (M @ f) + 1.1 * f
    return orc.frob_rel_err(ve, v)


@pytest.mark.slow
def test_spme_matches_ewald_reference_params_mono(orc):
    # tests/spme/test.sh:2-3 on the reference's own input: N1000_Phi0.1.xyz
    # with mono.mod, dim 256, porder 6, xi 0.8, rmax 8.371887
    pos, rdi, box = reference_system("N1000_Phi0.1.xyz", "mono.mod")
    err = _spme_vs_ewald(orc, pos, rdi, box, 0.8, 8.371887, 256, 6, nrhs=1)
    assert err < 1e-6, err


@pytest.mark.slow
def test_spme_matches_ewald_reference_params_poly(orc):
    # tests/spme/test.sh:8-9 on the reference's own input: N412.xyz with
    # poly.mod, dim 256, porder 8, xi 0.04, rmax 200.37188
    pos, rdi, box = reference_system("N412.xyz", "poly.mod")
    err = _spme_vs_ewald(orc, pos, rdi, box, 0.04, 200.37188, 256, 8, nrhs=2)
    assert err < 1e-6, err


def test_spme_matches_ewald_small(orc):
    pos, rdi = uniform_system(200, 20.0, 12)
    assert _spme_vs_ewald(orc, pos, rdi, 20.0, 0.8, 8.3, 128, 8) < 1e-7


def test_spme_matches_ewald_poly(orc):
    # polydisperse path (three mesh round trips, src/mob/spme.cc:771-838)
    rng = np.random.default_rng(13)
    pos = rng.uniform(0, 20.0, 600)
    rdi = rng.uniform(0.5, 1.5, 200)
    assert _spme_vs_ewald(orc, pos, rdi, 20.0, 0.8, 8.3, 128, 8) < 1e-7


def test_odd_order_nyquist_fix_converges_to_ewald(orc):
    # odd porder: the reference's B-spline modulus divides by ~1e-33 at the
    # Nyquist index (src/mob/spme.cc:110-128); with the neighbour mean there
    # SPME converges to the Ewald sum like the even orders around it
    pos, rdi = uniform_system(200, 20.0, 12)
    errs = {p: _spme_vs_ewald(orc, pos, rdi, 20.0, 0.8, 8.3, 128, p) for p in (6, 7, 8)}
    assert errs[8] < errs[7] < errs[6] < 1e-5, errs


def test_spme_converges_with_mesh(orc):
    pos, rdi = uniform_system(150, 18.0, 14)
    errs = [_spme_vs_ewald(orc, pos, rdi, 18.0, 1.0, 6.0, K, 6, nrhs=1) for K in (32, 64, 128)]
    assert errs[0] > errs[1] > errs[2]


def test_ewald_xi_independent_and_symmetric(orc):
    pos, rdi = uniform_system(60, 12.0, 15)
    M1 = orc.ewald_matrix(pos, rdi, 12.0, tol=1e-13)
    M2 = orc.ewald_matrix(pos, rdi, 12.0, xi=2.0 * orc.ewald_auto_xi(12.0), tol=1e-13)
    assert np.linalg.norm(M1 - M2) / np.linalg.norm(M1) < 1e-10
    assert np.abs(M1 - M1.T).max() == 0.0
    assert np.linalg.eigvalsh(M1).min() > 0.0


def test_ewald_free_space_limit(orc):
    # a pair at r = 10 in growing boxes: the off-diagonal block tends to the
    # free-space RPY tensor (6 pi eta a scaling, unit radii)
    #   xa = 3/(2r) - 1/r^3 along e,  ya = 3/(4r) + 1/(2r^3) across,
    # with the leading periodic correction decaying as 1/L
    r = 10.0
    xa = 1.5 / r - 1.0 / r ** 3
    ya = 0.75 / r + 0.5 / r ** 3
    dev = []
    for box in (200.0, 400.0, 800.0):
        pos = np.array([0.0, 0.0, 0.0, r, 0.0, 0.0])
        M = orc.ewald_matrix(pos, np.ones(2), box, tol=1e-13)
        blk = M[0:3, 3:6]
        dev.append((blk[0, 0] - xa, blk[1, 1] - ya))
        assert abs(blk[0, 1]) < 1e-12 and abs(blk[1, 2]) < 1e-12
    for c in range(2):
        assert abs(dev[0][c] / dev[1][c] - 2.0) < 0.1
        assert abs(dev[1][c] / dev[2][c] - 2.0) < 0.05
    assert abs(dev[2][0]) < 0.05 * xa and abs(dev[2][1]) < 0.05 * ya


def _pair_block(orc, r, ai, aj, box, spme=None):
    """Off-diagonal 3x3 block of the periodic mobility of two spheres at
    separation r along x in a cube of side `box`: from the dense Ewald sum,
    or from the SPME operator when spme = (xi, rmax, K, p)."""
    pos = np.array([0.3 * box, 0.5 * box, 0.5 * box, 0.3 * box + r, 0.5 * box, 0.5 * box])
    rdi = np.array([ai, aj])
    if spme is None:
        M = orc.ewald_matrix(pos, rdi, box, tol=1e-14)
        return M[0:3, 3:6]
    xi, rmax, K, p = spme
    m = orc.MobSpme(2, rdi, box, xi, rmax, K, p)
    assert m.init()
    m.update(pos, rdi)
    f = np.zeros((6, 3))
    f[3:6, :] = np.eye(3)
    return m.mul_vector(f)[0:3, :]


def _free_space_rpy(r, ai, aj):
    """(xa, ya) of the free-space RPY tensor, in units where the self
    mobility of a sphere of radius a is 1/a (6 pi eta = 1). Far apart:
    xa = 3/(2r) - (ai^2 + aj^2)/(2 r^3), ya = 3/(4r) + (ai^2 + aj^2)/(4 r^3).
    Overlapping equal spheres (r < 2a): (1/a)(1 - 9r/(32a)) I +
    (1/a)(3r/(32a)) e e, i.e. xa = (1/a)(1 - 3r/(16a)),
    ya = (1/a)(1 - 9r/(32a))."""
    if r >= ai + aj:
        s = ai * ai + aj * aj
        return 1.5 / r - s / (2 * r ** 3), 0.75 / r + s / (4 * r ** 3)
    assert ai == aj
    a = ai
    return (1.0 - 3.0 * r / (16.0 * a)) / a, (1.0 - 9.0 * r / (32.0 * a)) / a


def _check_one_over_l(boxes, devs, xa, ya, tol):
    # the leading periodic correction is c1/L (Hasimoto's -2.837/L on the
    # diagonal): doubling L about halves the deviation. Fitting
    # d(L) = d0 + c1/L + c3/L^3 through the three boxes removes the periodic
    # terms up to O(1/L^5); d0 must vanish.
    for c in range(2):
        assert abs(devs[0][c] / devs[1][c] - 2.0) < 0.25, devs
        assert abs(devs[1][c] / devs[2][c] - 2.0) < 0.1, devs
    A = np.array([[1.0, 1.0 / L, 1.0 / L ** 3] for L in boxes])
    d0 = [np.linalg.solve(A, np.array([d[c] for d in devs]))[0] for c in range(2)]
    assert abs(d0[0]) < tol * abs(xa) and abs(d0[1]) < tol * abs(ya), (d0, xa, ya)


@pytest.mark.parametrize("r,ai,aj", [(1.5, 1.0, 1.0), (0.5, 1.0, 1.0), (3.0, 2.0, 2.0),
                                     (12.0, 1.0, 2.0), (9.0, 0.5, 3.0)])
def test_ewald_pair_tends_to_free_space_rpy(orc, r, ai, aj):
    # overlap branch (src/mob/ewald.cc:177-195) for r < ai + aj and the
    # unequal-radius a^2 terms of scalars_ewald_F (src/mob/ewald.h:55-85)
    xa, ya = _free_space_rpy(r, ai, aj)
    devs = []
    boxes = (200.0, 400.0, 800.0)
    for box in boxes:
        blk = _pair_block(orc, r, ai, aj, box)
        assert abs(blk[0, 1]) < 1e-12 and abs(blk[1, 2]) < 1e-12 and abs(blk[0, 2]) < 1e-12
        assert abs(blk[1, 1] - blk[2, 2]) < 1e-12
        devs.append((blk[0, 0] - xa, blk[1, 1] - ya))
    _check_one_over_l(boxes, devs, xa, ya, 1e-6)


@pytest.mark.parametrize("r,ai,aj", [(1.5, 1.0, 1.0), (3.0, 2.0, 2.0), (6.0, 1.0, 2.0)])
def test_spme_pair_tends_to_free_space_rpy(orc, r, ai, aj):
    # the SPME operator's own real-space copy of the overlap branch and the
    # a^2 terms (src/mob/mob_spme.cc:151-197), in growing boxes; xi and the
    # mesh are chosen so that both Ewald parts are converged (< 1e-9)
    xa, ya = _free_space_rpy(r, ai, aj)
    devs = []
    boxes = (40.0, 80.0, 160.0)
    for box, K in zip(boxes, (64, 128, 256)):
        blk = _pair_block(orc, r, ai, aj, box, spme=(0.35, 14.0, K, 8))
        devs.append((blk[0, 0] - xa, blk[1, 1] - ya))
    _check_one_over_l(boxes, devs, xa, ya, 1e-4)


def test_spme_operator_symmetric(orc):
    pos, rdi = uniform_system(120, 16.0, 16)
    m = orc.MobSpme(120, rdi, 16.0, 0.9, 5.0, 32, 6)
    assert m.init()
    m.update(pos, rdi)
    rng = np.random.default_rng(0)
    x, y = rng.standard_normal(360), rng.standard_normal(360)
    Mx, My = m.mul_vector(x)[:, 0], m.mul_vector(y)[:, 0]
    assert abs(x @ My - y @ Mx) < 1e-12 * abs(x @ My)


def test_spread_interp_adjoint_and_brute_force(orc):
    rng = np.random.default_rng(17)
    n, K, p, box = 20, 16, 4, 10.0
    pos = rng.uniform(0, box, 3 * n)
    e = orc.SpmeEngine(n, None, box, 0.9, K, p)
    e.update(pos)
    f = rng.standard_normal(3 * n)
    G = e.spread(f)
    # brute force: mesh point m gets sum_i M_p(wrapped distance) f_i
    brute = np.zeros((3, K, K, K))
    for i in range(n):
        c = pos[3 * i:3 * i + 3] / box * K
        g = np.floor(c)
        base = (g - p // 2 + 1).astype(int) % K
        w = [[orc.bspline(p, g[d] - c[d] + x) for x in range(p)] for d in range(3)]
        for z in range(p):
            for y in range(p):
                for x in range(p):
                    brute[:, (base[2] + z) % K, (base[1] + y) % K, (base[0] + x) % K] += \
                        w[0][x] * w[1][y] * w[2][z] * f[3 * i:3 * i + 3]
    np.testing.assert_allclose(G, brute, atol=1e-13)
    H = rng.standard_normal((3, K, K, K))
    assert abs(e.interp(H) @ f - (H * G).sum()) < 1e-10 * abs((H * G).sum())


def test_lm2_closed_form(orc):
    # lm2 = ScalarRecip(|k|) |b_x b_y b_z|^2 (src/mob/spme.cc:71-160,
    # Essmann et al. 1995 moduli); checked at a few wavenumbers
    K, p, box, xi = 24, 6, 12.0, 0.7
    e = orc.SpmeEngine(4, None, box, xi, K, p)
    lm2 = e.lm2()
    sv = [orc.bspline(p, i) for i in range(p - 1)]

    def bmod2(m):
        s = sum(sv[i] * np.exp(2j * np.pi * m * i / K) for i in range(p - 1))
        return 1.0 / abs(s) ** 2

    for (mz, my, mx) in [(1, 0, 0), (3, -2, 5), (-7, 12, 12), (0, 1, 1), (-11, -5, 9)]:
        k = 2 * np.pi / box * np.array([mx, my, mz])
        kk = k @ k
        kx2 = kk / xi ** 2
        sr = 6 * np.pi * np.exp(-kx2 / 4) / kk * (1 + kx2 * (0.25 + kx2 / 8))
        ref = sr * bmod2(mx) * bmod2(my) * bmod2(mz)
        assert abs(lm2[mz % K, my % K, mx] - ref) < 1e-12 * ref


def _brute_pairs(pos, rdi, box, cutoff, mode):
    n = rdi.size
    P = pos.reshape(n, 3)
    x0 = np.mod(P, box)
    rows = []
    for i in range(n):
        d = np.abs(x0 - x0[i])
        d = np.where(d > box / 2.0, d - box, d)
        rr = (d ** 2).sum(1)
        if mode == "R":
            sel = np.sqrt(rr) <= cutoff
        else:
            rr = np.where(rr == 0.0, 1e-10, rr)
            sel = 2.0 * np.sqrt(rr) / (rdi[i] + rdi) < cutoff
        rows.append(set(np.nonzero(sel)[0].tolist()))
    return rows


@pytest.mark.parametrize("mode,cutoff", [("R", 5.0), ("R", 2.7), ("S", 2.0), ("S", 2.3)])
def test_pairlist_matches_brute_force(orc, mode, cutoff):
    pos, rdi = uniform_system(400, 24.0, 18)
    if mode == "S":
        rdi = np.random.default_rng(1).uniform(0.6, 1.4, 400)
    ia, ja = orc.pairlist(pos, rdi, 24.0, cutoff, mode)
    rows = _brute_pairs(pos, rdi, 24.0, cutoff, mode)
    for i in range(400):
        got = ja[ia[i]:ia[i + 1]]
        assert list(got) == sorted(got)
        assert set(got.tolist()) == rows[i], i
        assert i in set(got.tolist())   # self pair included


def test_steric_matches_direct(orc):
    # tests/steric/test.sh:2 on the reference's own input: N412.xyz with
    # poly.mod, cutoff 2.3, k0 125
    pos, rdi, box = reference_system("N412.xyz", "poly.mod")
    f = orc.steric(pos, rdi, box, 2.3, 125.0)
    fd = direct_steric(pos, rdi, box, 2.3, 125.0)
    assert np.linalg.norm(f) > 0
    assert np.linalg.norm(f - fd) / np.linalg.norm(fd) < 1e-6
    # Newton's third law
    assert np.abs(f.reshape(-1, 3).sum(0)).max() < 1e-9 * np.abs(f).max()


@pytest.fixture(scope="module")
def ewald1000(orc):
    # testlanczos.cc builds MobEwald on N1000_Phi0.1.xyz with mono.mod
    pos, rdi, box = reference_system("N1000_Phi0.1.xyz", "mono.mod")
    return orc.ewald_matrix(pos, rdi, box, tol=1e-12)


def _eig_sqrt(M):
    w, V = np.linalg.eigh(M)
    return (V * np.sqrt(w)) @ V.T


@pytest.mark.slow
def test_lanczos_matches_eigen_sqrt(orc, ewald1000):
    # tests/lanczos/testlanczos.cc:81-144: maxiters 100, maxnrhs 16, tol 1e-7,
    # z ~ Uniform(0, 1); single (1 rhs) and block (3 rhs), both < 1e-6
    M = ewald1000
    M2 = _eig_sqrt(M)
    z = np.random.default_rng(21).uniform(0, 1, (3000, 3))
    y1, it1 = orc.lanczos(M, z[:, :1], 100, 16, 1e-7)
    assert orc.frob_rel_err(M2 @ z[:, :1], y1) < 1e-6
    y3, it3 = orc.lanczos(M, z, 100, 16, 1e-7)
    assert orc.frob_rel_err(M2 @ z, y3) < 1e-6
    assert it1[0] > 0 and it3[0] > 0


def test_lanczos_iterate_rounding_sensitivity(orc):
    # The reference's block Lanczos (src/brwn/brwn_lanczos.cc:36-221) has no
    # reorthogonalisation. Stopped early (tol 1e-2, 8 iterations at N = 2000)
    # its iterate is stable to rounding; once the basis has lost
    # orthogonality (tol 1e-6, 48 iterations) a 1e-15 relative perturbation
    # of z moves the returned iterate by a fraction of the tolerance
    # (3e-7 to 9e-7 here; the threaded sums of the host products alone, run
    # twice on the same z, do the same and can stop an iteration apart).
    # Two correct implementations with different rounding (QR method,
    # summation order) agree to that level only.
    n, box = 2000, 30.0
    pos, rdi = uniform_system(n, box, 5)
    o = orc.MobSpme(n, rdi, box, 0.9, 5.0, 64, 6)
    assert o.init()
    o.update(pos, rdi)
    z = orc.gaussian(8, 1, 3 * n * 8).reshape(8, 3 * n).T.copy()
    z2 = z * (1.0 + 1e-15 * np.random.default_rng(0).standard_normal(z.shape))
    moved = {}
    for tol in (1e-2, 1e-6):
        y, it = orc.lanczos(o, z, 100, 8, tol)
        y2, it2 = orc.lanczos(o, z2, 100, 8, tol)
        assert abs(int(it[0]) - int(it2[0])) <= (0 if tol == 1e-2 else 1)
        moved[tol] = orc.frob_rel_err(y, y2)
    assert moved[1e-2] < 1e-12
    assert 1e-8 < moved[1e-6] < 1e-5


def test_lanczos_non_convergence_reported(orc):
    rng = np.random.default_rng(22)
    A = rng.standard_normal((90, 90))
    M = A @ A.T + 0.01 * np.eye(90)   # ill-conditioned
    y, it = orc.lanczos(M, rng.standard_normal((90, 2)), 3, 2, 1e-12)
    assert it[0] == -1


def test_cholesky_displacements(orc):
    # BrwnChol (src/brwn/brwn_chol.cc:52-83): y = L z with M = L L^T; not
    # pointwise equal to M^{1/2} z, but L L^T reproduces M
    pos, rdi = uniform_system(40, 10.0, 23)
    M = orc.ewald_matrix(pos, rdi, 10.0, tol=1e-12)
    eye = np.eye(120)
    Lz, info = orc.chol_displacements(M, eye)
    assert info == 0
    np.testing.assert_allclose(Lz @ Lz.T, M, atol=1e-12 * np.abs(M).max())


def test_bd_step_composition(orc):
    # one BD step without overlaps (f = 0): pos' = pos + sqrt(2 dt) M^{1/2} z
    # with z the first column of the Philox block of stream 0
    rng = np.random.default_rng(24)
    n, box = 30, 14.0
    pos = np.zeros(3 * n)
    k = 0
    while k < n:   # random sequential addition, no overlap
        c = rng.uniform(0, box, 3)
        ok = all(np.linalg.norm(np.remainder(c - pos[3 * j:3 * j + 3] + box / 2, box) - box / 2)
                 >= 2.05 for j in range(k))
        if ok:
            pos[3 * k:3 * k + 3] = c
            k += 1
    rdi = np.ones(n)
    K, p, xi, rmax, dt, mi = 32, 6, 0.9, 5.0, 1e-3, 4
    bd = orc.BDSim(pos, rdi, box, xi, rmax, K, p, 200, 4, 1e-10, 2.0, 125.0, dt, mi, 99)
    bd.advance(1)
    m = orc.MobSpme(n, rdi, box, xi, rmax, K, p)
    assert m.init()
    m.update(pos, rdi)
    Mdense = m.mul_vector(np.eye(3 * n))
    Mdense = 0.5 * (Mdense + Mdense.T)
    z = orc.gaussian(99, 0, 3 * n * mi)[:3 * n]
    expect = pos + np.sqrt(2 * dt) * (_eig_sqrt(Mdense) @ z)
    np.testing.assert_allclose(bd.positions(), expect, atol=1e-8)
