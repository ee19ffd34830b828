"""Pins the oracle's building blocks against libraries and closed forms.

FFT (the reference's MKL DFTI, src/mob/spme.cc:549-618) against numpy.fft and
a brute-force DFT; dense linear algebra (MKL LAPACKE,
src/brwn/brwn_lanczos.cc:56,107,167) against numpy/scipy; Philox against the
Random123 known-answer vectors; B-splines (src/mob/spme.cc:38-68) against
closed forms and the reference's own polynomial table tab_W
(src/mob/spme.h:89-104).
"""
import json
import os

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 6, 7, 8, 12, 30, 45, 64, 96, 128, 250])
def test_fft1d_matches_numpy(orc, n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    np.testing.assert_allclose(orc.fft1d(x, -1), np.fft.fft(x), rtol=0, atol=1e-12 * n)
    np.testing.assert_allclose(orc.fft1d(x, +1), np.fft.ifft(x) * n, rtol=0, atol=1e-12 * n)


def test_fft1d_brute_force_dft(orc):
    n = 24
    x = np.random.default_rng(1).standard_normal(n) + 0j
    k = np.arange(n)
    dft = np.exp(-2j * np.pi * np.outer(k, k) / n) @ x
    np.testing.assert_allclose(orc.fft1d(x, -1), dft, atol=1e-12)


@pytest.mark.parametrize("K", [8, 12, 24, 30, 32])
def test_rfft3d_matches_numpy(orc, K):
    r = np.random.default_rng(K).standard_normal((K, K, K))
    s = orc.rfftn(r)
    ref = np.fft.rfftn(r)
    assert np.abs(s - ref).max() <= 1e-12 * np.abs(ref).max()
    back = orc.irfftn(s, K)
    np.testing.assert_allclose(back, r, atol=1e-13)
    # Parseval on the half spectrum (interior bins count twice)
    w = np.full(K // 2 + 1, 2.0)
    w[0] = 1.0
    if K % 2 == 0:
        w[-1] = 1.0
    assert abs((np.abs(s) ** 2 * w).sum() / K ** 3 - (r ** 2).sum()) < 1e-10 * (r ** 2).sum()


def test_sym_eig_matches_numpy(orc):
    rng = np.random.default_rng(2)
    for n in [1, 2, 5, 40, 120]:
        A = rng.standard_normal((n, n))
        A = A + A.T
        w, V = orc.sym_eig(A)
        np.testing.assert_allclose(w, np.linalg.eigvalsh(A), atol=1e-11 * max(1, np.abs(w).max()))
        np.testing.assert_allclose(V @ np.diag(w) @ V.T, A, atol=1e-11 * np.abs(A).max())
        np.testing.assert_allclose(V.T @ V, np.eye(n), atol=1e-12)


def test_tridiag_eig_matches_scipy(orc):
    import scipy.linalg as sl
    rng = np.random.default_rng(3)
    d = rng.standard_normal(60)
    e = rng.standard_normal(59)
    w, V = orc.tridiag_eig(d, e)
    w2 = sl.eigvalsh_tridiagonal(d, e)
    np.testing.assert_allclose(w, w2, atol=1e-12)


def test_householder_qr(orc):
    rng = np.random.default_rng(4)
    A = rng.standard_normal((200, 9))
    Q, R = orc.householder_qr(A)
    np.testing.assert_allclose(Q @ R, A, atol=1e-12)
    np.testing.assert_allclose(Q.T @ Q, np.eye(9), atol=1e-13)
    assert np.abs(np.tril(R, -1)).max() == 0.0
    # same R as LAPACK up to row signs
    R2 = np.linalg.qr(A, mode="r")
    np.testing.assert_allclose(np.abs(R), np.abs(R2), atol=1e-12)


def test_cholesky(orc):
    import scipy.linalg as sl
    rng = np.random.default_rng(5)
    A = rng.standard_normal((30, 30))
    S = A @ A.T + 30 * np.eye(30)
    L, info = orc.cholesky(S)
    assert info == 0
    np.testing.assert_allclose(L, sl.cholesky(S, lower=True), atol=1e-12)
    _, info = orc.cholesky(-S)
    assert info == 1


def test_philox_known_answers(orc):
    # Random123 known-answer vectors for philox4x32-10
    with open(os.path.join(GOLDEN, "philox4x32_10_kat.json")) as fh:
        kats = json.load(fh)
    for kat in kats:
        ctr = [int(v, 16) for v in kat["ctr"]]
        key = [int(v, 16) for v in kat["key"]]
        out = orc.philox(ctr, key)
        assert [int(v) for v in out] == [int(v, 16) for v in kat["out"]]


def test_gaussian_moments(orc):
    g = orc.gaussian(2024, 1, 2_000_001)
    assert abs(g.mean()) < 5 / np.sqrt(g.size)
    assert abs(g.var() - 1.0) < 5 * np.sqrt(2.0 / g.size)
    assert abs(((g ** 4).mean()) - 3.0) < 0.02
    # distinct streams are distinct
    assert not np.allclose(orc.gaussian(2024, 2, 100), g[:100])


def _cardinal_bspline(p, u):
    """M_p(u) from the explicit sum (de Boor): sum_k (-1)^k C(p,k) (u-k)_+^(p-1) / (p-1)!"""
    from math import comb, factorial
    return sum((-1) ** k * comb(p, k) * max(u - k, 0.0) ** (p - 1) for k in range(p + 1)) / factorial(p - 1)


@pytest.mark.parametrize("p", [2, 3, 4, 5, 6, 7, 8])
def test_bspline_closed_form(orc, p):
    # BSpline(p, x) = M_p(x + 1)
    for x in np.linspace(-1.5, p + 0.5, 37):
        assert abs(orc.bspline(p, x) - _cardinal_bspline(p, x + 1)) < 1e-12
    # partition of unity of the p weights of any fractional offset
    for frac in [0.0, 0.1, 0.5, 0.93]:
        assert abs(sum(orc.bspline(p, -frac + k) for k in range(p)) - 1.0) < 1e-13


def test_bspline_reference_table(orc):
    # the reference's polynomial table tab_W (src/mob/spme.h:89-104): weight x
    # of fractional offset f is sum_y W[x][y] g^y with g = f - 1/2
    with open(os.path.join(GOLDEN, "bspline_tab_w.json")) as fh:
        tabs = json.load(fh)
    for p_str, rows in tabs.items():
        p = int(p_str)
        W = np.array(rows, dtype=float) / rows_denominator(p)
        for f in [0.0, 0.25, 0.5, 0.77, 0.999]:
            g = f - 0.5
            q = W @ np.array([g ** y for y in range(p)])
            ours = np.array([orc.bspline(p, -f + x) for x in range(p)])
            np.testing.assert_allclose(ours, q, atol=1e-14)


def rows_denominator(p):
    return {4: 48.0, 6: 3840.0}[p]
