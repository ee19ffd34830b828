"""Pins the BrwnRandom restatement (oracle/brwn_random.py, after
src/brwn/brwn_random.cc:118-263) by closed forms and its defining
invariant."""
import numpy as np

from oracle import brwn_random as BR


def test_scaled_identity_gives_exact_root():
    # M = mu I: first = sqrt(mu) P z, second = P z - z, alpha = sqrt(mu)
    rng = np.random.default_rng(1)
    dim, nv, mu = 40, 6, 2.7
    M = mu * np.eye(dim)
    z = rng.standard_normal((dim, 3))
    y = BR.apply(M, *BR.sample(M, rng.standard_normal((dim, nv))), z)
    assert np.allclose(y, np.sqrt(mu) * z, rtol=0, atol=1e-12)


def test_invariant_block_is_captured_exactly():
    # M = diag(M1, mu I) with the sample inside M1's coordinates: V spans
    # them, B^{1/2} = M1^{1/2}, alpha = sqrt(mu), so y = M^{1/2} z exactly
    rng = np.random.default_rng(2)
    dim, r, mu = 30, 6, 0.8
    A = rng.standard_normal((r, r))
    M1 = A @ A.T + 0.5 * np.eye(r)
    M = mu * np.eye(dim)
    M[:r, :r] = M1
    omega = np.zeros((dim, r))
    omega[:r] = rng.standard_normal((r, r))
    z = rng.standard_normal((dim, 4))
    y = BR.apply(M, *BR.sample(M, omega), z)
    w, V = np.linalg.eigh(M)
    root = (V * np.sqrt(w)) @ V.T
    assert np.allclose(y, root @ z, rtol=0, atol=1e-11)


def test_norm_matches_quadratic_form():
    rng = np.random.default_rng(3)
    dim, nv = 50, 10
    A = rng.standard_normal((dim, dim))
    M = A @ A.T / dim + 0.3 * np.eye(dim)
    z = rng.standard_normal((dim, 6))
    y = BR.apply(M, *BR.sample(M, rng.standard_normal((dim, nv))), z)
    assert np.allclose((y * y).sum(0), (z * (M @ z)).sum(0), rtol=1e-12)
    # and it is closer to M^{1/2} z than z's scaled copy is
    w, V = np.linalg.eigh(M)
    root = (V * np.sqrt(w)) @ V.T
    assert np.linalg.norm(y - root @ z) < np.linalg.norm(root @ z)
