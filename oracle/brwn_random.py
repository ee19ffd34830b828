"""Restatement of BrwnRandom (src/brwn/brwn_random.cc:118-263, randomised
sampling of M^{1/2} z) in numpy, on a dense matrix.

TEST INFRASTRUCTURE ONLY: the checker of the GPU BrwnRandom in tests/. The
product package (stokesdt_amd) never imports this module.

Pinned in tests/test_oracle_brwn_random.py by closed forms: M = mu I gives
y = sqrt(mu) z; a positive semi-definite M of rank <= num_vecs gives
y = M^{1/2} z; and |y|^2 = z^T M z holds for every column.
"""
import numpy as np


def sample(M, omega):
    """:123-147. omega (dim, nv). Returns (V, U, s): V orthonormal basis of
    M^3 omega, U = V uu and s = sqrt(ss) from svd(V^T M V) = uu ss vv^T."""
    v = M @ (M @ (M @ omega))
    q, _ = np.linalg.qr(v)                 # dgeqrf + dorgqr
    b = q.T @ (M @ q)
    uu, ss, _ = np.linalg.svd(b)           # dgesvd, left vectors
    return q, q @ uu, np.sqrt(ss)


def apply(M, V, U, s, z):
    """:149-263 for z (dim, nrhs)."""
    first = U @ (s[:, None] * (U.T @ z))
    second = V @ (V.T @ z) - z
    y = M @ z
    out = np.empty_like(z, dtype=np.float64)
    for k in range(z.shape[1]):
        beta2 = z[:, k] @ y[:, k]
        a = second[:, k] @ second[:, k]
        b = -2.0 * (first[:, k] @ second[:, k])
        c = first[:, k] @ first[:, k] - beta2
        alpha = (-b + np.sqrt(b * b - 4.0 * a * c)) / (2.0 * a)
        out[:, k] = first[:, k] - alpha * second[:, k]
    return out
