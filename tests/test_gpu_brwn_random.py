"""GPU BrwnRandom (src/brwn/brwn_random.cc:118-263) against the oracle's
restatement on the dense GPU SPME operator, with the same Gaussian sample
vectors (Philox stream of the seed, one stream index per sampling)."""
import numpy as np
import pytest

from conftest import uniform_system
from oracle import brwn_random as BR

pytestmark = pytest.mark.gpu


def _mob(sdt, n=200, box=16.0, seed=3):
    pos, rdi = uniform_system(n, box, seed)
    g = sdt.MobSpme(n, rdi, box, 0.9, 5.0, 32, 6, sdt.FP64)
    assert g.init()
    g.update(pos)
    return g, g.mul_vector(np.eye(3 * n))


def _omega(orc, seed, stream, dim, nv):
    return orc.gaussian(seed, stream, dim * nv).reshape(nv, dim).T


def test_brwn_random_matches_oracle(sdt, orc):
    g, M = _mob(sdt)
    dim, nv, seed = M.shape[0], 30, 5678
    z = np.random.default_rng(4).standard_normal((dim, 5))
    br = sdt.BrwnRandom(dim, nv, 8, seed)
    assert br.init()
    y = br.compute(g, z)
    ref_sub = BR.sample(M, _omega(orc, seed, 0, dim, nv))
    ref = BR.apply(M, *ref_sub, z)
    assert np.linalg.norm(y - ref) / np.linalg.norm(ref) < 1e-9
    # |y|^2 = z^T M z column by column
    assert np.allclose((y * y).sum(0), (z * (M @ z)).sum(0), rtol=1e-10)
    # a null mobility reuses the sampled subspace
    z2 = np.random.default_rng(5).standard_normal((dim, 2))
    y2 = br.compute(None, z2)
    assert np.linalg.norm(y2 - BR.apply(M, *ref_sub, z2)) / np.linalg.norm(y2) < 1e-9
    # a new sampling takes the next stream index
    y3 = br.compute(g, z)
    ref3 = BR.apply(M, *BR.sample(M, _omega(orc, seed, 1, dim, nv)), z)
    assert np.linalg.norm(y3 - ref3) / np.linalg.norm(ref3) < 1e-9
    assert np.linalg.norm(y3 - y) > 1e-6 * np.linalg.norm(y)


def test_brwn_random_approaches_root_with_more_vectors(sdt, orc):
    g, M = _mob(sdt, n=60, box=12.0, seed=6)
    dim = M.shape[0]
    w, V = np.linalg.eigh(0.5 * (M + M.T))
    root = (V * np.sqrt(w)) @ V.T
    z = np.random.default_rng(7).standard_normal((dim, 3))
    errs = []
    for nv in (4, 40, 128):
        br = sdt.BrwnRandom(dim, nv, 4, 11)
        assert br.init()
        y = br.compute(g, z)
        ref = BR.apply(M, *BR.sample(M, _omega(orc, 11, 0, dim, nv)), z)
        assert np.linalg.norm(y - ref) / np.linalg.norm(ref) < 1e-9
        errs.append(np.linalg.norm(y - root @ z) / np.linalg.norm(root @ z))
    # the sampled approximation improves with the number of vectors
    assert errs[2] < errs[1] < errs[0]


def test_brwn_random_errors(sdt):
    g, M = _mob(sdt, n=20, box=12.0, seed=8)
    dim = M.shape[0]
    assert not sdt.BrwnRandom(dim, dim + 1, 4).init()
    br = sdt.BrwnRandom(dim, 10, 4)
    assert br.init()
    with pytest.raises(Exception):
        br.compute(None, np.ones((dim, 1)))   # no sampling yet
