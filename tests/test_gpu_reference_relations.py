"""The reference's own test relations, run on the GPU with the reference's own
inputs (tests/data/N1000_Phi0.1.xyz and N412.xyz with mono.mod / poly.mod,
committed under tests/golden/reference_data):

- tests/spme/test.sh:2-9 with testspme.cc:103-141: v = 2.1 M f + 1.1 f for
  three Uniform(0, 1) vectors, SPME against the dense Ewald mobility
  (MobEwald, tol 1e-12), relative Frobenius error < 1e-6;
- tests/steric/test.sh:2 with teststeric.cc:34-79,155: cell-list steric
  forces against the O(N^2) sum, < 1e-6;
- tests/lanczos/test.sh with testlanczos.cc:81-144: block Lanczos (tol 1e-7,
  max_nrhs 16) on the dense Ewald mobility against its eigen-sqrt, 1 and 3
  right-hand sides, < 1e-6.

The reference draws its Uniform(0, 1) vectors from MKL (seed 1234), which
cannot be reproduced; a seeded numpy stream of the same distribution is used.
The dense Ewald mobility here is the GPU one, itself checked against the
oracle (test_gpu_ewald.py).
"""
import numpy as np
import pytest

from conftest import direct_steric, reference_system

pytestmark = pytest.mark.gpu


def _eig_sqrt(M):
    w, V = np.linalg.eigh(0.5 * (M + M.T))
    return (V * np.sqrt(np.maximum(w, 0.0))) @ V.T


@pytest.fixture(scope="module")
def n1000_ewald(sdt):
    pos, rdi, box = reference_system("N1000_Phi0.1.xyz", "mono.mod")
    me = sdt.MobEwald(rdi.size, rdi, box, 1e-12)
    assert me.init(), getattr(me, "init_error", "")
    me.update(pos)
    return pos, rdi, box, me


@pytest.mark.parametrize("xyz,mod,p,xi,rmax,prec,tol", [
    ("N1000_Phi0.1.xyz", "mono.mod", 6, 0.8, 8.371887, 0, 1e-6),
    ("N1000_Phi0.1.xyz", "mono.mod", 8, 0.8, 8.371887, 0, 1e-6),
    ("N412.xyz", "poly.mod", 8, 0.04, 200.37188, 0, 1e-6),
    ("N1000_Phi0.1.xyz", "mono.mod", 6, 0.8, 8.371887, 1, 1e-5),
    ("N412.xyz", "poly.mod", 8, 0.04, 200.37188, 1, 1e-5),
])
def test_testspme_relation(sdt, xyz, mod, p, xi, rmax, prec, tol):
    print("testspme", xyz, mod, p, prec, flush=True)
    pos, rdi, box = reference_system(xyz, mod)
    n = rdi.size
    f = np.random.default_rng(1234).uniform(0.0, 1.0, (3 * n, 3))
    me = sdt.MobEwald(n, rdi, box, 1e-12)
    assert me.init()
    me.update(pos)
    v_ewald = me.mul_vector(f, 2.1, 1.1, f)
    ms = sdt.MobSpme(n, rdi, box, xi, rmax, 256, p, prec)
    assert ms.init(), getattr(ms, "init_error", "")
    ms.update(pos)
    v_spme = ms.mul_vector(f, 2.1, 1.1, f)
    err = np.linalg.norm(v_ewald - v_spme) / np.linalg.norm(v_ewald)
    print("error = %.3e" % err, flush=True)
    assert err < tol


def test_teststeric_relation(sdt):
    pos, rdi, box = reference_system("N412.xyz", "poly.mod")
    st = sdt.StericForce(rdi.size, rdi, box, 2.3, 125.0)
    assert st.init()
    f = st.compute(pos)
    fd = direct_steric(pos, rdi, box, 2.3, 125.0)
    assert np.linalg.norm(fd) > 0
    assert np.linalg.norm(f - fd) / np.linalg.norm(fd) < 1e-6


def test_testlanczos_relation(sdt, n1000_ewald):
    pos, rdi, box, me = n1000_ewald
    nm = 3 * rdi.size
    M2 = _eig_sqrt(me.get_matrix())
    z = np.random.default_rng(1234).uniform(0.0, 1.0, (nm, 3))
    bl = sdt.BrwnLanczos(nm, 100, 16, 1e-7)
    assert bl.init()
    for nrhs in (1, 3):
        y, ok = bl.compute(me, z[:, :nrhs])
        assert ok
        err = np.linalg.norm(y - M2 @ z[:, :nrhs]) / np.linalg.norm(M2 @ z[:, :nrhs])
        print("error %d = %.3e" % (nrhs, err), flush=True)
        assert err < 1e-6


def test_mob_ewald_matches_oracle_and_matrix_entry(sdt, orc):
    pos, rdi, box = reference_system("N412.xyz", "poly.mod")
    n = rdi.size
    me = sdt.MobEwald(n, rdi, box, 1e-10)
    assert me.init()
    me.update(pos)
    M = me.get_matrix()
    ref = orc.ewald_matrix(pos, rdi, box, tol=1e-10)
    assert np.abs(M - ref).max() < 1e-12 * np.abs(ref).max()
    assert np.array_equal(M, sdt.rpy_ewald_matrix(pos, rdi, box, 1e-10))
    f = np.random.default_rng(3).standard_normal((3 * n, 5))
    v0 = np.random.default_rng(4).standard_normal((3 * n, 5))
    assert orc.frob_rel_err(0.5 * ref @ f - 2.0 * v0, me.mul_vector(f, 0.5, -2.0, v0)) < 1e-12
    # no real/recip split and no SPME-only entry points on a dense handle
    out = np.zeros(3 * n)
    fc = np.ascontiguousarray(f[:, 0])
    rc = sdt.lib().sdt_mob_spme_mulvector_real(me.h, 1, 1.0, 3 * n, sdt._ptr(fc), 0.0, 3 * n,
                                               sdt._ptr(out))
    assert rc == sdt.SDT_EINVAL


def test_brwn_chol_factor_and_oracle(sdt, orc, n1000_ewald):
    pos, rdi, box, me = n1000_ewald
    nm = 3 * rdi.size
    M = me.get_matrix()
    ch = sdt.BrwnChol(nm)
    assert ch.init()
    # y = L z for z = I is L itself: L L^T = M, L lower
    L = ch.compute(me, np.eye(nm))
    assert np.abs(np.triu(L, 1)).max() == 0.0
    assert np.abs(L @ L.T - M).max() < 1e-12 * np.abs(M).max()
    z = orc.gaussian(77, 0, nm * 4).reshape(4, nm).T.copy()
    y = ch.compute(me, z)
    yo, info = orc.chol_displacements(M, z)
    assert info == 0
    assert orc.frob_rel_err(yo, y) < 1e-11
    # both square roots reproduce the quadratic form of the covariance:
    # z^T M z = |L^T z|^2 = |M^{1/2} z|^2
    bl = sdt.BrwnLanczos(nm, 100, 4, 1e-10)
    assert bl.init()
    d, ok = bl.compute(me, z)
    assert ok
    q = np.einsum("ij,ij->j", z, M @ z)
    assert np.allclose((d * d).sum(0), q, rtol=1e-8)
    assert np.allclose(((L.T @ z) ** 2).sum(0), q, rtol=1e-10)


def test_brwn_chol_needs_matrix_mobility(sdt):
    n, box = 50, 12.0
    rdi = np.ones(n)
    pos = np.random.default_rng(1).uniform(0, box, 3 * n)
    ms = sdt.MobSpme(n, rdi, box, 0.9, 5.0, 24, 6, sdt.FP64)
    assert ms.init()
    ms.update(pos)
    ch = sdt.BrwnChol(3 * n)
    assert ch.init()
    with pytest.raises(sdt.SdtError):
        ch.compute(ms, np.ones((3 * n, 1)))
