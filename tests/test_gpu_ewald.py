"""GPU dense Ewald-summed RPY matrix (rpy_ewald_matrix_mex,
mex/rpy_ewald_matrix_mex.cc:10-64; MobDebug::GetMatrix over EwaldKernel,
src/mob/ewald.cc:86-330) against the oracle's restatement, and against the
GPU SPME operator the way tests/spme/testspme.cc compares SPME with the dense
Ewald sum."""
import numpy as np
import pytest

from conftest import poly_rsa_system, uniform_system

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


@pytest.mark.parametrize("mode", ["full", "real", "recip"])
def test_ewald_matrix_matches_oracle(sdt, orc, mode):
    pos, rdi = uniform_system(60, 10.0, 21)   # overlapping pairs included
    M = sdt.rpy_ewald_matrix(pos, rdi, 10.0, 1e-12, mode)
    full, re, rc = orc.ewald_matrix(pos, rdi, 10.0, tol=1e-12, parts=True)
    ref = {"full": full, "real": re, "recip": rc}[mode]
    assert M.shape == (180, 180)
    assert np.array_equal(M, M.T)
    assert _rel(M, ref) < 1e-12


def test_ewald_matrix_poly_and_explicit_xi(sdt, orc):
    rng = np.random.default_rng(22)
    pos = rng.uniform(0.0, 14.0, 3 * 80)
    rdi = rng.uniform(0.5, 1.8, 80)
    xi = 0.6
    M = sdt.rpy_ewald_matrix(pos, rdi, 14.0, 1e-10, "full", xi)
    ref = orc.ewald_matrix(pos, rdi, 14.0, xi=xi, tol=1e-10)
    assert _rel(M, ref) < 1e-12
    # the sum does not depend on the splitting (to the truncation tolerance)
    M2 = sdt.rpy_ewald_matrix(pos, rdi, 14.0, 1e-10, "full")
    assert _rel(M2, M) < 1e-8
    # positive definite (RPY)
    assert np.linalg.eigvalsh(M).min() > 0


def test_ewald_matrix_pos_layouts_and_errors(sdt):
    pos, rdi = uniform_system(20, 9.0, 23)
    a = sdt.rpy_ewald_matrix(pos, rdi, 9.0, 1e-12)
    b = sdt.rpy_ewald_matrix(pos.reshape(20, 3).T, rdi, 9.0, 1e-12)
    assert np.array_equal(a, b)
    with pytest.raises(Exception):
        sdt.rpy_ewald_matrix(pos, rdi, 9.0, 1e-12, "dense")
    with pytest.raises(Exception):
        sdt.rpy_ewald_matrix(pos, rdi, -1.0, 1e-12)


def test_ewald_matrix_matches_gpu_spme(sdt):
    # tests/spme/testspme.cc: SPME within 1e-6 of the dense Ewald sum
    pos, rdi = uniform_system(200, 20.0, 12)
    M = sdt.rpy_ewald_matrix(pos, rdi, 20.0, 1e-12)
    g = sdt.MobSpme(200, rdi, 20.0, 0.8, 8.3, 128, 8, sdt.FP64)
    assert g.init()
    g.update(pos)
    f = np.random.default_rng(7).uniform(0, 1, (600, 3))
    v = g.mul_vector(f)
    assert np.linalg.norm(v - M @ f) / np.linalg.norm(M @ f) < 1e-7


@pytest.mark.parametrize("prec,tol", [(0, 1e-6), (1, 1e-5)])
def test_spme_matches_ewald_n412_reference_params(sdt, prec, tol):
    # tests/spme/test.sh (N412 poly, radii 21..115 in a box of 1000): dim 256,
    # porder 8, xi 0.04, rmax 200.37188, alpha 2.1, beta 1.1, < 1e-6 (FP64)
    pos, rdi = poly_rsa_system(412, 1000.0, 19)
    # an explicit splitting: the default 10^(1/6) sqrt(pi) / L = 0.0026 needs
    # ~3e5 k vectors per pair in this dilute box (the sum does not depend on it)
    M = sdt.rpy_ewald_matrix(pos, rdi, 1000.0, 1e-12, "full", 0.01)
    g = sdt.MobSpme(412, rdi, 1000.0, 0.04, 200.37188, 256, 8, prec)
    assert g.init()
    g.update(pos)
    f = np.random.default_rng(7).uniform(0, 1, (1236, 3))
    v = g.mul_vector(f, alpha=2.1, beta=1.1, v=f)
    ve = 2.1 * (M @ f) + 1.1 * f
    assert np.linalg.norm(v - ve) / np.linalg.norm(ve) < tol


@pytest.mark.parametrize("p,tol", [(5, 2e-5), (7, 1e-6)])
def test_odd_order_spme_matches_ewald(sdt, p, tol):
    # odd porder with the Nyquist modulus taken as the neighbour mean
    pos, rdi = uniform_system(200, 20.0, 12)
    M = sdt.rpy_ewald_matrix(pos, rdi, 20.0, 1e-12)
    g = sdt.MobSpme(200, rdi, 20.0, 0.8, 8.3, 128, p, sdt.FP64)
    assert g.init()
    g.update(pos)
    f = np.random.default_rng(7).uniform(0, 1, (600, 3))
    v = g.mul_vector(f)
    assert np.linalg.norm(v - M @ f) / np.linalg.norm(M @ f) < tol
