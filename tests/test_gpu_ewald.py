"""GPU dense Ewald-summed RPY matrix (rpy_ewald_matrix_mex,
mex/rpy_ewald_matrix_mex.cc:10-64; MobDebug::GetMatrix over EwaldKernel,
src/mob/ewald.cc:86-330) against the oracle's restatement, and against the
GPU SPME operator the way tests/spme/testspme.cc compares SPME with the dense
Ewald sum."""
import numpy as np
import pytest

from conftest import uniform_system

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
