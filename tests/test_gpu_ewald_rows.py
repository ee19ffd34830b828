"""Ewald-summed M f on a sample of rows (EwaldVectorKernel,
src/mob/ewald.cc:333-536), the matrix never formed: the reciprocal sum runs
over structure factors of f instead of pair by pair.

- At one splitting it equals the rows of the dense Ewald matrix times f (the
  same image and k tables, so only the summation order differs), against the
  oracle's dense matrix and the GPU's, with overlapping pairs and unequal
  radii (the a_j^2 structure factor).
- The sum does not depend on the splitting beyond the truncation tolerance: a
  wrong sign or factor in either the real or the reciprocal part breaks this.
- At N = 20 000 it agrees with an accurate FP64 SPME product on the sampled
  rows (tests/spme/testspme.cc compares SPME and Ewald the same way).
"""
import numpy as np
import pytest

from conftest import box_for, reference_system, uniform_system

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def test_rows_equal_oracle_dense_matrix(sdt, orc):
    pos, rdi = uniform_system(60, 10.0, 21)   # overlapping pairs included
    f = np.random.default_rng(3).standard_normal(180)
    rows = np.array([0, 7, 13, 59, 31])
    xi = 0.7
    v, x = sdt.rpy_ewald_rows(pos, rdi, 10.0, 1e-12, f, rows, xi)
    assert x == xi
    M = orc.ewald_matrix(pos, rdi, 10.0, xi=xi, tol=1e-12)
    ref = (M @ f).reshape(-1, 3)[rows]
    assert _rel(v, ref) < 1e-12


@pytest.mark.parametrize("which", ["poly_overlapping", "n412"])
def test_rows_poly_equal_gpu_dense_and_split_independent(sdt, which):
    if which == "poly_overlapping":
        L = 14.0
        rng = np.random.default_rng(33)
        pos, rdi = rng.uniform(0.0, L, 3 * 150), rng.uniform(0.5, 1.8, 150)
    else:
        # the reference's N412 input with poly.mod radii
        pos, rdi, L = reference_system("N412.xyz", "poly.mod")
    n = rdi.size
    f = np.random.default_rng(4).standard_normal(3 * n)
    rows = np.random.default_rng(5).choice(n, 40, replace=False)
    xi = 3.0 / L
    M = sdt.rpy_ewald_matrix(pos, rdi, L, 1e-11, "full", xi)
    ref = (M @ f).reshape(-1, 3)[rows]
    v, _ = sdt.rpy_ewald_rows(pos, rdi, L, 1e-11, f, rows, xi)
    assert _rel(v, ref) < 1e-12
    # any other splitting, including the one the sample picks for itself
    for x in (None, 8.0 / L):
        v2, used = sdt.rpy_ewald_rows(pos, rdi, L, 1e-11, f, rows, x)
        assert used > 0
        assert _rel(v2, ref) < 1e-8, (x, used)


def test_rows_at_scale_match_fp64_spme(sdt):
    n = 20000
    L = box_for(n, 0.2)
    pos = sdt.rsa_positions(n, L, 1.0, 17)
    rdi = np.ones(n)
    f = np.random.default_rng(6).standard_normal(3 * n)
    rows = np.random.default_rng(7).choice(n, 300, replace=False)
    v, xi = sdt.rpy_ewald_rows(pos, rdi, L, 1e-10, f, rows)
    print("N=20000 L=%.3f xi=%.4f" % (L, xi), flush=True)
    # SPME with the real sum to erfc(4.3) and an 8th-order spline on a
    # spacing of 0.58 (xi h = 0.26); its error falls with xi h, and the
    # coarser setting (xi h = 0.35) is measurably further from the Ewald sum
    errs = []
    for xs, rc in ((0.45, 9.5), (0.6, 7.0)):
        m = sdt.MobSpme(n, rdi, L, xs, rc, 128, 8, sdt.FP64)
        assert m.init()
        m.update(pos)
        errs.append(_rel(v, m.mul_vector(f).reshape(-1, 3)[rows]))
    print("SPME vs Ewald rows:", errs, flush=True)
    assert errs[0] < 1e-5 and errs[1] > 2.0 * errs[0]


def test_rows_argument_errors(sdt):
    pos, rdi = uniform_system(20, 9.0, 23)
    f = np.zeros(60)
    with pytest.raises(Exception):
        sdt.rpy_ewald_rows(pos, rdi, 9.0, 1e-10, f, [20])
    with pytest.raises(Exception):
        sdt.rpy_ewald_rows(pos, rdi, -9.0, 1e-10, f, [0])
    with pytest.raises(Exception):
        sdt.rpy_ewald_rows(pos, rdi, 9.0, 1e-10, f[:30], [0])
