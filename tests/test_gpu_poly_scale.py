"""The polydisperse SPME path at scale (the three round trips of
src/mob/spme.cc:771-838, folded into two interpolated outputs here) at
N = 20 000 with radii uniform in [0.5, 1.5]:

- FP64 against the oracle's restatement (1e-10) and FP32 against FP64 (1e-5);
- a converged FP64 SPME operator against the dense Ewald sum on sampled rows
  (EwaldVectorKernel, src/mob/ewald.cc:333-536), which pins the unequal-radius
  a_i^2 + a_j^2 terms of both sums at this size;
- a BD interval with the polydisperse mobility against the FP64 path.
"""
import numpy as np
import pytest

from conftest import box_for

pytestmark = pytest.mark.gpu

N = 20_000


@pytest.fixture(scope="module")
def system(sdt):
    # sequential addition at the largest radius, then smaller radii drawn:
    # no pair overlaps
    L = box_for(N, 0.3, 1.5)
    pos = sdt.rsa_positions(N, L, 1.5, 91)
    rdi = np.random.default_rng(92).uniform(0.5, 1.5, N)
    return pos, rdi, L


def test_poly_fp64_matches_oracle_fp32_within_1e5(sdt, orc, system):
    pos, rdi, L = system
    f = np.random.default_rng(93).standard_normal((3 * N, 4))
    out = {}
    for prec in (sdt.FP64, sdt.FP32):
        g = sdt.MobSpme(N, rdi, L, 0.6, 6.0, 96, 6, prec)
        assert g.init()
        g.update(pos)
        out[prec] = g.mul_vector(f)
    o = orc.MobSpme(N, rdi, L, 0.6, 6.0, 96, 6)
    assert o.init()
    o.update(pos, rdi)
    ref = o.mul_vector(f)
    assert orc.frob_rel_err(ref, out[sdt.FP64]) < 1e-10
    assert orc.frob_rel_err(out[sdt.FP64], out[sdt.FP32]) < 1e-5


def test_poly_converged_spme_matches_ewald_rows(sdt, orc, system):
    pos, rdi, L = system
    f = np.random.default_rng(94).standard_normal(3 * N)
    rows = np.random.default_rng(95).choice(N, 400, replace=False)
    v_ew, xi = sdt.rpy_ewald_rows(pos, rdi, L, 1e-10, f, rows)
    print("Ewald rows xi %.4f" % xi, flush=True)
    g = sdt.MobSpme(N, rdi, L, 0.4, 10.0, 192, 8, sdt.FP64)
    assert g.init()
    g.update(pos)
    v = g.mul_vector(f).reshape(-1, 3)[rows]
    err = np.linalg.norm(v - v_ew) / np.linalg.norm(v_ew)
    print("converged poly SPME vs Ewald rows: %.3e" % err, flush=True)
    assert err < 1e-5
    # the radius-dependent part is not negligible at these radii: the same
    # rows with unit radii differ by far more than the tolerance
    v1, _ = sdt.rpy_ewald_rows(pos, np.ones(N), L, 1e-10, f, rows)
    assert np.linalg.norm(v1 - v_ew) > 100 * err * np.linalg.norm(v_ew)


def test_poly_bd_interval_fp32_vs_fp64(sdt, orc, system):
    pos, rdi, L = system
    runs = {}
    for prec in (sdt.FP64, sdt.FP32):
        bd = sdt.BDSim(pos, rdi, L, 0.6, 6.0, 96, 6, prec, 100, 8, 1e-6, 2.0, 125.0, 1e-3, 10, 96)
        bd.init()
        bd.advance(10)
        runs[prec] = bd.positions() - pos
    assert orc.frob_rel_err(runs[sdt.FP64], runs[sdt.FP32]) < 1e-5
