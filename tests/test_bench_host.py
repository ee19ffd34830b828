"""Host logic of bench.py (no GPU): the amortisation of a timed window over
mobility intervals, the algorithmic byte counts behind `roofline`, the
committed traffic summary it reads, and the seeded workload generators of
the ensemble and serving legs."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


def test_amortise_whole_intervals_and_composed():
    mi = 100
    # a window of exactly one interval: steps / time
    lat = [0.6] + [0.0004] * 99
    v, how = bench.amortise(lat, 100, mi)
    assert v == pytest.approx(100 / sum(lat)) and "whole" in how
    # 20 steps starting on the interval boundary: start + 99 x mean tail
    lat = [0.6] + [0.0004] * 19
    v, how = bench.amortise(lat, 200, mi)
    assert v == pytest.approx(100 / (0.6 + 99 * 0.0004)) and "composed" in how
    # a window that misses the boundary has no solve in it: refused
    with pytest.raises(RuntimeError):
        bench.amortise([0.0004] * 20, 101, mi)


def test_algorithmic_bytes_of_the_workload():
    c = bench.CFG
    n, K = c["n"], c["K"]
    # interp and spread at 100 columns, FP32: the meshes (3 K^3 x 4 bytes
    # per column), the particle offsets and bases, the FP64 vectors
    b100 = bench.alg_bytes("interp", 100, n, K, 4)
    assert b100 == 3 * 100 * K ** 3 * 4 + 3 * n * 8 + 3 * n * 100 * 8
    assert 2.75e9 < b100 < 2.77e9
    assert bench.alg_bytes("spread", 100, n, K, 4) == b100
    # forward transform: real meshes in, half spectra out (50.7 MB per column)
    f1 = bench.alg_bytes("fft_xy_fwd", 1, n, K, 4)
    assert f1 == 3 * (K ** 3 * 4 + K * K * (K // 2 + 1) * 8)
    assert bench.alg_bytes("fft_xy_fwd", 100, n, K, 4) == 100 * f1
    assert bench.alg_bytes("fft_z_influence", 1, n, K, 4) == 2 * 3 * K * K * (K // 2 + 1) * 8
    # polydisperse: six meshes per column
    assert bench.alg_bytes("fft_xy_inv", 1, n, K, 4, poly=True) == 2 * bench.alg_bytes(
        "fft_xy_inv", 1, n, K, 4)
    assert bench.alg_bytes("unknown", 1, n, K, 4) == 0


def test_committed_traffic_summary_is_read():
    t = bench.read_traffic("fft_xy_fwd")
    assert 100 in t and 1 in t
    # the counter traffic of the forward transform equals its algorithmic bytes
    alg = bench.alg_bytes("fft_xy_fwd", 100, bench.CFG["n"], bench.CFG["K"], 4)
    assert abs(t[100] - alg) < 0.01 * alg


def _smooth(k):
    for f in (2, 3, 5):
        while k % f == 0:
            k //= f
    return k == 1


def test_smooth_mesh():
    for L in (20.0, 37.3, 55.0, 76.8, 200.0):
        k = bench.smooth_mesh(L, 24, 96)
        assert 24 <= k <= 96 and k % 2 == 0 and _smooth(k)
        # no admissible mesh is nearer to L / 0.8
        t = L / 0.8
        assert all(abs(k - t) <= abs(j - t) for j in range(24, 97, 2) if _smooth(j))


def test_serving_requests_seeded_poisson():
    a = bench.serve_requests(2.0)
    b = bench.serve_requests(2.0)
    assert len(a) == len(b) and all(x["n"] == y["n"] and x["arrival"] == y["arrival"]
                                    for x, y in zip(a, b))
    # about 50 arrivals per second
    assert 60 <= len(a) <= 140
    c = bench.SERVE
    for r in a:
        assert c["nmin"] <= r["n"] <= c["nmax"] and r["dim"] % 2 == 0 and _smooth(r["dim"])
        assert 0.0 < r["arrival"] < 2.0
    # lattice placement: no two unit spheres overlap (checked on the smallest)
    r = min(a, key=lambda x: x["n"])
    p = r["pos"].reshape(-1, 3)
    d = p[:, None, :] - p[None, :, :]
    d -= r["L"] * np.round(d / r["L"])
    rr = np.sqrt((d * d).sum(-1)) + np.eye(len(p)) * 10.0
    assert rr.min() > 2.0


def test_lattice_positions_refuses_crowding():
    rng = np.random.default_rng(0)
    with pytest.raises(ValueError):
        bench.lattice_positions(rng, 1000, 15.0)
