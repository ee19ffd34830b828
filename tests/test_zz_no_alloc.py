"""sdt_diffusion_dev_noalloc, the allocation-free variant of sdt_diffusion
(tools/calc_diff/calc_diff.cc:100-140): device inputs and outputs, partial
sums in the library's static device array. Checked through the C ABI
against the allocating entry point on two calls in a row, the second with a
smaller input (the array keeps the first call's contents), bit for bit (the
same kernels in the same order).
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def _hip(sdt):
    # the HIP runtime the library itself is linked against (same soname, so
    # the loader hands back the instance already in the process)
    sdt.lib()
    h = ctypes.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    h.hipMalloc.restype = ctypes.c_int
    h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    h.hipMemcpy.restype = ctypes.c_int
    h.hipFree.argtypes = [ctypes.c_void_p]
    h.hipFree.restype = ctypes.c_int
    h.hipDeviceSynchronize.restype = ctypes.c_int
    return h


def _walk(rng, nframes, npos):
    steps = rng.standard_normal((nframes, npos, 3)) * 0.1
    return np.ascontiguousarray(np.cumsum(steps, axis=0) + rng.uniform(0, 20, (1, npos, 3)))


def test_noalloc_matches_allocating_entry(sdt):
    hip = _hip(sdt)
    L = sdt.lib()
    rng = np.random.default_rng(2026)
    cases = [(64, 1500, 0.01), (18, 300, 0.5)]   # (frames, particles, dt); second smaller
    cap = max(nf * n * 3 for nf, n, _ in cases)
    d_traj, d_D = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(d_traj), 8 * cap) == 0
    assert hip.hipMalloc(ctypes.byref(d_D), 8 * 64) == 0
    try:
        for nf, n, dt in cases:
            print("case frames=%d particles=%d" % (nf, n), flush=True)
            traj = _walk(rng, nf, n)
            ref = sdt.diffusion(traj, dt)
            assert hip.hipMemcpy(d_traj, traj.ctypes.data_as(ctypes.c_void_p), traj.nbytes,
                                 H2D) == 0
            nl = ctypes.c_int(0)
            rc = L.sdt_diffusion_dev_noalloc(d_traj, nf, n, dt, d_D, ctypes.byref(nl), None)
            assert rc == 0, L.sdt_last_error().decode()
            assert hip.hipDeviceSynchronize() == 0
            assert nl.value == nf // 2 - 1 == ref.size
            got = np.zeros(nl.value)
            assert hip.hipMemcpy(got.ctypes.data_as(ctypes.c_void_p), d_D, got.nbytes, D2H) == 0
            assert np.array_equal(got, ref), np.abs(got - ref).max()
        # sizes beyond the static array are refused before any launch
        nl = ctypes.c_int(0)
        rc = L.sdt_diffusion_dev_noalloc(d_traj, 2 * 600, 256 * 2048, 1.0, d_D, ctypes.byref(nl),
                                         None)
        assert rc == sdt.SDT_ECAP
        assert "static scratch array" in L.sdt_last_error().decode()
    finally:
        hip.hipFree(d_traj)
        hip.hipFree(d_D)
