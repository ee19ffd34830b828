"""GPU parity of the SPME mobility (MobSpme::MulVector / RealMulVector /
RecipMulVector, src/mob/mob_spme.cc:219-262, and rpy_spme_mex,
mex/rpy_spme_mex.cc:8-75) against the oracle.

Tolerances: FP64 GPU vs FP64 CPU <= 1e-10 relative 2-norm; FP32-mixed vs
FP64 GPU <= 1e-5 (the north-star gates).
"""
import os

import numpy as np
import pytest

from conftest import uniform_system

pytestmark = pytest.mark.gpu
G = np.load(os.path.join(os.path.dirname(__file__), "golden", "golden_small.npz"))

CASES = [
    # (n, box, xi, rmax, K, p, poly, nrhs)
    (64, 12.0, 0.9, 5.0, 24, 6, False, 3),
    (300, 20.0, 0.9, 5.0, 32, 6, False, 1),
    (300, 20.0, 0.9, 5.0, 32, 4, False, 2),
    (300, 20.0, 0.9, 5.0, 36, 5, False, 3),
    (300, 20.0, 0.9, 5.0, 32, 7, True, 2),
    (250, 20.0, 0.8, 8.3, 48, 8, False, 17),
    (200, 16.0, 1.0, 4.0, 40, 6, True, 4),
    (300, 20.0, 0.9, 5.0, 64, 6, True, 2),
    # K not a multiple of 4: the row-owner spread kernel
    (300, 20.0, 0.9, 5.0, 30, 6, False, 3),
    (300, 20.0, 0.9, 5.0, 50, 5, True, 2),
]


def _pair(sdt, orc, n, box, xi, rmax, K, p, poly, seed, prec):
    pos, rdi = uniform_system(n, box, seed)
    if poly:
        rdi = np.random.default_rng(seed + 1).uniform(0.7, 1.3, n)
    g = sdt.MobSpme(n, rdi, box, xi, rmax, K, p, prec)
    assert g.init(), getattr(g, "init_error", "")
    g.update(pos)
    o = orc.MobSpme(n, rdi, box, xi, rmax, K, p)
    assert o.init()
    o.update(pos, rdi)
    return g, o, pos, rdi


@pytest.mark.parametrize("case", CASES)
def test_fp64_matches_oracle(sdt, orc, case):
    n, box, xi, rmax, K, p, poly, nrhs = case
    print("case", case, flush=True)
    g, o, _, _ = _pair(sdt, orc, n, box, xi, rmax, K, p, poly, 100 + K, sdt.FP64)
    rng = np.random.default_rng(K)
    f = rng.uniform(0, 1, (3 * n, nrhs))
    v0 = rng.uniform(0, 1, (3 * n, nrhs))
    for mode, fn in (("full", g.mul_vector), ("real", g.real_mul_vector),
                     ("recip", g.recip_mul_vector)):
        out = fn(f, 2.1, 1.1, v0)
        ref = o.mul_vector(f, 2.1, 1.1, v0, mode=mode)
        assert orc.frob_rel_err(ref, out) < 1e-10, mode
    assert g.real_nnz() == o.real_nnz()


def test_fp64_matches_golden(sdt, orc):
    for tag in ("mono", "poly"):
        box, xi, rmax, K, p = G[tag + "_params"]
        pos, rdi, f = G[tag + "_pos"], G[tag + "_rdi"], G[tag + "_f"]
        g = sdt.MobSpme(rdi.size, rdi, box, xi, rmax, int(K), int(p), sdt.FP64)
        assert g.init()
        g.update(pos)
        for mode, fn in (("full", g.mul_vector), ("real", g.real_mul_vector),
                         ("recip", g.recip_mul_vector)):
            out = fn(f, 2.1, 1.1, G["mono_v0"]) if tag == "mono" else fn(f)
            assert orc.frob_rel_err(G[tag + "_" + mode], out) < 1e-10, (tag, mode)


@pytest.mark.parametrize("poly,p", [(False, 6), (True, 6), (False, 4), (False, 7), (True, 8)])
def test_fp32_within_1e5_of_fp64(sdt, orc, poly, p):
    # every spline order of the FP32 kernels (p = 7, 8 narrow the interp strip)
    n, box = 2000, 30.0
    pos, rdi = uniform_system(n, box, 7)
    if poly:
        rdi = np.random.default_rng(8).uniform(0.7, 1.3, n)
    f = np.random.default_rng(9).standard_normal((3 * n, 4))
    outs = []
    for prec in (sdt.FP64, sdt.FP32):
        g = sdt.MobSpme(n, rdi, box, 0.9, 5.0, 64, p, prec)
        assert g.init()
        g.update(pos)
        outs.append(g.mul_vector(f))
    assert orc.frob_rel_err(outs[0], outs[1]) < 1e-5


@pytest.mark.parametrize("prec,tol", [(0, 1e-12), (1, 1e-6)])
def test_spread_mesh_matches_oracle(sdt, orc, prec, tol):
    n, box, K, p = 300, 20.0, 32, 6
    pos, rdi = uniform_system(n, box, 3)
    g = sdt.MobSpme(n, rdi, box, 0.9, 5.0, K, p, prec)
    assert g.init()
    g.update(pos)
    e = orc.SpmeEngine(n, None, box, 0.9, K, p)
    e.update(pos)
    f = np.random.default_rng(4).standard_normal(3 * n)
    ref = e.spread(f)
    assert np.abs(g.debug_spread(f) - ref).max() <= tol * np.abs(ref).max()


@pytest.mark.parametrize("prec,tol,K,p", [(0, 1e-12, 32, 6), (1, 1e-6, 32, 6),
                                            (0, 1e-12, 64, 7), (1, 1e-6, 48, 8)])
def test_interp_matches_oracle(sdt, orc, prec, tol, K, p):
    # the interpolation step alone (src/mob/spme.cc:163-198) on random meshes
    n, box = 300, 20.0
    pos, rdi = uniform_system(n, box, 3)
    g = sdt.MobSpme(n, rdi, box, 0.9, 5.0, K, p, prec)
    assert g.init()
    g.update(pos)
    e = orc.SpmeEngine(n, None, box, 0.9, K, p)
    e.update(pos)
    meshes = np.random.default_rng(5).standard_normal((3, K, K, K))
    ref = e.interp(meshes, 1.7)
    out = g.debug_interp(meshes, 1.7)
    assert np.abs(out - ref).max() <= tol * np.abs(ref).max()


def test_rpy_spme_one_shot(sdt, orc):
    n, box = 200, 18.0
    pos, rdi = uniform_system(n, box, 5)
    f = np.random.default_rng(6).standard_normal((3 * n, 2))
    o = orc.MobSpme(n, rdi, box, 0.9, 5.0, 32, 6)
    assert o.init()
    o.update(pos, rdi)
    for mode in ("full", "real", "recip"):
        out = sdt.rpy_spme(pos.reshape(n, 3).T, rdi, box, f, mode, 0.9, 5.0, 32, 6)
        assert orc.frob_rel_err(o.mul_vector(f, mode=mode), out) < 1e-10


def test_positions_outside_box_are_wrapped(sdt):
    # the reference discards the periodic wrap in the SPME update
    # (src/mob/spme.cc:404-412); this path wraps like the pair lists
    n, box = 150, 16.0
    pos, rdi = uniform_system(n, box, 11)
    shift = np.random.default_rng(12).integers(-2, 3, 3 * n) * box
    f = np.random.default_rng(13).standard_normal((3 * n, 2))
    outs = []
    for p in (pos, pos + shift):
        g = sdt.MobSpme(n, rdi, box, 0.9, 5.0, 32, 6, sdt.FP64)
        assert g.init()
        g.update(p)
        outs.append(g.mul_vector(f))
    assert np.abs(outs[0] - outs[1]).max() < 1e-11 * np.abs(outs[0]).max()


def test_operator_symmetric_and_linear(sdt):
    n, box = 1500, 28.0
    pos, rdi = uniform_system(n, box, 14)
    g = sdt.MobSpme(n, rdi, box, 0.9, 5.0, 48, 6, sdt.FP64)
    assert g.init()
    g.update(pos)
    rng = np.random.default_rng(15)
    x, y = rng.standard_normal(3 * n), rng.standard_normal(3 * n)
    Mx, My = g.mul_vector(x)[:, 0], g.mul_vector(y)[:, 0]
    assert abs(x @ My - y @ Mx) < 1e-12 * abs(x @ My)
    Mxy = g.mul_vector(np.stack([x, y, 2 * x - 3 * y], 1))
    assert np.abs(Mxy[:, 2] - (2 * Mx - 3 * My)).max() < 1e-12 * np.abs(Mxy).max()


def test_init_and_state_errors(sdt):
    n = 10
    pos, rdi = uniform_system(n, 20.0, 1)
    g = sdt.MobSpme(n, rdi, 20.0, 0.9, 5.0, 20, 6, sdt.FP64)   # 4p > K
    assert not g.init() and "dim/4" in g.init_error
    g = sdt.MobSpme(n, rdi, 20.0, 0.9, 5.0, 28, 4, sdt.FP64)   # 28 = 4 * 7
    assert not g.init()
    g = sdt.MobSpme(n, rdi, 20.0, 0.9, 10.0, 32, 6, sdt.FP64)  # 2 rmax >= L
    assert not g.init()
    g = sdt.MobSpme(n, rdi, 20.0, 0.9, 5.0, 32, 6, sdt.FP64)
    assert g.init()
    with pytest.raises(sdt.SdtError) as e:
        g.mul_vector(np.zeros(3 * n))
    assert e.value.code == sdt.SDT_ESTATE


def test_large_system_fp32_vs_fp64(sdt, orc):
    # size-independent properties at the N = 100 000, phi = 0.3, K = 128,
    # p = 6 shape: FP32 within 1e-5 of FP64, FP64 operator symmetric. The 40
    # columns go through in one launch (as the block-Lanczos batches do): the
    # FP32 spread sums three column blocks, the last one partial, on the
    # matrix cores in one workgroup per mesh block.
    n = 100_000
    box = (4.0 * np.pi * n / (3.0 * 0.3)) ** (1.0 / 3.0)
    pos, rdi = uniform_system(n, box, 2024)
    rng = np.random.default_rng(1)
    f = rng.standard_normal((3 * n, 40))
    outs = []
    for prec in (sdt.FP64, sdt.FP32):
        g = sdt.MobSpme(n, rdi, box, 0.896, 5.0, 128, 6, prec)
        assert g.init()
        g.update(pos)
        g.set_launch_width(64)
        outs.append(g.mul_vector(f))
    assert orc.frob_rel_err(outs[0], outs[1]) < 1e-5
    x, y = f[:, 0], f[:, 1]
    assert abs(x @ outs[0][:, 1] - y @ outs[0][:, 0]) < 1e-11 * abs(x @ outs[0][:, 1])


@pytest.mark.parametrize("prec,tol", [(0, 1e-10), (1, 1e-5)])
def test_workload_size_matches_oracle(sdt, orc, prec, tol):
    # the N = 100 000, phi = 0.3 (L = 111.7693), K = 128, p = 6, rc = 5,
    # xi = 0.896 system from random sequential addition, 12 right-hand sides:
    # dense tiles put more particles into a spread workgroup's staging list
    # than one pass of it holds
    n = 100_000
    box = (4.0 * np.pi * n / (3.0 * 0.3)) ** (1.0 / 3.0)
    pos = sdt.rsa_positions(n, box, 1.0, 2024)
    rdi = np.ones(n)
    g = sdt.MobSpme(n, rdi, box, 0.896, 5.0, 128, 6, prec)
    assert g.init()
    g.update(pos)
    o = orc.MobSpme(n, rdi, box, 0.896, 5.0, 128, 6)
    assert o.init()
    o.update(pos, rdi)
    f = np.random.default_rng(11).standard_normal((3 * n, 12))
    for mode in ("recip", "real"):
        got = getattr(g, mode + "_mul_vector")(f)
        ref = o.mul_vector(f, mode=mode)
        assert orc.frob_rel_err(ref, got) < tol, mode


@pytest.mark.parametrize("K", [160, 192])
def test_large_mesh_fft_paths(sdt, orc, K):
    # K = 160: FP32 uses the wide fused-plane FFT variant, FP64 the separate
    # x / y passes (a plane no longer fits LDS); K = 192: both separate
    n, box = 400, 40.0
    g, o, pos, rdi = _pair(sdt, orc, n, box, 0.9, 6.0, K, 6, False, 21, sdt.FP64)
    f = np.random.default_rng(22).standard_normal((3 * n, 3))
    ref = o.mul_vector(f, mode="recip")
    got = g.recip_mul_vector(f)
    assert orc.frob_rel_err(ref, got) < 1e-10
    g32 = sdt.MobSpme(n, rdi, box, 0.9, 6.0, K, 6, sdt.FP32)
    assert g32.init()
    g32.update(pos)
    assert orc.frob_rel_err(got, g32.recip_mul_vector(f)) < 1e-5


@pytest.mark.parametrize("prec,tol", [(0, 1e-10), (1, 1e-5)])
def test_blocked_input_paths_match_oracle(sdt, orc, prec, tol):
    # 40 columns run as blocks of 16, 16 and 8: the spread and the real space
    # share one column-blocked copy of the input where their block widths
    # agree (FP32 16, FP64 8) and the spread reads the packed input otherwise
    n, box = 1500, 28.0
    g, o, _, _ = _pair(sdt, orc, n, box, 0.9, 5.0, 48, 6, False, 31, prec)
    f = np.random.default_rng(32).standard_normal((3 * n, 40))
    for mode, fn in (("full", g.mul_vector), ("recip", g.recip_mul_vector)):
        ref = o.mul_vector(f, mode=mode)
        assert orc.frob_rel_err(ref, fn(f)) < tol, mode


@pytest.mark.parametrize("n,box,K,p,prec,b", [
    (20000, 74.8, 128, 6, 1, 16),    # the workload's mesh and order
    (3000, 30.0, 48, 8, 1, 9),       # p = 8
    (1500, 28.0, 30, 6, 1, 12),      # tile side 6
    (800, 20.0, 50, 5, 0, 8),        # FP64, tile side 5
    (300, 20.0, 24, 4, 1, 3),        # narrow batch
])
def test_interp_meshes_per_pass_agree(sdt, orc, n, box, K, p, prec, b):
    # the interp kernel stages 1, 2 or 4 meshes per LDS pass (vector reads);
    # every width sums each stencil in the same order, so all agree with the
    # oracle to the precision's tolerance and with each other to rounding
    pos, rdi = uniform_system(n, box, 41)
    g = sdt.MobSpme(n, rdi, box, 0.9, 5.0, K, p, prec)
    assert g.init()
    g.update(pos)
    f = np.random.default_rng(42).standard_normal((3 * n, b))
    o = orc.MobSpme(n, rdi, box, 0.9, 5.0, K, p)
    assert o.init()
    o.update(pos, rdi)
    ref = o.mul_vector(f, mode="recip")
    out = {}
    for mv in ((1, 2, 4) if prec == 1 else (1, 2)):
        g.set_interp_form(mv)
        out[mv] = g.recip_mul_vector(f)
        assert orc.frob_rel_err(ref, out[mv]) < (1e-5 if prec == 1 else 1e-10), mv
    g.set_interp_form(0)
    for mv in out:
        assert orc.frob_rel_err(out[1], out[mv]) < (1e-6 if prec == 1 else 1e-14), mv


def test_interp_meshes_per_pass_refused(sdt):
    pos, rdi = uniform_system(200, 16.0, 43)
    g = sdt.MobSpme(200, rdi, 16.0, 1.0, 4.0, 40, 6, sdt.FP64)
    assert g.init()
    g.update(pos)
    g.set_interp_form(4)
    with pytest.raises(Exception):
        g.recip_mul_vector(np.ones((600, 8)))
    with pytest.raises(Exception):
        g.set_interp_form(3)


@pytest.mark.parametrize("prec", [0, 1])
def test_z_transform_launch_shapes_agree(sdt, orc, prec):
    # the z transform and influence run with 1024, 512 or 256 threads and
    # 16, 8 or 4 pencils per workgroup; each pencil's arithmetic is the same
    # in every shape, so the results are identical
    n, box = 2000, 30.0
    g, o, _, _ = _pair(sdt, orc, n, box, 0.9, 5.0, 64, 6, False, 51, prec)
    f = np.random.default_rng(52).standard_normal((3 * n, 5))
    ref = None
    for z in (1, 2, 3, 4, 5):
        g.set_forms(0, z)
        out = g.recip_mul_vector(f)
        if ref is None:
            ref = out
            assert orc.frob_rel_err(o.mul_vector(f, mode="recip"), out) < (1e-10 if prec == 0 else 1e-5)
        assert np.array_equal(out, ref), z
    g.set_forms(0, 0)
    with pytest.raises(Exception):
        g.set_forms(0, 6)


@pytest.mark.parametrize("prec,tol", [(0, 1e-13), (1, 1e-6)])
def test_wide_full_product_overlapped_streams(sdt, orc, prec, tol):
    # wide full products run the real space on a second stream beside the
    # spread and transforms, joined before the interpolation adds the
    # reciprocal part: the sum of the two parts, deterministic from call to
    # call, and equal to the oracle
    n, box = 3000, 30.0
    g, o, _, _ = _pair(sdt, orc, n, box, 0.9, 5.0, 48, 6, False, 61, prec)
    f = np.random.default_rng(62).standard_normal((3 * n, 24))
    v0 = np.random.default_rng(63).standard_normal((3 * n, 24))
    full = g.mul_vector(f, 1.5, 0.5, v0)
    again = g.mul_vector(f, 1.5, 0.5, v0)
    assert np.array_equal(full, again)
    parts = 1.5 * (g.real_mul_vector(f) + g.recip_mul_vector(f)) + 0.5 * v0
    assert orc.frob_rel_err(parts, full) < tol
    ref = o.mul_vector(f, mode="full")
    assert orc.frob_rel_err(1.5 * ref + 0.5 * v0, full) < (1e-10 if prec == 0 else 1e-5)


@pytest.mark.parametrize("prec,K,p", [(0, 48, 6), (1, 48, 6), (1, 100, 6), (1, 120, 8), (1, 64, 4)])
def test_spread_forms_identical(sdt, orc, prec, K, p):
    # the point-owner spread sums wide FP32 launches on the vector ALUs
    # (forces from scalar loads or LDS) or on the matrix cores (one
    # workgroup per column block, or per mesh block over all column blocks),
    # wide FP64 launches on the vector ALUs per column block or over all of
    # them (form 1 + 8, the built-in one), and stores its meshes per lane or
    # as whole rows through LDS. The
    # matrix cores take 4 candidates per step in list order with one
    # rounding per product-add, so every form gives the same sums. The full
    # product (b >= 8) spreads from the column-blocked copy, which the
    # matrix-core forms need; each product is one launch (b = 20 and 40 are
    # two and three column blocks, the last one partial); K = 48, 100 and
    # 120 leave a 16-, a 4- (one active wave) and a 24-point partial x block,
    # p = 4 and 8 the shortest and longest stencils.
    n, box = 3000, 30.0
    g, o, _, _ = _pair(sdt, orc, n, box, 0.9, 5.0, K, p, False, 81, prec)
    g.set_launch_width(64)
    forms = (0, 1, 2, 3, 4, 5, 6, 7, 9, 11, 15)
    for b in (1, 20, 40):
        f = np.random.default_rng(82 + b).standard_normal((3 * n, b))
        out = {}
        for form in forms:
            g.set_forms(0, 0, form)
            out[form] = (g.recip_mul_vector(f), g.mul_vector(f))
        g.set_forms(0, 0, 0)
        for form in forms[1:]:
            for k in range(2):
                assert np.array_equal(out[0][k], out[form][k]), (b, form, k)
        tol = 1e-10 if prec == 0 else 1e-5
        assert orc.frob_rel_err(o.mul_vector(f, mode="recip"), out[0][0]) < tol
        assert orc.frob_rel_err(o.mul_vector(f), out[0][1]) < tol
    with pytest.raises(Exception):
        g.set_forms(0, 0, 16)


def test_spread_stage_overflow_same_sums(sdt, orc):
    # a 16 KB stage holds 50 candidates of the matrix-core spread, and a
    # 32 x 4 x 4 block here has about 80: the blocks overflow it, and the
    # workgroup then stages and sums each column block in chunks. The sums
    # keep candidate order, so they equal the vector-ALU forms' (with their
    # own chunking) to the bit, in one launch over three column blocks.
    n, box = 3000, 30.0
    g, o, _, _ = _pair(sdt, orc, n, box, 0.9, 5.0, 48, 6, False, 83, sdt.FP32)
    g.set_launch_width(64)
    g.set_tuning(spread_stage_kb=16)
    f = np.random.default_rng(84).standard_normal((3 * n, 40))
    out = {}
    for form in (0, 1, 2):
        g.set_forms(0, 0, form)
        out[form] = g.mul_vector(f)
    assert np.array_equal(out[0], out[1]) and np.array_equal(out[0], out[2])
    assert orc.frob_rel_err(o.mul_vector(f), out[0]) < 1e-5
    with pytest.raises(Exception):
        g.set_tuning(spread_stage_kb=8)


@pytest.mark.parametrize("prec", [0, 1])
def test_fused_and_separate_xy_transforms_agree(sdt, orc, prec):
    # the x / y transforms run as one fused plane kernel where a plane fits
    # LDS and as separate x and y passes otherwise; forcing the separate
    # passes on the same mesh checks one against the other (FP64: the same
    # radix chain, identical; FP32 K = 128: the fused kernel's two-stage
    # 16 x 8 transform against the generic chain, rounding only)
    n, box = 2000, 30.0
    g, o, _, _ = _pair(sdt, orc, n, box, 0.9, 5.0, 128, 6, False, 91, prec)
    f = np.random.default_rng(92).standard_normal((3 * n, 9))
    g.set_forms(0, 0, 0, 1)
    sep = g.recip_mul_vector(f)
    g.set_forms(0, 0, 0, 0)
    fused = g.recip_mul_vector(f)
    if prec == 0:
        assert np.array_equal(sep, fused)
    else:
        assert orc.frob_rel_err(fused, sep) < 1e-6
    assert orc.frob_rel_err(o.mul_vector(f, mode="recip"), sep) < (1e-10 if prec == 0 else 1e-5)
