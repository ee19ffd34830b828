"""stokesdt_amd: MI355X (gfx950) Brownian-dynamics step of StokesDT.

Python mirror of the reference's class interfaces over the C ABI of
libstokesdt_amd.so (include/stokesdt_amd.h): MobSpme, PairListR/PairListS,
BrwnLanczos, StericForce, the rpy_spme_mex / pairlist_mex one-shot entry
points, and a BDSimBox-style step driver. All arithmetic runs on the GPU in
HIP kernels; if the library is missing every call raises.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libstokesdt_amd.so")

FP64 = 0
FP32 = 1

SDT_OK = 0
SDT_EINVAL = -1
SDT_ESTATE = -2
SDT_ENOMEM = -3
SDT_ECONVERGE = -4
SDT_EHIP = -5
SDT_ECAP = -6

_lib = None

_I = ctypes.c_int
_D = ctypes.c_double
_P = ctypes.c_void_p
_I64 = ctypes.c_int64
_PI64 = ctypes.POINTER(ctypes.c_int64)
_PI = ctypes.POINTER(ctypes.c_int)

# C signatures of include/stokesdt_amd.h and include/stokesdt_amd_debug.h
SIGNATURES = {
    "sdt_last_error": (ctypes.c_char_p, []),
    "sdt_version": (ctypes.c_char_p, []),
    "sdt_mob_spme_create": (_I, [_I, _P, _D, _D, _D, _I, _I, _I, ctypes.POINTER(_P)]),
    "sdt_mob_spme_init": (_I, [_P]),
    "sdt_mob_spme_update": (_I, [_P, _P, _P]),
    "sdt_mob_spme_update_dev": (_I, [_P, _P, _P]),
    "sdt_mob_spme_mulvector": (_I, [_P, _I, _D, _I, _P, _D, _I, _P]),
    "sdt_mob_spme_mulvector_real": (_I, [_P, _I, _D, _I, _P, _D, _I, _P]),
    "sdt_mob_spme_mulvector_recip": (_I, [_P, _I, _D, _I, _P, _D, _I, _P]),
    "sdt_mob_spme_mulvector_dev": (_I, [_P, _I, _I, _D, _I, _P, _D, _I, _P]),
    "sdt_mob_spme_dim": (_I, [_P, _PI]),
    "sdt_mob_spme_real_nnz": (_I, [_P, _PI64]),
    "sdt_mob_spme_destroy": (_I, [_P]),
    "sdt_rpy_spme": (_I, [_P, _P, _I, _D, _P, _I, _I, ctypes.c_char_p, _D, _D, _I, _I, _P, _I]),
    "sdt_rpy_spme_ex": (_I, [_P, _P, _I, _D, _P, _I, _I, ctypes.c_char_p, _D, _D, _I, _I, _I, _P,
                             _I]),
    "sdt_pairlist_create": (_I, [_I, _P, _D, _D, ctypes.c_char, ctypes.POINTER(_P)]),
    "sdt_pairlist_init": (_I, [_P, _PI64]),
    "sdt_pairlist_build": (_I, [_P, _P, _PI64]),
    "sdt_pairlist_get_pairs": (_I, [_P, _P, _P]),
    "sdt_pairlist_destroy": (_I, [_P]),
    "sdt_pairlist": (_I, [_P, _P, _I, _D, _D, ctypes.c_char, _P, _P, _I64, _PI64]),
    "sdt_brwn_lanczos_create": (_I, [_I, _I, _I, _D, ctypes.POINTER(_P)]),
    "sdt_brwn_lanczos_init": (_I, [_P]),
    "sdt_brwn_lanczos_compute": (_I, [_P, _P, _I, _I, _P, _I, _P]),
    "sdt_brwn_lanczos_compute_dev": (_I, [_P, _P, _I, _I, _P, _I, _P]),
    "sdt_brwn_lanczos_iters": (_I, [_P, _P, _I, _PI]),
    "sdt_brwn_lanczos_set_reorth": (_I, [_P, _I]),
    "sdt_brwn_lanczos_destroy": (_I, [_P]),
    "sdt_steric_create": (_I, [_I, _P, _D, _D, _D, ctypes.POINTER(_P)]),
    "sdt_steric_init": (_I, [_P]),
    "sdt_steric_compute": (_I, [_P, _P, _P, _P]),
    "sdt_steric_accumulate": (_I, [_P, _P, _P, _P]),
    "sdt_steric_destroy": (_I, [_P]),
    "sdt_bd_create": (_I, [_I, _P, _P, _D, _D, _D, _I, _I, _I, _I, _I, _D, _D, _D, _D, _I,
                           ctypes.c_uint64, ctypes.POINTER(_P)]),
    "sdt_bd_init": (_I, [_P]),
    "sdt_bd_advance": (_I, [_P, _I]),
    "sdt_bd_get_positions": (_I, [_P, _P]),
    "sdt_bd_last_iters": (_I, [_P, _P, _I, _PI]),
    "sdt_bd_destroy": (_I, [_P]),
    "sdt_force_steric": (_I, [_P, _P, _I, _D, _D, _D, _P]),
    "sdt_rpy_ewald_matrix": (_I, [_P, _P, _I, _D, _D, ctypes.c_char_p, _D, _P, _I]),
    "sdt_rpy_ewald_rows": (_I, [_P, _P, _I, _D, _D, _D, _I, _P, _P, _P, _P]),
    "sdt_diffusion": (_I, [_P, _I, _I, _D, _P, ctypes.POINTER(ctypes.c_int)]),
    "sdt_diffusion_dev_noalloc": (_I, [_P, _I, _I, _D, _P, ctypes.POINTER(ctypes.c_int), _P]),
    "sdt_brwn_random_create": (_I, [_I, _I, _I, ctypes.c_uint64, ctypes.POINTER(_P)]),
    "sdt_brwn_random_init": (_I, [_P]),
    "sdt_brwn_random_compute": (_I, [_P, _P, _I, _I, _P, _I, _P]),
    "sdt_brwn_random_compute_dev": (_I, [_P, _P, _I, _I, _P, _I, _P]),
    "sdt_brwn_random_destroy": (_I, [_P]),
    "sdt_simbox_create": (_I, [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p,
                               ctypes.POINTER(_P)]),
    "sdt_simbox_init": (_I, [_P]),
    "sdt_simbox_advance": (_I, [_P, _I]),
    "sdt_simbox_npos": (_I, [_P, _PI]),
    "sdt_simbox_state": (_I, [_P, ctypes.POINTER(_I64), ctypes.POINTER(_D)]),
    "sdt_simbox_params": (_I, [_P, _P, _P]),
    "sdt_simbox_get_positions": (_I, [_P, _P]),
    "sdt_simbox_get_radii": (_I, [_P, _P]),
    "sdt_simbox_write_xyz": (_I, [_P, ctypes.c_char_p]),
    "sdt_simbox_destroy": (_I, [_P]),
    "sdt_gaussian": (_I, [ctypes.c_uint64, ctypes.c_uint32, _I64, _P]),
    "sdt_device_synchronize": (_I, []),
    "sdt_debug_gram": (_I, [_I64, _I, _P, _P, _P]),
    "sdt_debug_gram_groups": (_I, [_I64, _I, _I, _P, _P, _P]),
    "sdt_debug_band_sqrt": (_I, [_I, _I, _P, _I, _P]),
    "sdt_debug_spme_spread": (_I, [_P, _P, _P]),
    "sdt_debug_spme_interp": (_I, [_P, _P, _D, _P]),
    "sdt_debug_spme_forms": (_I, [_P, _I, _I, _I, _I]),
    "sdt_debug_mob_launch_width": (_I, [_P, _I]),
    "sdt_debug_spme_tuning": (_I, [_P, _I, _I, _I, _I]),
    "sdt_bench_stage_times": (_I, [_P, _I, _I, _P]),
    "sdt_bench_copy": (_I, [_I64, _I, ctypes.POINTER(_D)]),
    "sdt_bench_bd_steps": (_I, [_P, _I, ctypes.POINTER(_D)]),
    "sdt_bd_set_profiling": (_I, [_P, _I]),
    "sdt_bd_profile": (_I, [_P, _P, _P, _I]),
    "sdt_bd_profile_columns": (_I, [_P, _P, _I]),
    "sdt_util_rsa_positions": (_I, [_I, _D, _D, ctypes.c_uint64, _I64, _P, _PI64]),
    "sdt_mob_ewald_create": (_I, [_I, _P, _D, _D, _D, ctypes.POINTER(_P)]),
    "sdt_mob_ewald_get_matrix": (_I, [_P, _I, _P]),
    "sdt_mob_init": (_I, [_P]),
    "sdt_mob_update": (_I, [_P, _P, _P]),
    "sdt_mob_mulvector": (_I, [_P, _I, _D, _I, _P, _D, _I, _P]),
    "sdt_mob_dim": (_I, [_P, _PI]),
    "sdt_mob_destroy": (_I, [_P]),
    "sdt_brwn_chol_create": (_I, [_I, ctypes.POINTER(_P)]),
    "sdt_brwn_chol_init": (_I, [_P]),
    "sdt_brwn_chol_compute": (_I, [_P, _P, _I, _I, _P, _I, _P]),
    "sdt_brwn_chol_compute_dev": (_I, [_P, _P, _I, _I, _P, _I, _P]),
    "sdt_brwn_chol_destroy": (_I, [_P]),
    "sdt_bd_set_graphs": (_I, [_P, _I]),
    "sdt_bd_autotune": (_I, [_P, ctypes.c_char_p, _P, _P]),
    "sdt_ensemble_create": (_I, [_I, _P, _P, _P, _P, _P, _P, _P, _I, _I, _I, _I, _D, _D, _D, _D,
                                 _I, ctypes.c_uint64, ctypes.POINTER(_P)]),
    "sdt_ensemble_init": (_I, [_P, _I, ctypes.c_char_p]),
    "sdt_ensemble_set_graphs": (_I, [_P, _I]),
    "sdt_ensemble_advance": (_I, [_P, _I, _P]),
    "sdt_ensemble_get_positions": (_I, [_P, _I, _P]),
    "sdt_ensemble_tuning": (_I, [_P, _I, _P, _P]),
    "sdt_ensemble_last_iters": (_I, [_P, _I, _P, _I, _PI]),
    "sdt_ensemble_destroy": (_I, [_P]),
    "sdt_serve_create": (_I, [_I, _D, _I, _I, _I, _I, _D, _D, _D, _D, _I, ctypes.POINTER(_P)]),
    "sdt_serve_footprint": (_I, [_P, _I, _I, _D, _D, ctypes.POINTER(_D)]),
    "sdt_serve_submit": (_I, [_P, _I, _P, _P, _D, _D, _D, _I, _I, ctypes.c_uint64, _D, _PI]),
    "sdt_serve_run": (_I, [_P, _D]),
    "sdt_serve_job": (_I, [_P, _I, _PI, _P, _PI, ctypes.POINTER(_D)]),
    "sdt_serve_job_positions": (_I, [_P, _I, _P]),
    "sdt_serve_job_step_ms": (_I, [_P, _I, _P, _I, _PI]),
    "sdt_serve_stats": (_I, [_P, _P]),
    "sdt_serve_destroy": (_I, [_P]),
    "sdt_device_memory": (_I, [ctypes.POINTER(_D), ctypes.POINTER(_D)]),
    "sdt_plain_mob_create": (_I, [_I, _P, _D, _D, _D, _I, _I, ctypes.POINTER(_P)]),
    "sdt_plain_mob_update": (_I, [_P, _P]),
    "sdt_plain_mob_mulvector": (_I, [_P, _I, _I, _D, _I, _P, _D, _I, _P]),
    "sdt_plain_lanczos": (_I, [_P, _I, _I, _D, _I, _I, _P, _I, _P, _P, _I, _PI]),
    "sdt_plain_mob_destroy": (_I, [_P]),
    "sdt_plain_bd_create": (_I, [_I, _P, _P, _D, _D, _D, _I, _I, _I, _I, _D, _D, _D, _D, _I,
                                 ctypes.c_uint64, ctypes.POINTER(_P)]),
    "sdt_plain_bd_init": (_I, [_P]),
    "sdt_plain_bd_advance": (_I, [_P, _I]),
    "sdt_plain_bd_get_positions": (_I, [_P, _P]),
    "sdt_plain_bd_last_iters": (_I, [_P, _P, _I, _PI]),
    "sdt_plain_bd_set_profiling": (_I, [_P, _I]),
    "sdt_plain_bd_profile": (_I, [_P, _P, _P, _P, _I]),
    "sdt_plain_bd_destroy": (_I, [_P]),
}

STAGES = ["spread", "fft_xy_fwd", "fft_z_influence", "fft_xy_inv", "interp", "real_space",
          "krylov_block", "steric", "mob_update", "input_copy"]


class SdtError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("stokesdt_amd error %d: %s" % (code, msg))
        self.code = code


def lib():
    """Loads libstokesdt_amd.so; raises if it was not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libstokesdt_amd.so is not built: run __graft_entry__.build() "
                               "or python stokesdt_amd/build.py")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _check(rc, allow=()):
    if rc != SDT_OK and rc not in allow:
        raise SdtError(rc, lib().sdt_last_error().decode())
    return rc


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _colmajor(a):
    a = np.asarray(a, dtype=np.float64)
    if a.ndim == 1:
        a = a.reshape(-1, 1)
    return np.asfortranarray(a)


def version():
    return lib().sdt_version().decode()


class MobSpme:
    """MobSpme (src/mob/mob_spme.h:21-125) on the GPU.

    MobSpme(npos, rdi, box_size, xi, rmax, dim, porder, precision=FP64);
    init() -> bool; update(pos, rdi); mul_vector / real_mul_vector /
    recip_mul_vector(f, alpha, beta, v) with f, v of shape (3 npos, nrhs).
    """

    def __init__(self, npos, rdi, box_size, xi, rmax, dim, porder, precision=FP64):
        self.npos = int(npos)
        self.K = int(dim)
        self.rdi = _f64(rdi)
        h = _P()
        _check(lib().sdt_mob_spme_create(self.npos, _ptr(self.rdi), box_size, xi, rmax, dim,
                                         porder, precision, ctypes.byref(h)))
        self.h = h

    def init(self):
        rc = lib().sdt_mob_spme_init(self.h)
        if rc != SDT_OK:
            self.init_error = lib().sdt_last_error().decode()
            return False
        return True

    def update(self, pos, rdi=None):
        pos = _f64(pos)
        r = self.rdi if rdi is None else _f64(rdi)
        _check(lib().sdt_mob_spme_update(self.h, _ptr(pos), _ptr(r)))

    def dim(self):
        d = ctypes.c_int(0)
        _check(lib().sdt_mob_spme_dim(self.h, ctypes.byref(d)))
        return d.value

    def real_nnz(self):
        n = ctypes.c_int64(0)
        _check(lib().sdt_mob_spme_real_nnz(self.h, ctypes.byref(n)))
        return n.value

    def _mul(self, fn, f, alpha, beta, v):
        f = _colmajor(f)
        n, nrhs = f.shape
        out = np.zeros((n, nrhs), order="F") if v is None else _colmajor(v).copy(order="F")
        _check(fn(self.h, nrhs, alpha, n, _ptr(f), beta, n, _ptr(out)))
        return out

    def mul_vector(self, f, alpha=1.0, beta=0.0, v=None):
        return self._mul(lib().sdt_mob_spme_mulvector, f, alpha, beta, v)

    def real_mul_vector(self, f, alpha=1.0, beta=0.0, v=None):
        return self._mul(lib().sdt_mob_spme_mulvector_real, f, alpha, beta, v)

    def recip_mul_vector(self, f, alpha=1.0, beta=0.0, v=None):
        return self._mul(lib().sdt_mob_spme_mulvector_recip, f, alpha, beta, v)

    def debug_spread(self, f):
        """Spread of one 3N force vector; returns meshes [3][K][K][K]."""
        f = _f64(f)
        K = self.K
        mesh = np.zeros(3 * K * K * K)
        _check(lib().sdt_debug_spme_spread(self.h, _ptr(f), _ptr(mesh)))
        return mesh.reshape(3, K, K, K)

    def debug_interp(self, meshes, alpha=1.0):
        """Interpolation of three meshes [3][K][K][K] at the particles
        (monodisperse); returns 3N values in the reference layout."""
        m = _f64(np.asarray(meshes).reshape(-1))
        v = np.zeros(3 * self.npos)
        _check(lib().sdt_debug_spme_interp(self.h, _ptr(m), alpha, _ptr(v)))
        return v

    def set_forms(self, interp=0, z=0, spread=0, fft=0):
        """Launch forms (testing and measurement): interp meshes per LDS pass
        (0 built-in, 1, 2 or 4 (FP32)); z transform shape (0 built-in,
        1 .. 5); spread form (0 built-in, 1 forces in LDS, 2 scalar force
        reads, 3 matrix-core sums; plus 4 for row-tiled mesh stores, plus 8
        for column blocks in turn); x / y transforms (0 fused planes where
        they fit, 1 separate passes)."""
        _check(lib().sdt_debug_spme_forms(self.h, int(interp), int(z), int(spread), int(fft)))

    def set_interp_form(self, form):
        self.set_forms(interp=form)

    def set_launch_width(self, width):
        """Right-hand sides per launch of mul_vector (default 16)."""
        _check(lib().sdt_debug_mob_launch_width(self.h, int(width)))

    def set_tuning(self, fft_order=0, interp_sxt=0, interp_ch=0, spread_stage_kb=0):
        """Launch choices for every width (0 = built-in): FFT radix order,
        interp strip width and particle chunk, spread stage size in KB."""
        _check(lib().sdt_debug_spme_tuning(self.h, int(fft_order), int(interp_sxt),
                                           int(interp_ch), int(spread_stage_kb)))

    def stage_times(self, b, reps=5):
        t = np.zeros(7)
        _check(lib().sdt_bench_stage_times(self.h, b, reps, _ptr(t)))
        return t

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            lib().sdt_mob_spme_destroy(self.h)
            self.h = None


class MobEwald:
    """MobEwald(npos, rdi, box_size, tol, xi=None) (src/mob/mob_ewald.h): the
    dense Ewald-summed RPY matrix on the GPU. init() -> bool; update(pos,
    rdi); mul_vector(f, alpha, beta, v); get_matrix(). Usable wherever a
    mobility is (BrwnLanczos, BrwnRandom, BrwnChol)."""

    def __init__(self, npos, rdi, box_size, tol, xi=None):
        self.npos = int(npos)
        self.rdi = _f64(rdi)
        h = _P()
        _check(lib().sdt_mob_ewald_create(self.npos, _ptr(self.rdi), box_size, tol,
                                          -1.0 if xi is None else xi, ctypes.byref(h)))
        self.h = h

    def init(self):
        rc = lib().sdt_mob_init(self.h)
        if rc != SDT_OK:
            self.init_error = lib().sdt_last_error().decode()
            return False
        return True

    def update(self, pos, rdi=None):
        pos = _f64(pos)
        r = self.rdi if rdi is None else _f64(rdi)
        _check(lib().sdt_mob_update(self.h, _ptr(pos), _ptr(r)))

    def dim(self):
        d = ctypes.c_int(0)
        _check(lib().sdt_mob_dim(self.h, ctypes.byref(d)))
        return d.value

    def mul_vector(self, f, alpha=1.0, beta=0.0, v=None):
        f = _colmajor(f)
        n, nrhs = f.shape
        out = np.zeros((n, nrhs), order="F") if v is None else _colmajor(v).copy(order="F")
        _check(lib().sdt_mob_mulvector(self.h, nrhs, alpha, n, _ptr(f), beta, n, _ptr(out)))
        return out

    def get_matrix(self):
        n = 3 * self.npos
        mat = np.zeros((n, n), order="F")
        _check(lib().sdt_mob_ewald_get_matrix(self.h, n, _ptr(mat)))
        return mat

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            lib().sdt_mob_destroy(self.h)
            self.h = None


class BrwnChol:
    """BrwnChol(dim) (src/brwn/brwn_chol.cc:29-83): y = L z, M = L L^T, on a
    dense Ewald mobility."""

    def __init__(self, dim):
        h = _P()
        _check(lib().sdt_brwn_chol_create(dim, ctypes.byref(h)))
        self.h = h

    def init(self):
        return lib().sdt_brwn_chol_init(self.h) == SDT_OK

    def compute(self, mob, z):
        z = _colmajor(z)
        n, nrhs = z.shape
        y = np.zeros((n, nrhs), order="F")
        _check(lib().sdt_brwn_chol_compute(self.h, mob.h, nrhs, n, _ptr(z), n, _ptr(y)))
        return y

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            lib().sdt_brwn_chol_destroy(self.h)
            self.h = None


def rpy_spme(pos, rdi, L, f, mode, xi, rmax, dim, porder, precision=FP64):
    """v = rpy_spme_mex(pos, rdi, L, f, mode, xi, rmax, dim, porder)
    (mex/rpy_spme_mex.cc:8-75). pos: (3, npos) or flat AoS; f: (3 npos, nrhs)."""
    pos = _f64(np.asarray(pos).T.reshape(-1) if np.ndim(pos) == 2 and np.shape(pos)[0] == 3
               else np.asarray(pos).reshape(-1))
    rdi = _f64(rdi).reshape(-1)
    f = _colmajor(f)
    n, nrhs = f.shape
    v = np.zeros((n, nrhs), order="F")
    _check(lib().sdt_rpy_spme_ex(_ptr(pos), _ptr(rdi), rdi.size, L, _ptr(f), nrhs, n,
                                 mode.encode(), xi, rmax, dim, porder, precision, _ptr(v), n))
    return v


def rpy_ewald_matrix(pos, rdi, L, tol, mode="full", xi=None):
    """mat = rpy_ewald_matrix_mex(pos, rdi, L, tol, mode, xi)
    (mex/rpy_ewald_matrix_mex.cc:10-64). pos: (3, npos) or flat AoS; xi None
    selects the reference's default splitting. Returns the 3 npos x 3 npos
    matrix."""
    pos = _f64(np.asarray(pos).T.reshape(-1) if np.ndim(pos) == 2 and np.shape(pos)[0] == 3
               else np.asarray(pos).reshape(-1))
    rdi = _f64(rdi).reshape(-1)
    n = 3 * rdi.size
    mat = np.zeros((n, n), order="F")
    _check(lib().sdt_rpy_ewald_matrix(_ptr(pos), _ptr(rdi), rdi.size, L, tol, mode.encode(),
                                      -1.0 if xi is None else xi, _ptr(mat), n))
    return mat


def rpy_ewald_rows(pos, rdi, L, tol, f, rows, xi=None):
    """EwaldVectorKernel (src/mob/ewald.cc:333-536), v = M f, for the
    particles in rows only, with no size limit. pos: flat AoS or (npos, 3);
    f: 3 npos values. Returns (v as (len(rows), 3), xi used)."""
    pos = _f64(pos).reshape(-1)
    rdi = _f64(rdi).reshape(-1)
    f = _f64(f).reshape(-1)
    rows = np.ascontiguousarray(rows, dtype=np.int32).reshape(-1)
    if f.size != 3 * rdi.size or pos.size != 3 * rdi.size:
        raise ValueError("pos and f need 3 values per particle")
    v = np.zeros(3 * rows.size)
    x = ctypes.c_double(0.0)
    _check(lib().sdt_rpy_ewald_rows(_ptr(pos), _ptr(rdi), rdi.size, L, tol,
                                    -1.0 if xi is None else xi, rows.size,
                                    rows.ctypes.data_as(_P), _ptr(f), _ptr(v), ctypes.byref(x)))
    return v.reshape(-1, 3), x.value


def diffusion(pos, dt):
    """Diffusion coefficients of tools/calc_diff (calc_diff.cc:100-140).
    pos: (nframes, npos, 3) positions, dt: time between frames. Returns D for
    lags 1 .. nframes/2 - 1."""
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    nf, n = pos.shape[0], pos.shape[1]
    D = np.zeros(max(nf // 2 - 1, 1))
    nl = ctypes.c_int(0)
    _check(lib().sdt_diffusion(_ptr(pos), nf, n, dt, _ptr(D), ctypes.byref(nl)))
    return D[:nl.value].copy()


class PairList:
    """PairListR ('R') / PairListS ('S') (src/pairlist/pair_base.h:30-69)."""

    def __init__(self, npos, rdi, box_size, cutoff, mode):
        self.npos = int(npos)
        self.rdi = _f64(rdi)
        h = _P()
        _check(lib().sdt_pairlist_create(self.npos, _ptr(self.rdi), box_size, cutoff,
                                         mode.encode()[:1], ctypes.byref(h)))
        self.h = h

    def init(self):
        est = ctypes.c_int64(0)
        rc = lib().sdt_pairlist_init(self.h, ctypes.byref(est))
        return est.value if rc == SDT_OK else 0

    def build(self, pos):
        pos = _f64(pos)
        n = ctypes.c_int64(0)
        _check(lib().sdt_pairlist_build(self.h, _ptr(pos), ctypes.byref(n)))
        self.nnz = n.value
        return n.value

    def get_pairs(self):
        ia = np.zeros(self.npos + 1, dtype=np.int32)
        ja = np.zeros(max(self.nnz, 1), dtype=np.int32)
        _check(lib().sdt_pairlist_get_pairs(self.h, _ptr(ia), _ptr(ja)))
        return ia, ja[:self.nnz]

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            lib().sdt_pairlist_destroy(self.h)
            self.h = None


def PairListR(npos, rdi, box_size, cutoff):
    return PairList(npos, rdi, box_size, cutoff, "R")


def PairListS(npos, rdi, box_size, cutoff):
    return PairList(npos, rdi, box_size, cutoff, "S")


def pairlist(pos, rdi, L, cutoff, mode):
    """[ia, ja] = pairlist_mex(pos, rdi, L, cutoff, mode) (mex/pairlist_mex.cc:9-80)."""
    pos = _f64(np.asarray(pos).T.reshape(-1) if np.ndim(pos) == 2 and np.shape(pos)[0] == 3
               else np.asarray(pos).reshape(-1))
    rdi = _f64(rdi).reshape(-1)
    n = rdi.size
    ia = np.zeros(n + 1, dtype=np.int32)
    nnz = ctypes.c_int64(0)
    cap = 64 * n + 64
    while True:
        ja = np.zeros(cap, dtype=np.int32)
        rc = lib().sdt_pairlist(_ptr(pos), _ptr(rdi), n, L, cutoff, mode.encode()[:1], _ptr(ia),
                                _ptr(ja), cap, ctypes.byref(nnz))
        if rc == SDT_ECAP:
            cap = nnz.value
            continue
        _check(rc)
        return ia, ja[:nnz.value].copy()


class BrwnRandom:
    """BrwnRandom(dim, num_vecs, max_nrhs) (src/brwn/brwn_random.h:16-86);
    the sample vectors come from the Gaussian stream of seed."""

    def __init__(self, dim, num_vecs, max_nrhs, seed=5678):
        h = _P()
        _check(lib().sdt_brwn_random_create(dim, num_vecs, max_nrhs, seed, ctypes.byref(h)))
        self.h = h
        self.dim = dim

    def init(self):
        return lib().sdt_brwn_random_init(self.h) == SDT_OK

    def compute(self, mob, z):
        """y = sampled approximation of M^{1/2} z; z: (dim, nrhs). mob None
        reuses the last sampling."""
        z = _colmajor(z)
        n, nrhs = z.shape
        y = np.zeros((n, nrhs), order="F")
        _check(lib().sdt_brwn_random_compute(self.h, mob.h if mob is not None else None, nrhs, n,
                                             _ptr(z), n, _ptr(y)))
        return y

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            lib().sdt_brwn_random_destroy(self.h)
            self.h = None


class BrwnLanczos:
    """BrwnLanczos(dim, max_iters, max_nrhs, tol) (src/brwn/brwn_lanczos.h:45-46)."""

    def __init__(self, dim, max_iters, max_nrhs, tol):
        h = _P()
        _check(lib().sdt_brwn_lanczos_create(dim, max_iters, max_nrhs, tol, ctypes.byref(h)))
        self.h = h
        self.dim = dim

    def init(self):
        return lib().sdt_brwn_lanczos_init(self.h) == SDT_OK

    def set_reorth(self, on):
        _check(lib().sdt_brwn_lanczos_set_reorth(self.h, 1 if on else 0))

    def compute(self, mob, z):
        """y ~= M^{1/2} z; z: (dim, nrhs). Returns (y, converged)."""
        z = _colmajor(z)
        n, nrhs = z.shape
        y = np.zeros((n, nrhs), order="F")
        rc = lib().sdt_brwn_lanczos_compute(self.h, mob.h, nrhs, n, _ptr(z), n, _ptr(y))
        _check(rc, allow=(SDT_ECONVERGE,))
        return y, rc == SDT_OK

    def iters(self):
        out = np.zeros(4096, dtype=np.int32)
        ng = ctypes.c_int(0)
        _check(lib().sdt_brwn_lanczos_iters(self.h, _ptr(out), 4096, ctypes.byref(ng)))
        return out[:ng.value].copy()

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            lib().sdt_brwn_lanczos_destroy(self.h)
            self.h = None


class StericForce:
    """StericForce(npos, rdi, box_size, r0, k0) (src/force/force_steric.cc:56-112)."""

    def __init__(self, npos, rdi, box_size, r0, k0):
        self.npos = int(npos)
        self.rdi = _f64(rdi)
        h = _P()
        _check(lib().sdt_steric_create(self.npos, _ptr(self.rdi), box_size, r0, k0,
                                       ctypes.byref(h)))
        self.h = h

    def init(self):
        return lib().sdt_steric_init(self.h) == SDT_OK

    def compute(self, pos):
        pos = _f64(pos)
        f = np.zeros(3 * self.npos)
        _check(lib().sdt_steric_compute(self.h, _ptr(pos), _ptr(self.rdi), _ptr(f)))
        return f

    def accumulate(self, pos, f):
        pos = _f64(pos)
        f = _f64(f).copy()
        _check(lib().sdt_steric_accumulate(self.h, _ptr(pos), _ptr(self.rdi), _ptr(f)))
        return f

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            lib().sdt_steric_destroy(self.h)
            self.h = None


def force_steric(pos, rdi, L, r0, k0):
    """f = force_steric_mex(pos, rdi, L, r0, k0) (mex/force_steric_mex.cc:8-50).
    pos: (3, npos) or flat AoS; returns the flat AoS force vector."""
    pos = _f64(np.asarray(pos).T.reshape(-1) if np.ndim(pos) == 2 and np.shape(pos)[0] == 3
               else np.asarray(pos).reshape(-1))
    rdi = _f64(rdi).reshape(-1)
    f = np.zeros(pos.size)
    _check(lib().sdt_force_steric(_ptr(pos), _ptr(rdi), rdi.size, L, r0, k0, _ptr(f)))
    return f


class SimBox:
    """BDSimBox(config_file, model_file, xyz_file) (src/simbox/simbox_bd.cc:31-366)
    with the reference's .cfg / .mod / XYZ formats, on the GPU BD step.
    Construction parses the files only; init() allocates the device state."""

    DKEYS = ("spme_xi", "spme_rmax", "lanczos_tol", "steric_r0", "steric_k0", "delta_t",
             "ewald_tol")
    IKEYS = ("spme_dim", "spme_porder", "mob_interval", "lanczos_maxiters", "lanczos_maxnrhs",
             "traj_interval", "traj_start", "precision", "seed", "mob_method", "brownian_method")

    def __init__(self, config_file, model_file, xyz_file):
        h = _P()
        _check(lib().sdt_simbox_create(str(config_file).encode(), str(model_file).encode(),
                                       str(xyz_file).encode(), ctypes.byref(h)))
        self.h = h
        n = ctypes.c_int(0)
        _check(lib().sdt_simbox_npos(self.h, ctypes.byref(n)))
        self.npos = n.value

    def init(self):
        _check(lib().sdt_simbox_init(self.h))

    def advance(self, nsteps):
        _check(lib().sdt_simbox_advance(self.h, int(nsteps)))

    def state(self):
        st, t = ctypes.c_int64(0), ctypes.c_double(0.0)
        _check(lib().sdt_simbox_state(self.h, ctypes.byref(st), ctypes.byref(t)))
        return st.value, t.value

    def params(self):
        d = np.zeros(7)
        i = np.zeros(11, dtype=np.int64)
        _check(lib().sdt_simbox_params(self.h, _ptr(d), _ptr(i)))
        out = dict(zip(self.DKEYS, d.tolist()))
        out.update(zip(self.IKEYS, i.tolist()))
        out["seed"] = int(np.uint64(i[8]))
        return out

    def positions(self):
        out = np.zeros(3 * self.npos)
        _check(lib().sdt_simbox_get_positions(self.h, _ptr(out)))
        return out

    def radii(self):
        out = np.zeros(self.npos)
        _check(lib().sdt_simbox_get_radii(self.h, _ptr(out)))
        return out

    def write_xyz(self, path):
        _check(lib().sdt_simbox_write_xyz(self.h, str(path).encode()))

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            lib().sdt_simbox_destroy(self.h)
            self.h = None


class BDSim:
    """BDSimBox::Advance (src/simbox/simbox_bd.cc:296-366) on the GPU."""

    def __init__(self, pos, rdi, box_size, xi, rmax, dim, porder, precision, max_iters, max_nrhs,
                 tol, steric_r0, steric_k0, dt, mob_interval, seed):
        pos = _f64(pos)
        self.rdi = _f64(rdi)
        self.npos = self.rdi.size
        h = _P()
        _check(lib().sdt_bd_create(self.npos, _ptr(pos), _ptr(self.rdi), box_size, xi, rmax, dim,
                                   porder, precision, max_iters, max_nrhs, tol, steric_r0,
                                   steric_k0, dt, mob_interval, seed, ctypes.byref(h)))
        self.h = h

    def init(self):
        _check(lib().sdt_bd_init(self.h))
        return True

    def advance(self, nsteps):
        _check(lib().sdt_bd_advance(self.h, nsteps))

    def timed_advance(self, nsteps):
        ms = ctypes.c_double(0.0)
        _check(lib().sdt_bench_bd_steps(self.h, nsteps, ctypes.byref(ms)))
        return ms.value

    def positions(self):
        out = np.zeros(3 * self.npos)
        _check(lib().sdt_bd_get_positions(self.h, _ptr(out)))
        return out

    def last_iters(self):
        out = np.zeros(4096, dtype=np.int32)
        ng = ctypes.c_int(0)
        _check(lib().sdt_bd_last_iters(self.h, _ptr(out), 4096, ctypes.byref(ng)))
        return out[:ng.value].copy()

    def set_profiling(self, on):
        _check(lib().sdt_bd_set_profiling(self.h, 1 if on else 0))

    def set_graphs(self, on):
        """Step tails as one replayed HIP graph per mobility interval (default)."""
        _check(lib().sdt_bd_set_graphs(self.h, 1 if on else 0))

    def autotune(self, cache=None):
        """Autotuned SPME launch choices for single-column and Lanczos-batch
        launches: ((narrow), (wide), [default ms, tuned ms] x 2)."""
        c = np.zeros(8, dtype=np.int32)
        ms = np.zeros(4)
        _check(lib().sdt_bd_autotune(self.h, None if cache is None else str(cache).encode(),
                                     _ptr(c), _ptr(ms)))
        return tuple(int(x) for x in c[:4]), tuple(int(x) for x in c[4:]), ms

    def profile(self):
        """{stage: (total ms, launches, right-hand sides over the launches)}
        recorded since profiling was enabled."""
        ms = np.zeros(len(STAGES))
        cnt = np.zeros(len(STAGES), dtype=np.int64)
        cols = np.zeros(len(STAGES), dtype=np.int64)
        _check(lib().sdt_bd_profile(self.h, _ptr(ms), _ptr(cnt), len(STAGES)))
        _check(lib().sdt_bd_profile_columns(self.h, _ptr(cols), len(STAGES)))
        return {name: (float(ms[i]), int(cnt[i]), int(cols[i])) for i, name in enumerate(STAGES)}

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            lib().sdt_bd_destroy(self.h)
            self.h = None


def gaussian(seed, stream, n):
    out = np.zeros(n)
    _check(lib().sdt_gaussian(seed, stream, n, _ptr(out)))
    return out


def debug_gram(A, B):
    A = _f64(A)
    B = _f64(B)
    n, b = A.shape
    G = np.zeros((b, b))
    _check(lib().sdt_debug_gram(n, b, _ptr(A), _ptr(B), _ptr(G)))
    return G


def bench_copy(n, reps=3):
    """Mean time (ms) of a plain device copy of n doubles."""
    ms = _D()
    _check(lib().sdt_bench_copy(int(n), int(reps), ctypes.byref(ms)))
    return ms.value


def debug_gram_groups(A, B, bs):
    """[A_g^T B_g] for the column groups of width bs of n x btot A, B."""
    A = _f64(A)
    B = _f64(B)
    n, btot = A.shape
    ng = (btot + bs - 1) // bs
    G = np.zeros((ng, bs, bs))
    _check(lib().sdt_debug_gram_groups(n, btot, bs, _ptr(A), _ptr(B), _ptr(G)))
    return G


def band_sqrt(T, kd, kb):
    """T^{1/2} E_1 for a symmetric band matrix (host routine of block Lanczos)."""
    T = _f64(T)
    m = T.shape[0]
    S = np.zeros((m, kb))
    _check(lib().sdt_debug_band_sqrt(m, kd, _ptr(T), kb, _ptr(S)))
    return S


def rsa_positions(n, box, a=1.0, seed=0, max_attempts=None):
    """Seeded random sequential addition of n non-overlapping spheres."""
    pos = np.zeros(3 * n)
    placed = ctypes.c_int64(0)
    if max_attempts is None:
        max_attempts = 2000 * n
    _check(lib().sdt_util_rsa_positions(n, box, a, seed, max_attempts, _ptr(pos),
                                        ctypes.byref(placed)))
    return pos


class PlainMobSpme:
    """The plain FP64 GPU form of MobSpme (per-RHS loop, stored P/ind tables,
    atomic spread, library FFTs, stored 3x3 blocks + SpMV): the baseline of
    the optimised path. create runs the reference's Init checks."""

    def __init__(self, npos, rdi, box_size, xi, rmax, dim, porder):
        self.npos = int(npos)
        self.rdi = _f64(rdi)
        h = _P()
        _check(lib().sdt_plain_mob_create(self.npos, _ptr(self.rdi), box_size, xi, rmax, dim,
                                          porder, ctypes.byref(h)))
        self.h = h

    def update(self, pos):
        pos = _f64(pos)
        _check(lib().sdt_plain_mob_update(self.h, _ptr(pos)))

    def mul_vector(self, f, alpha=1.0, beta=0.0, v=None, mode=0):
        f = _colmajor(f)
        n, nrhs = f.shape
        out = np.zeros((n, nrhs), order="F") if v is None else _colmajor(v).copy(order="F")
        _check(lib().sdt_plain_mob_mulvector(self.h, mode, nrhs, alpha, n, _ptr(f), beta, n,
                                             _ptr(out)))
        return out

    def lanczos(self, z, max_iters, max_nrhs, tol):
        """(y ~= M^{1/2} z, iteration of convergence per group) by the
        reference's BlockLanczos, group after group."""
        z = _colmajor(z)
        n, nrhs = z.shape
        y = np.zeros((n, nrhs), order="F")
        it = np.zeros(4096, dtype=np.int32)
        ng = ctypes.c_int(0)
        _check(lib().sdt_plain_lanczos(self.h, max_iters, max_nrhs, tol, nrhs, n, _ptr(z), n,
                                       _ptr(y), _ptr(it), 4096, ctypes.byref(ng)))
        return y, it[:ng.value].copy()

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            lib().sdt_plain_mob_destroy(self.h)
            self.h = None


class PlainBDSim:
    """BDSimBox::Advance (src/simbox/simbox_bd.cc:296-366) on the plain FP64
    GPU path."""

    def __init__(self, pos, rdi, box_size, xi, rmax, dim, porder, max_iters, max_nrhs, tol,
                 steric_r0, steric_k0, dt, mob_interval, seed):
        pos = _f64(pos)
        self.rdi = _f64(rdi)
        self.npos = self.rdi.size
        h = _P()
        _check(lib().sdt_plain_bd_create(self.npos, _ptr(pos), _ptr(self.rdi), box_size, xi, rmax,
                                         dim, porder, max_iters, max_nrhs, tol, steric_r0,
                                         steric_k0, dt, mob_interval, seed, ctypes.byref(h)))
        self.h = h

    def init(self):
        _check(lib().sdt_plain_bd_init(self.h))
        return True

    def advance(self, nsteps):
        _check(lib().sdt_plain_bd_advance(self.h, nsteps))

    def positions(self):
        out = np.zeros(3 * self.npos)
        _check(lib().sdt_plain_bd_get_positions(self.h, _ptr(out)))
        return out

    def last_iters(self):
        out = np.zeros(4096, dtype=np.int32)
        ng = ctypes.c_int(0)
        _check(lib().sdt_plain_bd_last_iters(self.h, _ptr(out), 4096, ctypes.byref(ng)))
        return out[:ng.value].copy()

    def set_profiling(self, on):
        _check(lib().sdt_plain_bd_set_profiling(self.h, 1 if on else 0))

    def profile(self):
        ms = np.zeros(len(STAGES))
        cnt = np.zeros(len(STAGES), dtype=np.int64)
        cols = np.zeros(len(STAGES), dtype=np.int64)
        _check(lib().sdt_plain_bd_profile(self.h, _ptr(ms), _ptr(cnt), _ptr(cols), len(STAGES)))
        return {name: (float(ms[i]), int(cnt[i]), int(cols[i])) for i, name in enumerate(STAGES)}

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            lib().sdt_plain_bd_destroy(self.h)
            self.h = None


class Ensemble:
    """Independent BD systems of uneven size, box and mesh advanced together:
    each mobility interval's updates and Lanczos solves side by side, every
    other step four HIP graph launches (lanes of systems) over all systems'
    step tails.
    systems: list of dicts with keys pos (flat AoS), rdi, box, xi, rmax, dim."""

    def __init__(self, systems, porder, precision, max_iters, max_nrhs, tol, steric_r0,
                 steric_k0, dt, mob_interval, seed):
        self.npos = np.array([len(s["rdi"]) for s in systems], dtype=np.int32)
        pos = _f64(np.concatenate([np.asarray(s["pos"], dtype=np.float64).reshape(-1)
                                   for s in systems]))
        rdi = _f64(np.concatenate([np.asarray(s["rdi"], dtype=np.float64) for s in systems]))
        box = _f64([s["box"] for s in systems])
        xi = _f64([s["xi"] for s in systems])
        rmax = _f64([s["rmax"] for s in systems])
        dim = np.array([s["dim"] for s in systems], dtype=np.int32)
        h = _P()
        _check(lib().sdt_ensemble_create(len(systems), _ptr(self.npos), _ptr(pos), _ptr(rdi),
                                         _ptr(box), _ptr(xi), _ptr(rmax), _ptr(dim), porder,
                                         precision, max_iters, max_nrhs, tol, steric_r0,
                                         steric_k0, dt, mob_interval, seed, ctypes.byref(h)))
        self.h = h

    def init(self, autotune=False, cache=None):
        _check(lib().sdt_ensemble_init(self.h, 1 if autotune else 0,
                                       None if cache is None else str(cache).encode()))

    def set_graphs(self, on):
        _check(lib().sdt_ensemble_set_graphs(self.h, 1 if on else 0))

    def advance(self, nsteps):
        """Returns each step's host wall time in ms."""
        ms = np.zeros(max(int(nsteps), 1))
        _check(lib().sdt_ensemble_advance(self.h, int(nsteps), _ptr(ms)))
        return ms[:nsteps]

    def positions(self, s):
        out = np.zeros(3 * int(self.npos[s]))
        _check(lib().sdt_ensemble_get_positions(self.h, s, _ptr(out)))
        return out

    def tuning(self, s):
        """((narrow choices), (wide choices), [default ms, tuned ms] x 2)."""
        c = np.zeros(8, dtype=np.int32)
        ms = np.zeros(4)
        _check(lib().sdt_ensemble_tuning(self.h, s, _ptr(c), _ptr(ms)))
        return tuple(c[:4]), tuple(c[4:]), ms

    def last_iters(self, s):
        out = np.zeros(4096, dtype=np.int32)
        ng = ctypes.c_int(0)
        _check(lib().sdt_ensemble_last_iters(self.h, s, _ptr(out), 4096, ctypes.byref(ng)))
        return out[:ng.value].copy()

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            lib().sdt_ensemble_destroy(self.h)
            self.h = None


SERVE_FP64, SERVE_FP32, SERVE_PLAIN = 0, 1, 2


def device_memory():
    """(free, total) bytes of the current device."""
    f, t = _D(), _D()
    _check(lib().sdt_device_memory(ctypes.byref(f), ctypes.byref(t)))
    return f.value, t.value
JOB_STATES = ("queued", "running", "done", "failed", "stopped")


class Server:
    """Request loop for BD jobs with continuous batching (sdt_serve_*):
    jobs are admitted first come first served while a slot is free and their
    predicted device footprint fits the capacity left; each runs on its own
    stream, resident on the device."""

    def __init__(self, kind, capacity_bytes, max_active, porder=6, max_iters=100, max_nrhs=8,
                 tol=1e-6, steric_r0=2.0, steric_k0=125.0, dt=1e-3, mob_interval=100):
        h = _P()
        _check(lib().sdt_serve_create(kind, float(capacity_bytes), max_active, porder, max_iters,
                                      max_nrhs, tol, steric_r0, steric_k0, dt, mob_interval,
                                      ctypes.byref(h)))
        self.h = h
        self.npos = []

    def footprint(self, npos, dim, box, rmax):
        b = _D()
        _check(lib().sdt_serve_footprint(self.h, npos, dim, box, rmax, ctypes.byref(b)))
        return b.value

    def submit(self, pos, rdi, box, xi, rmax, dim, nsteps, seed, arrival):
        pos = _f64(pos)
        rdi = _f64(rdi)
        j = ctypes.c_int(0)
        _check(lib().sdt_serve_submit(self.h, rdi.size, _ptr(pos), _ptr(rdi), box, xi, rmax, dim,
                                      nsteps, seed, arrival, ctypes.byref(j)))
        self.npos.append(rdi.size)
        return j.value

    def run(self, deadline=0.0):
        _check(lib().sdt_serve_run(self.h, float(deadline)))

    def job(self, j):
        st, nd = ctypes.c_int(0), ctypes.c_int(0)
        t = np.zeros(3)
        b = _D()
        _check(lib().sdt_serve_job(self.h, j, ctypes.byref(st), _ptr(t), ctypes.byref(nd),
                                   ctypes.byref(b)))
        out = dict(state=JOB_STATES[st.value], arrival=t[0], admit=t[1], finish=t[2],
                   steps=nd.value, bytes=b.value)
        if st.value == 3:
            out["error"] = lib().sdt_last_error().decode()
        return out

    def positions(self, j):
        out = np.zeros(3 * self.npos[j])
        _check(lib().sdt_serve_job_positions(self.h, j, _ptr(out)))
        return out

    def step_ms(self, j):
        n = ctypes.c_int(0)
        _check(lib().sdt_serve_job_step_ms(self.h, j, None, 0, ctypes.byref(n)))
        out = np.zeros(max(n.value, 1))
        _check(lib().sdt_serve_job_step_ms(self.h, j, _ptr(out), n.value, ctypes.byref(n)))
        return out[:n.value]

    def stats(self):
        s = np.zeros(3)
        _check(lib().sdt_serve_stats(self.h, _ptr(s)))
        return dict(wall_s=s[0], peak_bytes=s[1], peak_active=int(s[2]))

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            lib().sdt_serve_destroy(self.h)
            self.h = None
