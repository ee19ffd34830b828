"""ctypes binding of the oracle (liboracle_sdt.so).

TEST INFRASTRUCTURE ONLY: a CPU FP64 C++/OpenMP restatement of the
xing-liu/stokesdt Brownian-dynamics path, used by tests/ and smoke() as the
checker and by bench.py as the CPU baseline. The product package
(stokesdt_amd) never imports this module.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle_sdt.so")
_lib = None

_D = ctypes.c_double
_I = ctypes.c_int
_P = ctypes.c_void_p
_DP = ctypes.POINTER(ctypes.c_double)
_IP = ctypes.POINTER(ctypes.c_int)


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError("oracle library not built: run `make -C oracle`")
        L = ctypes.CDLL(_LIB_PATH)
        sig = {
            "orc_bspline": (_D, [_I, _D]),
            "orc_philox": (None, [_P, _P, _P]),
            "orc_gaussian": (None, [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_int64, _P]),
            "orc_num_threads": (_I, []),
            "orc_set_num_threads": (None, [_I]),
            "orc_fft3d_forward": (None, [_I, _P]),
            "orc_fft3d_backward": (None, [_I, _P]),
            "orc_fft1d": (None, [_I, _P, _I]),
            "orc_sym_eig": (None, [_I, _P, _P, _P]),
            "orc_tridiag_eig": (None, [_I, _P, _P, _P, _P]),
            "orc_householder_qr": (None, [_I, _I, _P, _P]),
            "orc_cholesky": (_I, [_I, _P]),
            "orc_spme_create": (_P, [_I, _P, _D, _D, _I, _I]),
            "orc_spme_destroy": (None, [_P]),
            "orc_spme_update": (None, [_P, _P]),
            "orc_spme_ld3": (_I, [_P]),
            "orc_spme_spread": (None, [_P, _P, _P]),
            "orc_spme_interp": (None, [_P, _P, _D, _P]),
            "orc_spme_lm2": (None, [_P, _P]),
            "orc_spme_recip": (None, [_P, _I, _D, _I, _P, _D, _I, _P]),
            "orc_mob_spme_create": (_P, [_I, _P, _D, _D, _D, _I, _I]),
            "orc_mob_spme_init": (_I, [_P]),
            "orc_mob_spme_update": (None, [_P, _P, _P]),
            "orc_mob_spme_mulvector": (None, [_P, _I, _I, _D, _I, _P, _D, _I, _P]),
            "orc_mob_spme_real_nnz": (_I, [_P]),
            "orc_mob_spme_destroy": (None, [_P]),
            "orc_pairlist_create": (_P, [_I, _P, _D, _D, ctypes.c_char]),
            "orc_pairlist_init": (_I, [_P]),
            "orc_pairlist_build": (_I, [_P, _P]),
            "orc_pairlist_get": (None, [_P, _P, _P]),
            "orc_pairlist_nc1": (_I, [_P]),
            "orc_pairlist_destroy": (None, [_P]),
            "orc_steric_compute": (_I, [_I, _P, _P, _D, _D, _D, _P]),
            "orc_ewald_auto_xi": (_D, [_D]),
            "orc_ewald_matrix": (None, [_I, _P, _P, _D, _D, _D, _P, _P, _P, _P, _P]),
            "orc_lanczos": (_I, [_I, _P, _I, _I, _I, _D, _I, _I, _P, _I, _P, _P]),
            "orc_chol_disp": (_I, [_I, _P, _I, _I, _P, _I, _P]),
            "orc_bd_create": (_P, [_I, _P, _P, _D, _D, _D, _I, _I, _I, _I, _D, _D, _D,
                                   _D, _I, ctypes.c_uint64]),
            "orc_bd_advance": (None, [_P, _I]),
            "orc_bd_get_pos": (None, [_P, _P]),
            "orc_bd_last_iters": (_I, [_P, _P, _I]),
            "orc_bd_destroy": (None, [_P]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def bspline(p, x):
    return lib().orc_bspline(p, float(x))


def philox(ctr, key):
    c = np.asarray(ctr, dtype=np.uint32)
    k = np.asarray(key, dtype=np.uint32)
    out = np.zeros(4, dtype=np.uint32)
    lib().orc_philox(_ptr(c), _ptr(k), _ptr(out))
    return out


def gaussian(seed, stream, n):
    out = np.empty(n, dtype=np.float64)
    lib().orc_gaussian(seed, stream, n, _ptr(out))
    return out


def set_num_threads(n):
    lib().orc_set_num_threads(int(n))


def num_threads():
    return lib().orc_num_threads()


def fft1d(x, sign=-1):
    a = np.ascontiguousarray(x, dtype=np.complex128).copy()
    lib().orc_fft1d(a.size, _ptr(a), sign)
    return a


def rfftn(real):
    """3D R2C of a K^3 real array indexed [z][y][x]; returns [z][y][K/2+1]."""
    K = real.shape[0]
    ldx = 2 * (K // 2 + 1)
    mesh = np.zeros((K, K, ldx))
    mesh[:, :, :K] = real
    lib().orc_fft3d_forward(K, _ptr(mesh))
    return mesh[:, :, 0::2] + 1j * mesh[:, :, 1::2]


def irfftn(spec, K):
    ldx = 2 * (K // 2 + 1)
    mesh = np.zeros((K, K, ldx))
    mesh[:, :, 0::2] = spec.real
    mesh[:, :, 1::2] = spec.imag
    lib().orc_fft3d_backward(K, _ptr(mesh))
    return mesh[:, :, :K].copy()


def sym_eig(A):
    A = _f64(A)
    n = A.shape[0]
    w = np.zeros(n)
    V = np.zeros((n, n))
    lib().orc_sym_eig(n, _ptr(A), _ptr(w), _ptr(V))
    return w, V


def tridiag_eig(d, e):
    d = _f64(d)
    e = _f64(e)
    n = d.size
    w = np.zeros(n)
    V = np.zeros((n, n))
    lib().orc_tridiag_eig(n, _ptr(d), _ptr(e), _ptr(w), _ptr(V))
    return w, V


def householder_qr(A):
    """A: (m, n) array; returns Q (m, n), R (n, n)."""
    m, n = A.shape
    buf = np.asfortranarray(A, dtype=np.float64).copy(order="F")
    R = np.zeros((n, n), order="F")
    lib().orc_householder_qr(m, n, buf.ctypes.data_as(ctypes.c_void_p),
                             R.ctypes.data_as(ctypes.c_void_p))
    return buf, R


def cholesky(A):
    L = _f64(A).copy()
    info = lib().orc_cholesky(L.shape[0], _ptr(L))
    return L, info


class SpmeEngine:
    """Reference SPME engine pieces (mesh layout [z][y][2(K/2+1)])."""

    def __init__(self, npos, rdi, box, xi, dim, porder):
        self.rdi = None if rdi is None else _f64(rdi)
        self.h = lib().orc_spme_create(npos, _ptr(self.rdi), box, xi, dim, porder)
        if not self.h:
            raise ValueError("oracle SPME engine rejected the parameters")
        self.npos, self.K = npos, dim
        self.ld3 = lib().orc_spme_ld3(self.h)

    def update(self, pos):
        self.pos = _f64(pos)
        lib().orc_spme_update(self.h, _ptr(self.pos))

    def spread(self, f):
        f = _f64(f)
        out = np.zeros(3 * self.ld3)
        lib().orc_spme_spread(self.h, _ptr(f), _ptr(out))
        K = self.K
        return out.reshape(3, K, K, 2 * (K // 2 + 1))[:, :, :, :K].copy()

    def interp(self, meshes, alpha=1.0):
        K = self.K
        full = np.zeros((3, K, K, 2 * (K // 2 + 1)))
        full[:, :, :, :K] = meshes
        v = np.zeros(3 * self.npos)
        lib().orc_spme_interp(self.h, _ptr(full), alpha, _ptr(v))
        return v

    def lm2(self):
        K = self.K
        out = np.zeros(K * K * (K // 2 + 1))
        lib().orc_spme_lm2(self.h, _ptr(out))
        return out.reshape(K, K, K // 2 + 1)

    def recip(self, f, alpha=1.0, beta=0.0, v=None):
        f = np.asfortranarray(f, dtype=np.float64)
        if f.ndim == 1:
            f = f.reshape(-1, 1, order="F")
        n, nrhs = f.shape
        out = np.zeros((n, nrhs), order="F") if v is None else np.asfortranarray(v).copy(order="F")
        lib().orc_spme_recip(self.h, nrhs, alpha, n, f.ctypes.data_as(ctypes.c_void_p), beta, n,
                             out.ctypes.data_as(ctypes.c_void_p))
        return out

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_spme_destroy(self.h)
            self.h = None


class MobSpme:
    """MobSpme restated (src/mob/mob_spme.cc)."""

    MODES = {"full": 0, "real": 1, "recip": 2}

    def __init__(self, npos, rdi, box, xi, rmax, dim, porder):
        self.rdi = _f64(rdi)
        self.npos = npos
        self.h = lib().orc_mob_spme_create(npos, _ptr(self.rdi), box, xi, rmax, dim, porder)

    def init(self):
        return bool(lib().orc_mob_spme_init(self.h))

    def update(self, pos, rdi=None):
        self.pos = _f64(pos)
        r = self.rdi if rdi is None else _f64(rdi)
        lib().orc_mob_spme_update(self.h, _ptr(self.pos), _ptr(r))

    def mul_vector(self, f, alpha=1.0, beta=0.0, v=None, mode="full"):
        f = np.asfortranarray(f, dtype=np.float64)
        if f.ndim == 1:
            f = f.reshape(-1, 1, order="F")
        n, nrhs = f.shape
        out = np.zeros((n, nrhs), order="F") if v is None else \
            np.asfortranarray(v, dtype=np.float64).reshape(n, nrhs, order="F").copy(order="F")
        lib().orc_mob_spme_mulvector(self.h, self.MODES[mode], nrhs, alpha, n,
                                     f.ctypes.data_as(ctypes.c_void_p), beta, n,
                                     out.ctypes.data_as(ctypes.c_void_p))
        return out

    def real_nnz(self):
        return lib().orc_mob_spme_real_nnz(self.h)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_mob_spme_destroy(self.h)
            self.h = None


def pairlist(pos, rdi, box, cutoff, mode):
    """Returns (ia, ja) int32 CSR (0-based, self included), like pairlist_mex."""
    pos = _f64(pos)
    rdi = _f64(rdi)
    n = rdi.size
    h = lib().orc_pairlist_create(n, _ptr(rdi), box, cutoff, mode.encode()[:1])
    try:
        if lib().orc_pairlist_init(h) <= 0:
            raise ValueError("pair list init failed")
        nnz = lib().orc_pairlist_build(h, _ptr(pos))
        ia = np.zeros(n + 1, dtype=np.int32)
        ja = np.zeros(nnz, dtype=np.int32)
        lib().orc_pairlist_get(h, _ptr(ia), _ptr(ja))
        return ia, ja
    finally:
        lib().orc_pairlist_destroy(h)


def steric(pos, rdi, box, r0, k0):
    pos = _f64(pos)
    rdi = _f64(rdi)
    f = np.zeros(3 * rdi.size)
    nnz = lib().orc_steric_compute(rdi.size, _ptr(pos), _ptr(rdi), box, r0, k0, _ptr(f))
    if nnz < 0:
        raise ValueError("steric init failed")
    return f


def ewald_auto_xi(box):
    return lib().orc_ewald_auto_xi(box)


def ewald_matrix(pos, rdi, box, xi=None, tol=1e-12, parts=False):
    pos = _f64(pos)
    rdi = _f64(rdi)
    n = rdi.size
    if xi is None:
        xi = ewald_auto_xi(box)
    mat = np.zeros((3 * n, 3 * n))
    re = np.zeros((3 * n, 3 * n)) if parts else None
    rc = np.zeros((3 * n, 3 * n)) if parts else None
    nr = ctypes.c_int(0)
    nk = ctypes.c_int(0)
    lib().orc_ewald_matrix(n, _ptr(pos), _ptr(rdi), box, xi, tol, _ptr(mat), _ptr(re), _ptr(rc),
                           ctypes.byref(nr), ctypes.byref(nk))
    if parts:
        return mat, re, rc
    return mat


def lanczos(op, z, max_iters=100, max_nrhs=16, tol=1e-7):
    """op: dense symmetric matrix (ndarray) or MobSpme. z: (dim, nrhs)."""
    z = np.asfortranarray(z, dtype=np.float64)
    if z.ndim == 1:
        z = z.reshape(-1, 1, order="F")
    dim, nrhs = z.shape
    y = np.zeros((dim, nrhs), order="F")
    iters = np.zeros(nrhs + 1, dtype=np.int32)
    if isinstance(op, np.ndarray):
        mat = _f64(op)
        ng = lib().orc_lanczos(0, _ptr(mat), dim, max_iters, max_nrhs, tol, nrhs, dim,
                               z.ctypes.data_as(ctypes.c_void_p), dim,
                               y.ctypes.data_as(ctypes.c_void_p), _ptr(iters))
    else:
        ng = lib().orc_lanczos(1, op.h, dim, max_iters, max_nrhs, tol, nrhs, dim,
                               z.ctypes.data_as(ctypes.c_void_p), dim,
                               y.ctypes.data_as(ctypes.c_void_p), _ptr(iters))
    if ng < 0:
        raise ValueError("lanczos init failed")
    return y, iters[:ng].copy()


def chol_displacements(mat, z):
    mat = _f64(mat)
    z = np.asfortranarray(z, dtype=np.float64)
    n, nrhs = z.shape
    y = np.zeros((n, nrhs), order="F")
    info = lib().orc_chol_disp(n, _ptr(mat), nrhs, n, z.ctypes.data_as(ctypes.c_void_p), n,
                               y.ctypes.data_as(ctypes.c_void_p))
    return y, info


class BDSim:
    """BDSimBox::Advance restated with SPME + block Lanczos + steric."""

    def __init__(self, pos, rdi, box, xi, rmax, dim, porder, max_iters, max_nrhs, tol,
                 r0, k0, dt, mob_interval, seed):
        pos = _f64(pos)
        self.rdi = _f64(rdi)
        self.npos = self.rdi.size
        self.h = lib().orc_bd_create(self.npos, _ptr(pos), _ptr(self.rdi), box, xi, rmax, dim,
                                     porder, max_iters, max_nrhs, tol, r0, k0, dt,
                                     mob_interval, seed)
        if not self.h:
            raise ValueError("oracle BD init failed")

    def advance(self, nsteps):
        lib().orc_bd_advance(self.h, nsteps)

    def positions(self):
        out = np.zeros(3 * self.npos)
        lib().orc_bd_get_pos(self.h, _ptr(out))
        return out

    def last_iters(self):
        out = np.zeros(1024, dtype=np.int32)
        n = lib().orc_bd_last_iters(self.h, _ptr(out), 1024)
        return out[:n].copy()

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_bd_destroy(self.h)
            self.h = None


def frob_rel_err(a, b):
    """||a - b||_F / ||a||_F (src/util/matrix_io.cc:91-117)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm(a - b) / np.linalg.norm(a))
