"""The C-ABI library loads and exports every function declared in
include/*.h; host-side logic runs without a GPU."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INCLUDE = os.path.join(ROOT, "include")


def _declared():
    names = set()
    for fn in os.listdir(INCLUDE):
        if fn.endswith(".h"):
            text = open(os.path.join(INCLUDE, fn)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            names.update(re.findall(r"\b(sdt_[a-z0-9_]+)\s*\(", text))
    return names


def test_library_exports_every_declared_symbol(sdt):
    lib = ctypes.CDLL(sdt.LIB_PATH)
    declared = _declared()
    assert len(declared) >= 40
    missing = [n for n in declared if not hasattr(lib, n)]
    assert not missing, missing
    # and the Python binding covers every declared function
    assert declared <= set(sdt.SIGNATURES), declared - set(sdt.SIGNATURES)


def test_version(sdt):
    assert "gfx950" in sdt.version()


def test_band_sqrt_host_routine(sdt):
    rng = np.random.default_rng(3)
    # widths 4, 8 and 16 take fixed-width loops, the others the general one;
    # (440, 8, 8) is the size of a converging 8-column Lanczos group
    for m, kd, kb in [(8, 2, 2), (64, 8, 8), (200, 8, 8), (50, 1, 1), (33, 5, 3), (16, 16, 16),
                      (120, 4, 4), (90, 6, 4), (440, 8, 8), (70, 3, 5)]:
        A = rng.standard_normal((m, m))
        B = A @ A.T / m + 2 * np.eye(m)
        mask = np.abs(np.subtract.outer(np.arange(m), np.arange(m))) <= kd
        T = np.where(mask, B, 0.0)
        w, V = np.linalg.eigh(T)
        ref = (V * np.sqrt(w)) @ V.T[:, :kb]
        S = sdt.band_sqrt(T, kd, kb)
        assert np.abs(S - ref).max() <= 1e-13 * np.abs(ref).max()
    # a band matrix that splits: equal diagonal blocks with no coupling give
    # every eigenvalue twice and zero off-diagonals in the middle (the QL
    # sweeps deflate there); T^{1/2} then acts on each block alone
    for m2, kd, kb in [(40, 4, 4), (60, 8, 8), (25, 3, 2)]:
        A = rng.standard_normal((m2, m2))
        B = A @ A.T / m2 + np.eye(m2)
        mask = np.abs(np.subtract.outer(np.arange(m2), np.arange(m2))) <= kd
        B = np.where(mask, B, 0.0)
        T = np.zeros((2 * m2, 2 * m2))
        T[:m2, :m2] = B
        T[m2:, m2:] = B
        w, V = np.linalg.eigh(B)
        ref = np.zeros((2 * m2, kb))
        ref[:m2] = ((V * np.sqrt(w)) @ V.T)[:, :kb]
        S = sdt.band_sqrt(T, kd, kb)
        assert np.abs(S - ref).max() <= 1e-13 * np.abs(ref).max()


def test_rpy_spme_argument_errors(sdt):
    # the MEX argument checks (mex/rpy_spme_mex.cc:37-51) run before any
    # device call
    pos = np.zeros(6)
    rdi = np.ones(2)
    f = np.zeros((6, 1))
    with pytest.raises(sdt.SdtError, match="rmax must be less than L/2"):
        sdt.rpy_spme(pos, rdi, 10.0, f, "full", 0.5, 5.0, 32, 6)
    with pytest.raises(sdt.SdtError, match="porder must be greater than 1"):
        sdt.rpy_spme(pos, rdi, 10.0, f, "full", 0.5, 2.0, 32, 1)
    with pytest.raises(sdt.SdtError, match="Scalar arguments must be positive"):
        sdt.rpy_spme(pos, rdi, -1.0, f, "full", 0.5, 2.0, 32, 6)
    with pytest.raises(sdt.SdtError, match="mode must be full, real, or recip"):
        sdt.rpy_spme(pos, rdi, 10.0, f, "both", 0.5, 2.0, 32, 6)


def test_pairlist_argument_errors(sdt):
    pos = np.zeros(6)
    rdi = np.ones(2)
    with pytest.raises(sdt.SdtError, match="cutoff must be less than L/2"):
        sdt.pairlist(pos, rdi, 10.0, 5.0, "R")
    with pytest.raises(sdt.SdtError, match="mode must be R or S"):
        sdt.pairlist(pos, rdi, 10.0, 2.0, "Q")
