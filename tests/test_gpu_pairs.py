"""GPU pair lists (PairListR/PairListS, src/pairlist/pair_r.cc:95-170,
pair_s.cc:99-186; pairlist_mex, mex/pairlist_mex.cc:9-80) must be
bit-identical to the oracle's, and the steric force
(src/force/force_steric.cc:14-112) must match it."""
import ctypes
import os

import numpy as np
import pytest

from conftest import poly_system, uniform_system

pytestmark = pytest.mark.gpu
G = np.load(os.path.join(os.path.dirname(__file__), "golden", "golden_small.npz"))


@pytest.mark.parametrize("n,box,cutoff,mode,poly", [
    (500, 20.0, 5.0, "R", False),
    (500, 20.0, 2.7, "R", False),
    (3000, 40.0, 5.0, "R", True),
    (500, 20.0, 2.0, "S", False),
    (3000, 40.0, 2.3, "S", True),
    (20000, 120.0, 5.0, "R", False),
    (412, 1000.0, 2.3, "S", "prot"),
])
def test_pairlist_bit_exact(sdt, orc, n, box, cutoff, mode, poly):
    print("case", n, box, cutoff, mode, poly, flush=True)
    if poly == "prot":
        pos, rdi = poly_system(n, box, 41)
    else:
        pos, rdi = uniform_system(n, box, n)
        if poly:
            rdi = np.random.default_rng(n).uniform(0.5, 1.5, n)
    ia, ja = sdt.pairlist(pos, rdi, box, cutoff, mode)
    oia, oja = orc.pairlist(pos, rdi, box, cutoff, mode)
    assert np.array_equal(ia, oia)
    assert np.array_equal(ja, oja)


def test_pairlist_golden(sdt):
    pos, rdi = G["pl_pos"], G["pl_rdi"]
    for mode, cut in zip("RS", G["pl_cutoffs"]):
        ia, ja = sdt.pairlist(pos, rdi, 20.0, cut, mode)
        assert np.array_equal(ia, G["pl_%s_ia" % mode])
        assert np.array_equal(ja, G["pl_%s_ja" % mode])


def test_pairlist_boundary_ties(sdt, orc):
    # a cubic lattice of spacing 2.5 in a box of 25: particles sit exactly on
    # cell boundaries and many pair distances equal the cutoff 5.0 exactly
    # (inclusive test for 'R'), so membership depends on the same FP64
    # operations as the reference (cell assignment, 'all' stencil entries)
    g = np.arange(10) * 2.5
    P = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    pos = P.reshape(-1)
    rdi = np.ones(P.shape[0])
    for cutoff, mode in ((5.0, "R"), (2.5, "R"), (2.5, "S"), (5.0 / 1.0, "S")):
        if 2 * cutoff >= 25.0:
            continue
        ia, ja = sdt.pairlist(pos, rdi, 25.0, cutoff, mode)
        oia, oja = orc.pairlist(pos, rdi, 25.0, cutoff, mode)
        assert np.array_equal(ia, oia) and np.array_equal(ja, oja), (cutoff, mode)
    # negative and out-of-box coordinates are wrapped the same way
    pos2 = pos + np.repeat(np.array([-25.0, 50.0, 0.0])[None, :], P.shape[0], 0).reshape(-1)
    ia, ja = sdt.pairlist(pos2, rdi, 25.0, 5.0, "R")
    oia, oja = orc.pairlist(pos2, rdi, 25.0, 5.0, "R")
    assert np.array_equal(ia, oia) and np.array_equal(ja, oja)


def test_isolated_particles_have_only_self(sdt):
    pos = np.array([1.0, 1.0, 1.0, 10.0, 10.0, 10.0, 19.0, 1.0, 10.0])
    ia, ja = sdt.pairlist(pos, np.ones(3), 30.0, 2.0, "R")
    assert list(ia) == [0, 1, 2, 3] and list(ja) == [0, 1, 2]


def test_pairlist_capacity_error(sdt):
    pos, rdi = uniform_system(200, 15.0, 3)
    ia = np.zeros(201, dtype=np.int32)
    ja = np.zeros(5, dtype=np.int32)
    nnz = ctypes.c_int64(0)
    P = ctypes.c_void_p
    rc = sdt.lib().sdt_pairlist(pos.ctypes.data_as(P), rdi.ctypes.data_as(P), 200, 15.0, 5.0,
                                b"R", ia.ctypes.data_as(P), ja.ctypes.data_as(P), 5,
                                ctypes.byref(nnz))
    assert rc == sdt.SDT_ECAP and nnz.value > 5


def test_pairlist_handle_interface(sdt, orc):
    pos, rdi = uniform_system(800, 25.0, 9)
    pl = sdt.PairListR(800, rdi, 25.0, 4.0)
    assert pl.init() > 0
    pl.build(pos)
    ia, ja = pl.get_pairs()
    oia, oja = orc.pairlist(pos, rdi, 25.0, 4.0, "R")
    assert np.array_equal(ia, oia) and np.array_equal(ja, oja)
    # rebuild after a move keeps working
    pos2 = pos + 0.3
    pl.build(pos2)
    ia, ja = pl.get_pairs()
    oia, oja = orc.pairlist(pos2, rdi, 25.0, 4.0, "R")
    assert np.array_equal(ia, oia) and np.array_equal(ja, oja)


def test_steric_matches_oracle_and_golden(sdt, orc):
    pos, rdi = poly_system(412, 1000.0, 19)
    fo = sdt.StericForce(412, rdi, 1000.0, 2.3, 125.0)
    assert fo.init()
    f = fo.compute(pos)
    ref = orc.steric(pos, rdi, 1000.0, 2.3, 125.0)
    assert np.linalg.norm(ref) > 0
    assert np.linalg.norm(f - ref) / np.linalg.norm(ref) < 1e-12
    assert np.abs(f.reshape(-1, 3).sum(0)).max() < 1e-9 * np.abs(f).max()
    # Accumulate adds to the input
    base = np.full(3 * 412, 0.5)
    assert np.abs(fo.accumulate(pos, base) - (f + 0.5)).max() < 1e-12 * np.abs(f).max()
    fg = sdt.StericForce(300, G["pl_rdi"], 20.0, 2.2, 125.0)
    assert fg.init()
    fgold = fg.compute(G["pl_pos"])
    assert np.abs(fgold - G["steric_f"]).max() < 1e-12 * np.abs(G["steric_f"]).max()


def test_steric_rejects_nonpositive_k0(sdt):
    fo = sdt.StericForce(10, np.ones(10), 20.0, 2.0, 0.0)
    assert not fo.init()


def test_force_steric_one_shot(sdt, orc):
    # force_steric_mex semantics: (3, npos) positions, same forces as the
    # class interface
    pos, rdi = poly_system(412, 1000.0, 23)
    f = sdt.force_steric(pos.reshape(-1, 3).T, rdi, 1000.0, 2.3, 125.0)
    ref = orc.steric(pos, rdi, 1000.0, 2.3, 125.0)
    assert np.linalg.norm(f - ref) / np.linalg.norm(ref) < 1e-12
    with pytest.raises(sdt.SdtError):
        sdt.force_steric(pos, rdi, 1000.0, 2.3, 0.0)


@pytest.mark.parametrize("n,phi,r0,ncells", [(20_000, 0.3, 2.4, "many"), (60, 0.05, 6.0, "fallback")])
def test_steric_contact_cells_match_oracle(sdt, orc, n, phi, r0, ncells):
    # a dense random-sequential-addition system with contacts inside r0
    # (cells of the contact distance, 27 neighbours each), and a box under
    # three contact distances wide, where the force falls back to the pair list
    box = (4.0 * np.pi * n / (3.0 * phi)) ** (1.0 / 3.0)
    assert (box / r0 >= 3) == (ncells == "many")
    pos = sdt.rsa_positions(n, box, 1.0, 5)
    pos = pos + np.random.default_rng(6).uniform(-0.05, 0.05, pos.size)
    rdi = np.ones(n)
    fo = sdt.StericForce(n, rdi, box, r0, 125.0)
    assert fo.init()
    f = fo.compute(pos)
    ref = orc.steric(pos, rdi, box, r0, 125.0)
    assert np.linalg.norm(ref) > 0
    assert np.linalg.norm(f - ref) / np.linalg.norm(ref) < 1e-12
    # repeat calls give identical bits
    assert np.array_equal(fo.compute(pos), f)
