"""Test configuration: the `gpu` marker and shared system generators."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU (runs the HIP library)")
    config.addinivalue_line("markers", "slow: CPU test that takes more than a few seconds")


@pytest.fixture(scope="session")
def orc():
    from oracle import oracle as O
    O.lib()
    return O


@pytest.fixture(scope="session")
def sdt():
    import stokesdt_amd as S
    S.lib()
    return S


REF_DATA = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "reference_data")


def reference_system(xyz, mod):
    """One of the reference's own test inputs (tests/data/*.xyz with
    mono.mod / poly.mod, committed under tests/golden/reference_data):
    returns (flat AoS positions, radii, box). The model's `par name radius`
    lines give the radii; types without a line have radius 1 (the XYZ reader
    and model rules of src/util/molecule_io.cc:39-231)."""
    par = {}
    with open(os.path.join(REF_DATA, mod)) as fh:
        for line in fh:
            t = line.split()
            if len(t) >= 3 and t[0] == "par":
                par[t[1]] = float(t[2])
    with open(os.path.join(REF_DATA, xyz)) as fh:
        lines = fh.read().splitlines()
    n = int(lines[0])
    box = float(lines[1].split()[2])
    pos, rdi = [], []
    for line in lines[2:2 + n]:
        t = line.split()
        rdi.append(par.get(t[0], 1.0))
        pos.extend(float(v) for v in t[1:4])
    return np.array(pos), np.array(rdi), box


def uniform_system(n, box, seed):
    """Uniform positions in [0, box)^3 (overlaps allowed, like the
    reference's tests/data/N1000_Phi0.1.xyz), unit radii."""
    rng = np.random.default_rng(seed)
    return rng.uniform(0.0, box, 3 * n), np.ones(n)


def poly_rsa_system(n, box, seed):
    """Radii from the 15 sizes of poly_system, largest first, placed by
    random sequential addition without overlaps (periodic): a system like the
    reference's tests/data/N412.xyz, whose SPME test needs every overlapping
    pair inside the real-space cutoff. Returns (flat AoS positions, radii)."""
    rng = np.random.default_rng(seed)
    sizes = np.geomspace(21.4, 115.2, 15)
    # one sphere of every size, the rest drawn with weights ~ r^-2: volume
    # fraction ~0.13 at 412 spheres in a box of 1000 (uniform weights give
    # ~0.57, near the jamming limit of sequential addition)
    w = sizes ** -2.0
    rdi = np.concatenate([sizes, rng.choice(sizes, size=n - sizes.size, p=w / w.sum())])
    rdi = np.sort(rdi)[::-1].copy()
    pos = np.zeros((n, 3))
    for i in range(n):
        while True:
            cand = rng.uniform(0.0, box, (64, 3))
            d = cand[:, None, :] - pos[None, :i, :]
            d -= box * np.round(d / box)
            ok = np.all(np.sqrt((d * d).sum(2)) > rdi[i] + rdi[None, :i], axis=1)
            if ok.any():
                pos[i] = cand[np.argmax(ok)]
                break
    return pos.reshape(-1), rdi


def poly_system(n, box, seed):
    """Polydisperse radii drawn from 15 sizes spanning 21..115 (the spread of
    tests/data/poly.mod), uniform positions."""
    rng = np.random.default_rng(seed)
    sizes = np.geomspace(21.4, 115.2, 15)
    rdi = rng.choice(sizes, size=n)
    return rng.uniform(0.0, box, 3 * n), rdi


def direct_steric(pos, rdi, box, cutoff, k0):
    """O(N^2) steric force of tests/steric/teststeric.cc:34-79: pairs with
    2r/(a_i + a_j) < cutoff (r = 0 counts as 1e-10) push apart with
    k0 ((a_i + a_j) cutoff / 2 - r) along the min-image separation."""
    n = rdi.size
    P = pos.reshape(n, 3)
    f = np.zeros((n, 3))
    for i in range(n):
        d = np.remainder(P[i] - P + box / 2.0, box) - box / 2.0
        r = np.sqrt((d ** 2).sum(1))
        r = np.where(r == 0.0, 1e-10, r)
        s = 2.0 * r / (rdi[i] + rdi)
        sel = (s - cutoff < 0.0)
        sel[i] = False
        force = k0 * ((rdi[i] + rdi) * cutoff / 2.0 - r)
        f[i] = ((force / r)[sel, None] * d[sel]).sum(0)
    return f.reshape(-1)


def box_for(n, phi, a=1.0):
    """Side of the periodic cube holding n spheres of radius a at volume
    fraction phi."""
    return (4.0 * np.pi * n * a ** 3 / (3.0 * phi)) ** (1.0 / 3.0)
