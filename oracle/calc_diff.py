"""Restatement of tools/calc_diff/calc_diff.cc (diffusion coefficients from an
XYZ trajectory) in Python/numpy.

TEST INFRASTRUCTURE ONLY: the checker of sdt_diffusion and of the calc_diff
command-line tool in tests/. The product package (stokesdt_amd) never imports
this module.

Pinned in tests/test_oracle_calc_diff.py by the closed form of a uniform
drift (D(i) = <|v|^2> i dt / 6), by a hand-counted two-particle case, and by
the parsing conventions of the reference's own reader.
"""
import numpy as np


def read_trajectory(path):
    """calc_diff.cc:11-98. Returns (pos[nsteps, np, 3], tslen). Lines holding
    only white space are skipped; every particle has radius 1."""
    with open(path) as fh:
        lines = [l for l in fh if l.split()]
    np_ = int(lines[0].split()[0])                       # :15-22
    assert np_ > 0
    per = np_ + 2
    assert len(lines) % per == 0                         # :44
    nsteps = len(lines) // per
    assert nsteps > 2                                    # :47
    ts0 = float(lines[1].split()[1])                     # :38-40
    tslen = float(lines[per + 1].split()[1]) - ts0       # :40-42
    pos = np.empty((nsteps, np_, 3))
    for f in range(nsteps):                              # :61-96
        for k in range(np_):
            t = lines[f * per + 2 + k].split()
            pos[f, k] = [float(t[1]), float(t[2]), float(t[3])]
    return pos, tslen


def diffusion(pos, tslen):
    """calc_diff.cc:100-140 for the single radius 1: D[i-1] for lags
    i = 1 .. nsteps/2 - 1, origins j < nsteps/2, in the reference's order
    (lag, origin, particle; x, y, z terms added one by one)."""
    nsteps, np_ = pos.shape[0], pos.shape[1]
    half = nsteps // 2
    out = []
    for i in range(1, half):
        acc = 0.0
        for j in range(half):
            d = pos[j + i] - pos[j]
            for k in range(np_):
                acc += d[k, 0] * d[k, 0]
                acc += d[k, 1] * d[k, 1]
                acc += d[k, 2] * d[k, 2]
        out.append(acc / (6.0 * np_ * half * i * tslen) * 1.0)
    return np.array(out)


def report(pos, tslen):
    """The lines calc_diff prints (:48, :99, :103-108, :134-140)."""
    D = diffusion(pos, tslen)
    lines = ["nsteps = %d, step-length = %.3e" % (pos.shape[0], tslen),
             "Have 1 different radii", "", "tau, r = 1.000"]
    for i, d in enumerate(D, start=1):
        lines.append("%.3f, %.6f, %.6f" % (i * tslen, d, d * i * tslen))
    return lines
