"""Pins the calc_diff restatement (oracle/calc_diff.py, after
tools/calc_diff/calc_diff.cc) by a closed form, a hand count and the
reference reader's conventions."""
import numpy as np

from oracle import calc_diff as CD


def test_uniform_drift_closed_form():
    # r_k(t) = r_k(0) + v_k t: every origin sees |v_k|^2 (i dt)^2, so
    # D(i) = <|v|^2> i dt / 6 exactly
    rng = np.random.default_rng(1)
    nf, n, dt = 12, 7, 0.25
    r0 = rng.uniform(0, 10, (n, 3))
    v = rng.standard_normal((n, 3))
    pos = r0[None] + v[None] * (dt * np.arange(nf))[:, None, None]
    D = CD.diffusion(pos, dt)
    lags = np.arange(1, nf // 2)
    assert D.shape == (nf // 2 - 1,)
    assert np.allclose(D, (v ** 2).sum(1).mean() * lags * dt / 6.0, rtol=1e-12, atol=0)


def test_hand_counted_two_particles():
    # 4 frames -> origins 0, 1 and the single lag 1; particle 0 moves 1 along
    # x per frame, particle 1 sits still then jumps 2 along z at frame 2
    pos = np.zeros((4, 2, 3))
    pos[:, 0, 0] = [0, 1, 2, 3]
    pos[:, 1, 2] = [0, 0, 2, 2]
    # sum over origins 0, 1: (1 + 0) + (1 + 4) = 6; D = 6 / (6 * 2 * 2 * 1 * dt)
    assert np.allclose(CD.diffusion(pos, 0.5), [6.0 / (6 * 2 * 2 * 1 * 0.5)])


def test_reader_conventions(tmp_path):
    # the reference writer's frame layout (molecule_io.cc:395-401) with blank
    # lines between frames, which the reader skips
    pos = np.arange(5 * 3 * 3, dtype=float).reshape(5, 3, 3) / 7.0
    with open(tmp_path / "t.xyz", "w") as fh:
        for f in range(5):
            fh.write("3\n%d %g 10 10 10\n" % (f, 0.2 * f + 1.0))
            for k in range(3):
                fh.write("N %.10g %.10g %.10g\n" % tuple(pos[f, k]))
            fh.write("\n  \n")
    p, tslen = CD.read_trajectory(tmp_path / "t.xyz")
    assert p.shape == (5, 3, 3)
    assert abs(tslen - 0.2) < 1e-12
    assert np.allclose(p, pos, rtol=1e-9)
    out = CD.report(p, tslen)
    assert out[0] == "nsteps = 5, step-length = 2.000e-01"
    assert out[3] == "tau, r = 1.000" and len(out) == 4 + 1
