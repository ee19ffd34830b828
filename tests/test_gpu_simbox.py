"""BD driver from files on the GPU: the same trajectory as the BD step API,
restart continuity, trajectory frames (BDSimBox::Advance,
src/simbox/simbox_bd.cc:296-366) and the command-line driver
(bin/main/main.cc:12-158)."""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

CFG = """mob-method spme
brownian-method lanczos
delta-t 0.001
mob-interval 4
spme-nfft-grids 32
spme-porder 6
spme-xi 0.9
spme-rmax 5.0
lanczos-maxiters 100
lanczos-maxnrhs 4
lanczos-tol 1.0e-8
seed 77
"""


def _system(sdt, tmp_path, n=300, box=18.0, extra=""):
    pos = sdt.rsa_positions(n, box, 1.0, 5)
    xyz = tmp_path / "start.xyz"
    with open(xyz, "w") as fh:
        fh.write("%d\n0 0.0 %.17g %.17g %.17g\n" % (n, box, box, box))
        for i in range(n):
            fh.write("X %.17g %.17g %.17g\n" % tuple(pos[3 * i:3 * i + 3]))
    cfg = tmp_path / "run.cfg"
    cfg.write_text(CFG + extra)
    mod = tmp_path / "run.mod"
    mod.write_text("par X 1.0 0.0\nsteric 2.0 125.0\n")
    return pos, cfg, mod, xyz


def test_simbox_matches_bd_api(sdt, tmp_path):
    pos, cfg, mod, xyz = _system(sdt, tmp_path)
    sb = sdt.SimBox(cfg, mod, xyz)
    sb.init()
    sb.advance(6)
    bd = sdt.BDSim(pos, np.ones(pos.size // 3), 18.0, 0.9, 5.0, 32, 6, sdt.FP64, 100, 4, 1e-8,
                   2.0, 125.0, 1e-3, 4, 77)
    bd.init()
    bd.advance(6)
    assert np.array_equal(sb.positions(), bd.positions())
    step, t = sb.state()
    assert step == 6 and abs(t - 0.006) < 1e-15


def test_restart_at_interval_boundary_is_exact(sdt, tmp_path):
    pos, cfg, mod, xyz = _system(sdt, tmp_path)
    ref = sdt.SimBox(cfg, mod, xyz)
    ref.init()
    ref.advance(8)
    a = sdt.SimBox(cfg, mod, xyz)
    a.init()
    a.advance(4)
    r = tmp_path / "restart.xyz"
    a.write_xyz(r)
    b = sdt.SimBox(cfg, mod, r)
    assert b.state()[0] == 4
    b.init()
    b.advance(4)
    assert b.state()[0] == 8
    assert np.array_equal(b.positions(), ref.positions())


def test_restart_inside_an_interval_continues(sdt, tmp_path):
    # a start at step 6 (mob-interval 4) recomputes the interval's
    # displacements from the restart positions: close to, not equal to, the
    # uninterrupted run
    pos, cfg, mod, xyz = _system(sdt, tmp_path)
    ref = sdt.SimBox(cfg, mod, xyz)
    ref.init()
    ref.advance(8)
    a = sdt.SimBox(cfg, mod, xyz)
    a.init()
    a.advance(6)
    r = tmp_path / "restart6.xyz"
    a.write_xyz(r)
    b = sdt.SimBox(cfg, mod, r)
    b.init()
    b.advance(2)
    d_ref = ref.positions() - a.positions()
    d = b.positions() - a.positions()
    assert np.linalg.norm(d - d_ref) < 0.1 * np.linalg.norm(d_ref)


def test_trajectory_frames(sdt, tmp_path):
    traj = tmp_path / "traj.xyz"
    pos, cfg, mod, xyz = _system(sdt, tmp_path, extra="traj-interval 2\ntraj-startstep 1\n"
                                 "traj-output %s\n" % traj)
    sb = sdt.SimBox(cfg, mod, xyz)
    sb.init()
    snaps = {}
    for k in range(1, 7):
        sb.advance(1)
        snaps[k] = sb.positions()
    del sb
    with open(traj) as fh:
        lines = fh.read().splitlines()
    n = int(lines[0])
    frames = [lines[i:i + n + 2] for i in range(0, len(lines), n + 2)]
    assert [int(f[1].split()[0]) for f in frames] == [2, 4, 6]
    for f in frames:
        step = int(f[1].split()[0])
        p = np.array([[float(v) for v in l.split()[1:4]] for l in f[2:]]).ravel()
        assert np.array_equal(p, snaps[step])


def test_command_line_driver(sdt, tmp_path):
    pos, cfg, mod, xyz = _system(sdt, tmp_path)
    out = tmp_path / "end.xyz"
    exe = os.path.join(ROOT, "stokesdt_amd", "stokesdt_bd")
    r = subprocess.run([exe, "--config", str(cfg), "--model", str(mod), "--xyz", str(xyz),
                        "-n", "5", "--restart", str(out)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    assert "BD simulation finished" in r.stdout
    sb = sdt.SimBox(cfg, mod, xyz)
    sb.init()
    sb.advance(5)
    with open(out) as fh:
        lines = fh.read().splitlines()
    assert int(lines[1].split()[0]) == 5
    p = np.array([[float(v) for v in l.split()[1:4]] for l in lines[2:]]).ravel()
    assert np.array_equal(p, sb.positions())


def test_simbox_random_method_matches_composed_oracle(sdt, orc, tmp_path):
    # brownian-method random (simbox_bd.cc:153-157): D = BrwnRandom(Z) on the
    # mobility at the interval start; without forces every step moves the
    # particles by sqrt(2 dt) D[:, step]
    from oracle import brwn_random as BR
    n, box, nv = 120, 14.0, 20
    pos = sdt.rsa_positions(n, box, 1.0, 6)
    xyz = tmp_path / "start.xyz"
    with open(xyz, "w") as fh:
        fh.write("%d\n0 0.0 %.17g %.17g %.17g\n" % (n, box, box, box))
        for i in range(n):
            fh.write("X %.17g %.17g %.17g\n" % tuple(pos[3 * i:3 * i + 3]))
    cfg = tmp_path / "run.cfg"
    cfg.write_text(CFG.replace("brownian-method lanczos", "brownian-method random") +
                   "random-numvecs %d\n" % nv)
    mod = tmp_path / "run.mod"
    mod.write_text("par X 1.0 0.0\n")
    sb = sdt.SimBox(cfg, mod, xyz)
    sb.init()
    sb.advance(4)
    g = sdt.MobSpme(n, np.ones(n), box, 0.9, 5.0, 32, 6, sdt.FP64)
    assert g.init()
    g.update(pos)
    M = g.mul_vector(np.eye(3 * n))
    dim = 3 * n
    Z = orc.gaussian(77, 0, 4 * dim).reshape(4, dim).T
    omega = orc.gaussian(77 ^ 0x5678, 0, dim * nv).reshape(nv, dim).T
    D = BR.apply(M, *BR.sample(M, omega), Z)
    expect = pos + np.sqrt(2e-3) * D.sum(axis=1)
    disp = sb.positions() - pos
    assert np.abs(disp - (expect - pos)).max() < 1e-9 * np.abs(expect - pos).max()



def test_reference_example_config_runs(sdt):
    # the reference's data/example.cfg + example.mod on its N1000_Phi0.1.xyz:
    # the dense Ewald mobility (tol 1e-4) with Cholesky displacements. Three
    # steps composed from the pieces: D = L Z for the interval's Gaussian
    # block (simbox_bd.cc:323-331), then per step f = steric force,
    # pos += dt M f + sqrt(2 dt) D[:, k] with M at the interval's positions
    ref = os.path.join(ROOT, "tests", "golden", "reference_data")
    sb = sdt.SimBox(os.path.join(ref, "example.cfg"), os.path.join(ref, "example.mod"),
                    os.path.join(ref, "N1000_Phi0.1.xyz"))
    sb.init()
    pos0 = sb.positions()
    rdi = sb.radii()
    p = sb.params()
    sb.advance(3)
    n, L = rdi.size, None
    with open(os.path.join(ref, "N1000_Phi0.1.xyz")) as fh:
        fh.readline()
        L = float(fh.readline().split()[2])
    me = sdt.MobEwald(n, rdi, L, p["ewald_tol"])
    assert me.init()
    me.update(pos0)
    mi = p["mob_interval"]
    Z = sdt.gaussian(p["seed"], 0, 3 * n * mi).reshape(mi, 3 * n).T.copy()
    ch = sdt.BrwnChol(3 * n)
    assert ch.init()
    D = ch.compute(me, Z)
    st = sdt.StericForce(n, rdi, L, 2.0, 125.0)
    assert st.init()
    pos = pos0.copy()
    dt = p["delta_t"]
    for k in range(3):
        f = st.compute(pos)
        disp = me.mul_vector(f, dt, np.sqrt(2.0 * dt), D[:, k:k + 1])[:, 0]
        pos = pos + disp
    got = sb.positions()
    assert np.linalg.norm(got - pos) <= 1e-12 * np.linalg.norm(pos - pos0)
