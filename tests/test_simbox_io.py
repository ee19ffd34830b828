"""BD driver file formats (MoleculeIO, src/util/molecule_io.cc:25-414) and
the BDSimBox configuration rules (src/simbox/simbox_bd.cc:62-289), on the
host: parsing, defaults and ranges, rejected inputs, restart frames. The
XYZ / model fixtures under tests/golden/reference_data are the reference's
own test inputs (tests/data, data/)."""
import os

import numpy as np
import pytest

REF = os.path.join(os.path.dirname(__file__), "golden", "reference_data")

BASE_CFG = """# test config
delta-t            0.002
mob-method         spme
brownian-method    lanczos
mob-interval       4
spme-nfft-grids    32
spme-porder        6
spme-xi            0.9
spme-rmax          5.0
lanczos-maxiters   100
lanczos-maxnrhs    4
lanczos-tol        1.0e-6
"""


def _write(tmp_path, name, text):
    p = tmp_path / name
    p.write_text(text)
    return p


def _read_xyz(path):
    with open(path) as fh:
        lines = fh.read().splitlines()
    n = int(lines[0])
    head = lines[1].split()
    types, pos = [], []
    for line in lines[2:2 + n]:
        t = line.split()
        types.append(t[0])
        pos.extend(float(v) for v in t[1:4])
    return n, head, types, np.array(pos)


def test_reference_fixture_poly(sdt, tmp_path):
    cfg = _write(tmp_path, "a.cfg", BASE_CFG)
    sb = sdt.SimBox(cfg, os.path.join(REF, "poly.mod"), os.path.join(REF, "N412.xyz"))
    n, head, types, pos = _read_xyz(os.path.join(REF, "N412.xyz"))
    assert sb.npos == n == 412
    assert np.array_equal(sb.positions(), pos)
    par = {}
    with open(os.path.join(REF, "poly.mod")) as fh:
        for line in fh:
            t = line.split()
            if len(t) >= 3 and t[0] == "par":
                par[t[1]] = float(t[2])
    assert np.array_equal(sb.radii(), np.array([par.get(t, 1.0) for t in types]))
    assert sb.state() == (0, 0.0)
    p = sb.params()
    assert p["delta_t"] == 0.002 and p["mob_interval"] == 4 and p["spme_dim"] == 32
    assert p["lanczos_maxnrhs"] == 4 and p["lanczos_tol"] == 1e-6


def test_defaults_and_ranges(sdt, tmp_path):
    # keys missing -> the reference defaults; out of range -> default
    cfg = _write(tmp_path, "b.cfg", "mob-method spme\nbrownian-method lanczos\n"
                 "spme-porder 8\ndelta-t -1.0\nmob-interval 0\n")
    mod = _write(tmp_path, "m.mod", "par X 1.0 0.0\nsteric 2.0 125.0\n")
    sb = sdt.SimBox(cfg, mod, os.path.join(REF, "N1000_Phi0.1.xyz"))
    p = sb.params()
    assert p["spme_xi"] == 0.5 and p["spme_rmax"] == 4.0 and p["spme_dim"] == 128
    assert p["spme_porder"] == 4            # [4, 6] through the config
    assert p["delta_t"] == 0.001 and p["mob_interval"] == 100
    assert p["lanczos_maxiters"] == 50 and p["lanczos_maxnrhs"] == 20
    assert p["lanczos_tol"] == 1e-3
    assert p["traj_interval"] == -1 and p["traj_start"] == 0
    assert p["steric_r0"] == 2.0 and p["steric_k0"] == 125.0
    assert p["precision"] == 0 and p["seed"] == 1234
    assert np.all(sb.radii() == 1.0) and sb.npos == 1000


@pytest.mark.parametrize("cfg_extra,mod,msg", [
    ("delta-t 0.1\n", "", "duplicate key"),
    ("", "bond 1 2 3.0 125\n", "bond"),
    ("", "steric 2 125\nsteric 2 125\n", "steric"),
])
def test_rejected_inputs(sdt, tmp_path, cfg_extra, mod, msg):
    cfg = _write(tmp_path, "c.cfg", BASE_CFG + cfg_extra)
    m = _write(tmp_path, "c.mod", mod)
    with pytest.raises(sdt.SdtError) as e:
        sdt.SimBox(cfg, m, os.path.join(REF, "N3_Phi0.1.xyz"))
    assert msg in str(e.value)


def test_reference_example_config_accepted(sdt, tmp_path):
    # the reference's own data/example.cfg: mob-method ewald, brownian-method
    # cholesky (any other value than lanczos / random selects Cholesky, as
    # simbox_bd.cc:130-139), ewald-tol 1e-4, lanczos-maxnrhs 20
    sb = sdt.SimBox(os.path.join(REF, "example.cfg"), os.path.join(REF, "example.mod"),
                    os.path.join(REF, "N1000_Phi0.1.xyz"))
    p = sb.params()
    assert p["mob_method"] == 1 and p["brownian_method"] == 2 and p["ewald_tol"] == 1e-4
    # Cholesky needs the dense mobility (simbox_bd.cc:141-144)
    cfg = _write(tmp_path, "d.cfg", BASE_CFG.replace("lanczos\n", "chol\n", 1))
    with pytest.raises(sdt.SdtError) as e:
        sdt.SimBox(cfg, os.path.join(REF, "mono.mod"), os.path.join(REF, "N3_Phi0.1.xyz"))
    assert "explicit matrix" in str(e.value)


def test_bad_xyz(sdt, tmp_path):
    cfg = _write(tmp_path, "e.cfg", BASE_CFG)
    mod = os.path.join(REF, "mono.mod")
    for text in ["3\n0 0.0 5 5 5\nX 1 2 3\n",                 # fewer particles than npos
                 "2\n0 0.0 5 6 5\nX 1 2 3\nX 1 1 1\n",        # not cubic
                 "0\n0 0.0 5 5 5\n"]:
        x = _write(tmp_path, "bad.xyz", text)
        with pytest.raises(sdt.SdtError):
            sdt.SimBox(cfg, mod, x)


def test_restart_frame_round_trip(sdt, tmp_path):
    cfg = _write(tmp_path, "f.cfg", BASE_CFG + "seed 987654321987\n")
    sb = sdt.SimBox(cfg, os.path.join(REF, "poly.mod"), os.path.join(REF, "N412.xyz"))
    out = tmp_path / "restart.xyz"
    sb.write_xyz(out)
    n, head, types, pos = _read_xyz(out)
    assert n == 412 and head[0] == "0" and int(head[5]) == 987654321987
    assert np.array_equal(pos, sb.positions())            # %.17g round-trips
    # the frame is itself a valid input: same particles, radii and seed
    sb2 = sdt.SimBox(cfg, os.path.join(REF, "poly.mod"), out)
    assert np.array_equal(sb2.positions(), sb.positions())
    assert np.array_equal(sb2.radii(), sb.radii())
    assert sb2.params()["seed"] == 987654321987
