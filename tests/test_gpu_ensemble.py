"""Graph-captured step tails and the batched ensemble of independent systems.

- A BD run whose step tails are replayed HIP graphs equals the eager run bit
  for bit (the graph holds the same launches).
- Every member of an ensemble (uneven particle counts, boxes and meshes; the
  interval work side by side, the tails of all systems in four lane graphs)
  equals the same system run alone (BDSimBox::Advance,
  src/simbox/simbox_bd.cc:296-366) bit for bit, and an FP64 member matches the
  oracle's step.
- The autotuner's choices are valid, change results only by rounding, and are
  read back from its cache.
"""
import numpy as np
import pytest

from conftest import box_for, uniform_system

pytestmark = pytest.mark.gpu

SYS = [(300, 0.30, 24), (1200, 0.15, 48), (700, 0.25, 30), (500, 0.10, 40), (950, 0.35, 32)]


def _systems(sdt, spec=SYS, seed=5):
    out = []
    for k, (n, phi, K) in enumerate(spec):
        L = box_for(n, phi)
        out.append(dict(pos=sdt.rsa_positions(n, L, 1.0, seed + k), rdi=np.ones(n), box=L,
                        xi=0.896, rmax=5.0, dim=K))
    return out


def test_bd_graph_tail_bit_identical(sdt):
    n, box = 400, 21.0
    pos, rdi = uniform_system(n, box, 12)
    runs = []
    for graphs in (True, False):
        bd = sdt.BDSim(pos, rdi, box, 0.9, 5.0, 32, 6, sdt.FP32, 100, 8, 1e-6, 2.0, 125.0, 1e-3,
                       5, 31)
        bd.init()
        bd.set_graphs(graphs)
        bd.advance(7)
        bd.advance(6)
        runs.append(bd.positions())
    assert np.array_equal(runs[0], runs[1])


def test_ensemble_members_equal_standalone_runs(sdt):
    systems = _systems(sdt)
    args = (6, sdt.FP32, 100, 4, 1e-6, 2.0, 125.0, 1e-3, 6)
    ens = sdt.Ensemble(systems, *args, 900)
    ens.init(autotune=False)
    ms = ens.advance(8)
    ms2 = ens.advance(7)
    assert ms.size == 8 and ms2.size == 7 and (ms > 0).all()
    for s, sy in enumerate(systems):
        print("system", s, flush=True)
        bd = sdt.BDSim(sy["pos"], sy["rdi"], sy["box"], sy["xi"], sy["rmax"], sy["dim"], 6,
                       sdt.FP32, 100, 4, 1e-6, 2.0, 125.0, 1e-3, 6, 900 + s)
        bd.init()
        bd.advance(15)
        assert np.array_equal(ens.positions(s), bd.positions()), s
        assert list(ens.last_iters(s)) == list(bd.last_iters())


def test_ensemble_eager_equals_graphs(sdt):
    systems = _systems(sdt, SYS[:3], seed=9)
    res = []
    for graphs in (True, False):
        ens = sdt.Ensemble(systems, 6, sdt.FP32, 100, 4, 1e-6, 2.0, 125.0, 1e-3, 4, 77)
        ens.init()
        ens.set_graphs(graphs)
        ens.advance(10)
        res.append([ens.positions(s) for s in range(len(systems))])
    for a, b in zip(*res):
        assert np.array_equal(a, b)


def test_ensemble_lanes_equal_eager(sdt):
    # nine systems over the four graph lanes (dealt largest first in snake
    # order: lanes of 3, 2, 2 and 2 systems), two mobility intervals, so each
    # lane's graph is captured twice: graph replays equal eager tails and the
    # same systems run alone, bit for bit
    spec = [(200 + 60 * k, 0.1 + 0.025 * k, (24, 30, 32)[k % 3]) for k in range(9)]
    systems = _systems(sdt, spec, seed=31)
    res = []
    for graphs in (True, False):
        ens = sdt.Ensemble(systems, 6, sdt.FP32, 100, 4, 1e-6, 2.0, 125.0, 1e-3, 4, 50)
        ens.init(autotune=False)
        ens.set_graphs(graphs)
        ens.advance(9)
        res.append([ens.positions(s) for s in range(9)])
    for s, sy in enumerate(systems):
        assert np.array_equal(res[0][s], res[1][s]), s
        bd = sdt.BDSim(sy["pos"], sy["rdi"], sy["box"], sy["xi"], sy["rmax"], sy["dim"], 6,
                       sdt.FP32, 100, 4, 1e-6, 2.0, 125.0, 1e-3, 4, 50 + s)
        bd.init()
        bd.advance(9)
        assert np.array_equal(res[0][s], bd.positions()), s


def test_ensemble_fp64_member_matches_oracle(sdt, orc):
    systems = _systems(sdt, SYS[:2], seed=21)
    ens = sdt.Ensemble(systems, 6, sdt.FP64, 100, 4, 1e-7, 2.0, 125.0, 1e-3, 5, 300)
    ens.init()
    ens.advance(7)
    for s, sy in enumerate(systems):
        ob = orc.BDSim(sy["pos"], sy["rdi"], sy["box"], sy["xi"], sy["rmax"], sy["dim"], 6, 100,
                       4, 1e-7, 2.0, 125.0, 1e-3, 5, 300 + s)
        ob.advance(7)
        d_ref = ob.positions() - sy["pos"]
        assert orc.frob_rel_err(d_ref, ens.positions(s) - sy["pos"]) < 1e-6


def test_ensemble_autotune_choices_and_cache(sdt, orc, tmp_path):
    systems = _systems(sdt, SYS[:3], seed=13)
    cache = tmp_path / "tune.txt"
    args = (6, sdt.FP32, 100, 4, 1e-6, 2.0, 125.0, 1e-3, 5, 44)
    ens = sdt.Ensemble(systems, *args)
    ens.init(autotune=True, cache=cache)
    ens.advance(6)
    plain = sdt.Ensemble(systems, *args)
    plain.init(autotune=False)
    plain.advance(6)
    for s in range(len(systems)):
        narrow, wide, ms = ens.tuning(s)
        for c in (narrow, wide):
            assert c[0] in (0, 1, 2) and c[1] in (0, 1, 2, 4)
            assert c[2] == 0 or 32 <= c[2] <= 128
            assert c[3] in (0, 32, 48, 64)
        assert ms[0] > 0 and 0 < ms[1] <= ms[0]
        sy = systems[s]
        d0 = plain.positions(s) - sy["pos"]
        assert orc.frob_rel_err(d0, ens.positions(s) - sy["pos"]) < 1e-5
    lines = cache.read_text().splitlines()
    assert len(lines) >= 2 and all(ln.startswith("spme-v1 ") for ln in lines)
    again = sdt.Ensemble(systems, *args)
    again.init(autotune=True, cache=cache)
    for s in range(len(systems)):
        assert again.tuning(s)[:2] == ens.tuning(s)[:2]
