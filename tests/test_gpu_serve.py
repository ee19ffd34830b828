"""The serving loop (continuous batching of BD jobs on one GPU): jobs admitted
by predicted device footprint while others run, their steps overlapping on
the device. Each job's result in the mix must equal its result alone (the
same BD step, src/simbox/simbox_bd.cc:296-366, run by itself): bit for bit on
the optimised path; to 1e-10 of the displacement on the plain FP64 path,
whose atomic spread sums in a run-dependent order.
"""
import numpy as np
import pytest

from conftest import box_for

pytestmark = pytest.mark.gpu

JOBS = [(300, 0.2, 24, 0.00), (650, 0.15, 32, 0.00), (420, 0.25, 24, 0.00),
        (500, 0.1, 30, 0.05), (250, 0.3, 24, 0.10), (700, 0.2, 32, 0.15)]


def _job_systems(sdt, spec=JOBS, seed=61):
    out = []
    for k, (n, phi, K, arr) in enumerate(spec):
        L = box_for(n, phi)
        out.append(dict(pos=sdt.rsa_positions(n, L, 1.0, seed + k), rdi=np.ones(n), box=L, dim=K,
                        arrival=arr, seed=1000 + k))
    return out


def test_serve_results_equal_alone(sdt):
    jobs = _job_systems(sdt)
    nsteps, mi = 12, 5
    sv = sdt.Server(sdt.SERVE_FP32, 1.0, 3, max_nrhs=4, mob_interval=mi)
    fp = [sv.footprint(len(j["rdi"]), j["dim"], j["box"], 5.0) for j in jobs]
    # room for the two largest jobs at once: the others queue for memory
    cap = sorted(fp)[-1] + sorted(fp)[-2] + 1.0
    sv = sdt.Server(sdt.SERVE_FP32, cap, 3, max_nrhs=4, mob_interval=mi)
    for j in jobs:
        sv.submit(j["pos"], j["rdi"], j["box"], 0.896, 5.0, j["dim"], nsteps, j["seed"],
                  j["arrival"])
    sv.run()
    st = sv.stats()
    assert st["peak_bytes"] <= cap and st["peak_active"] >= 2
    info = [sv.job(k) for k in range(len(jobs))]
    for k, j in enumerate(jobs):
        assert info[k]["state"] == "done" and info[k]["steps"] == nsteps, info[k]
        assert info[k]["admit"] >= j["arrival"]
        assert sv.step_ms(k).size == nsteps
        bd = sdt.BDSim(j["pos"], j["rdi"], j["box"], 0.896, 5.0, j["dim"], 6, sdt.FP32, 100, 4,
                       1e-6, 2.0, 125.0, 1e-3, mi, j["seed"])
        bd.init()
        bd.advance(nsteps)
        assert np.array_equal(sv.positions(k), bd.positions()), k
    # jobs overlapped on the device, and some waited for memory to free up
    overlap = any(a["admit"] < b["finish"] and b["admit"] < a["finish"]
                  for i, a in enumerate(info) for b in info[i + 1:])
    waited = any(any(o["finish"] <= a["admit"] for o in info if o is not a) for a in info)
    assert overlap and waited


def test_serve_plain_kind_equals_alone(sdt):
    jobs = _job_systems(sdt, JOBS[:2], seed=71)
    sv = sdt.Server(sdt.SERVE_PLAIN, 64e9, 4, max_nrhs=4, mob_interval=4)
    for j in jobs:
        sv.submit(j["pos"], j["rdi"], j["box"], 0.896, 5.0, j["dim"], 6, j["seed"], 0.0)
    sv.run()
    for k, j in enumerate(jobs):
        assert sv.job(k)["state"] == "done"
        pb = sdt.PlainBDSim(j["pos"], j["rdi"], j["box"], 0.896, 5.0, j["dim"], 6, 100, 4, 1e-6,
                            2.0, 125.0, 1e-3, 4, j["seed"])
        pb.init()
        pb.advance(6)
        # the plain path spreads with FP64 atomics, whose summation order
        # varies from run to run: two runs agree to rounding, not bit for bit
        d_alone = pb.positions() - j["pos"]
        d_mix = sv.positions(k) - j["pos"]
        assert np.linalg.norm(d_mix - d_alone) <= 1e-10 * np.linalg.norm(d_alone)


def test_serve_refusal_and_deadline(sdt):
    jobs = _job_systems(sdt, JOBS[:3], seed=81)
    sv = sdt.Server(sdt.SERVE_FP32, 1.0, 2, max_nrhs=4, mob_interval=5)
    small = sv.footprint(300, 24, jobs[0]["box"], 5.0)
    sv = sdt.Server(sdt.SERVE_FP32, 1.2 * small, 2, max_nrhs=4, mob_interval=5)
    # 3000 particles on a 48^3 mesh: a footprint larger than the capacity
    Lb = box_for(3000, 0.2)
    assert sv.footprint(3000, 48, Lb, 5.0) > 1.2 * small
    sv.submit(sdt.rsa_positions(3000, Lb, 1.0, 5), np.ones(3000), Lb, 0.896, 5.0, 48, 5, 1, 0.0)
    j0 = jobs[0]
    sv.submit(j0["pos"], j0["rdi"], j0["box"], 0.896, 5.0, j0["dim"], 1_000_000, 2, 0.0)
    sv.submit(j0["pos"], j0["rdi"], j0["box"], 0.896, 5.0, j0["dim"], 5, 3, 100.0)
    sv.run(deadline=2.0)
    a, b, c = sv.job(0), sv.job(1), sv.job(2)
    assert a["state"] == "failed" and "capacity" in a["error"]
    assert b["state"] == "stopped" and 0 < b["steps"] < 1_000_000
    assert c["state"] == "queued" and c["admit"] < 0
    assert sv.stats()["wall_s"] < 30.0
