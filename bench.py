#!/usr/bin/env python
"""BD steps/s of the StokesDT Brownian-dynamics step on one MI355X.

Workload (the single-system configuration the headline metric is quoted on):
N = 100 000 monodisperse spheres (a = 1) at volume fraction 0.3 in a
periodic cube (L = 111.77), seeded random-sequential-addition positions,
SPME mesh 128^3, B-spline order 6, real-space cutoff 5, Ewald xi = 0.896,
block Lanczos (block 8, tol 1e-6, FP64 Krylov basis and small algebra),
mobility interval 100, steric r0 = 2, k0 = 125, dt = 1e-3. A step is one
iteration of BDSimBox::Advance (src/simbox/simbox_bd.cc:296-366).

Timing: after the warm-up steps the run continues untimed to the next
mobility-interval boundary, so the `--steps` timed steps always begin with a
step that holds the mobility update and the block-Lanczos solve of all 100
columns. Each step is timed (the call returns after the device finished).
value = amortised steps/s over whole intervals: steps / time when the
window holds whole intervals, else one interval composed from the window's
interval-start step and the mean of its other steps.

Prints one JSON line: value = FP32-mixed steps/s; also the optimised FP64
GPU path and the plain FP64 GPU path (the reference algorithm as structured
on the CPU) on the same amortisation, the FP32-vs-FP64 errors, the error of M f against
the dense Ewald sum on 2000 sampled particles, the polydisperse path at
N = 100 000 (matvec stage times and FP32 steps/s), per-stage
device times (HIP events on the launching stream), the roofline of the
dominant kernel and the CPU baseline (the FP64 C++/OpenMP oracle timed on a
bounded sample).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("BD steps/s at N=100k (FP32-mixed vs FP64) at ≤1e-5 err; "
          "ensemble p50/p99 latency")
HBM_PEAK_GBS = 8000.0

CFG = dict(n=100_000, phi=0.3, K=128, p=6, rmax=5.0, xi=0.896, b=8, tol=1e-6, max_iters=100,
           mob_interval=100, dt=1e-3, r0=2.0, k0=125.0, seed=2024)


# batched ensemble (64 independent systems)
ENS = dict(nsys=64, nmin=500, nmax=20_000, phimin=0.05, phimax=0.35, kmin=24, kmax=96, p=6,
           xi=0.896, rmax=5.0, b=4, tol=1e-6, max_iters=100, mob_interval=100, dt=1e-3, r0=2.0,
           k0=125.0, seed=33)


# serving loop: seeded Poisson arrivals
SERVE = dict(rate=50.0, nmin=1000, nmax=200_000, phimin=0.05, phimax=0.3, h=0.8, p=6, xi=0.896,
             rmax=5.0, b=8, tol=1e-6, max_iters=100, steps=100, dt=1e-3, mob_interval=100,
             r0=2.0, k0=125.0, seed=55, horizon=6.0, deadline=20.0, max_active=16,
             capacity=288e9)


def box_size(n, phi):
    return (4.0 * np.pi * n / (3.0 * phi)) ** (1.0 / 3.0)


# the sampled dense-Ewald check of M f: rows and truncation tolerance
EWALD_ROWS, EWALD_TOL = 2000, 1e-10
ACCURATE = dict(K=256, p=8, xi=0.5, rmax=9.5)
T_START = time.perf_counter()


def note(msg):
    """A progress line on stderr (the JSON result goes to stdout)."""
    print("[bench %6.1f s] %s" % (time.perf_counter() - T_START, msg), file=sys.stderr,
          flush=True)


def smooth_mesh(L, kmin, kmax, h=0.8):
    """The even 2^a 3^b 5^c in [kmin, kmax] nearest to L / h (larger on a tie)."""
    def smooth(k):
        for f in (2, 3, 5):
            while k % f == 0:
                k //= f
        return k == 1
    cand = [k for k in range(kmin, kmax + 1) if k % 2 == 0 and smooth(k)]
    t = L / h
    return min(cand, key=lambda k: (abs(k - t), -k))


def ensemble_systems(S):
    """64 systems: N log-uniform in [500, 20000], phi uniform in [0.05, 0.35],
    mesh the nearest even 2^a 3^b 5^c to L / 0.8 in [24, 96] (seeded)."""
    e = ENS
    rng = np.random.default_rng(e["seed"])
    out = []
    for k in range(e["nsys"]):
        n = int(round(np.exp(rng.uniform(np.log(e["nmin"]), np.log(e["nmax"])))))
        phi = float(rng.uniform(e["phimin"], e["phimax"]))
        L = box_size(n, phi)
        out.append(dict(pos=S.rsa_positions(n, L, 1.0, e["seed"] + k), rdi=np.ones(n), box=L,
                        xi=e["xi"], rmax=e["rmax"], dim=smooth_mesh(L, e["kmin"], e["kmax"]),
                        phi=phi))
    return out


def lattice_positions(rng, n, L):
    """n non-overlapping unit spheres: a random subset of the sites of an
    m^3 simple cubic lattice (m = ceil(n^(1/3)), spacing L/m > 2 for phi <=
    0.3 and n >= 1000), each jittered within the free gap."""
    m = int(np.ceil(n ** (1.0 / 3.0) - 1e-9))
    while m ** 3 < n:
        m += 1
    a = L / m
    if a <= 2.0:
        raise ValueError("lattice spacing %.3f leaves no room for the spheres" % a)
    sites = rng.choice(m ** 3, size=n, replace=False)
    ijk = np.stack([sites // (m * m), (sites // m) % m, sites % m], axis=1).astype(np.float64)
    j = 0.45 * (a - 2.0)
    return ((ijk + 0.5) * a + rng.uniform(-j, j, (n, 3))).reshape(-1)


def serve_requests(horizon):
    """Requests of the serving run: Poisson arrivals at 50 per second over
    `horizon` seconds; N log-uniform in [1000, 200000], phi uniform in
    [0.05, 0.3], mesh the even 2^a 3^b 5^c nearest to L / 0.8 (at least
    24), 100 steps each (seeded)."""
    c = SERVE
    rng = np.random.default_rng(c["seed"])
    out = []
    t = 0.0
    while True:
        t += float(rng.exponential(1.0 / c["rate"]))
        if t >= horizon:
            break
        n = int(round(np.exp(rng.uniform(np.log(c["nmin"]), np.log(c["nmax"])))))
        phi = float(rng.uniform(c["phimin"], c["phimax"]))
        L = box_size(n, phi)
        out.append(dict(arrival=t, n=n, L=L, dim=smooth_mesh(L, 24, 512, c["h"]),
                        pos=lattice_positions(rng, n, L), seed=int(rng.integers(1 << 62))))
    return out


def run_serving(S, kind, requests):
    """One serving run over the request trace until the deadline."""
    c = SERVE
    free, total = S.device_memory()
    cap = min(c["capacity"], 0.92 * free)
    sv = S.Server(kind, cap, c["max_active"], c["p"], c["max_iters"], c["b"], c["tol"], c["r0"],
                  c["k0"], c["dt"], c["mob_interval"])
    for r in requests:
        sv.submit(r["pos"], np.ones(r["n"]), r["L"], c["xi"], c["rmax"], r["dim"], c["steps"],
                  r["seed"], r["arrival"])
    sv.run(c["deadline"])
    st = sv.stats()
    jobs = [sv.job(k) for k in range(len(requests))]
    lat = np.concatenate([sv.step_ms(k) for k in range(len(requests))] + [np.zeros(0)])
    steps = sum(j["steps"] for j in jobs)
    psteps = sum(j["steps"] * r["n"] for j, r in zip(jobs, requests))
    done = [j for j in jobs if j["state"] == "done"]
    adm = [j for j in jobs if j["admit"] >= 0 and j["state"] != "failed"]
    wait = [j["admit"] - j["arrival"] for j in adm]
    return dict(
        requests=len(requests), completed=len(done),
        completed_per_s=round(len(done) / st["wall_s"], 3),
        admitted=len(adm), refused=sum(j["state"] == "failed" for j in jobs),
        still_queued=sum(j["state"] == "queued" for j in jobs),
        steps_per_s=round(steps / st["wall_s"], 2),
        particle_steps_per_s=round(psteps / st["wall_s"], 1),
        step_latency_ms_p50=round(float(np.percentile(lat, 50)), 3) if lat.size else None,
        step_latency_ms_p99=round(float(np.percentile(lat, 99)), 3) if lat.size else None,
        queue_wait_s_p50=round(float(np.percentile(wait, 50)), 3) if wait else None,
        wall_s=round(st["wall_s"], 3), capacity_gb=round(cap / 1e9, 1),
        peak_reserved_gb=round(st["peak_bytes"] / 1e9, 2), peak_jobs_on_device=st["peak_active"])


def run_ensemble(S, steps, warmup):
    """Config-3 ensemble (FP32-mixed, b = 4 per system, autotuned): warmup
    steps and on to an interval boundary untimed, then `steps` timed steps
    (host wall time per step, every system advanced); then the same number of
    tail steps again with eager launches instead of graphs, for comparison."""
    import tempfile
    e = ENS
    mi = e["mob_interval"]
    systems = ensemble_systems(S)
    ens = S.Ensemble(systems, e["p"], S.FP32, e["max_iters"], e["b"], e["tol"], e["r0"], e["k0"],
                     e["dt"], mi, e["seed"])
    t = time.perf_counter()
    with tempfile.TemporaryDirectory() as tmp:
        ens.init(autotune=True, cache=os.path.join(tmp, "tune.txt"))
    t_tune = time.perf_counter() - t
    first = warmup + (-warmup) % mi
    ens.advance(first)
    lat = ens.advance(steps)
    value, how = amortise(list(lat / 1000.0), first, mi)
    tails = [x for j, x in enumerate(lat) if (first + j) % mi != 0]
    # eager tails over the same kind of steps (the window ends inside an
    # interval unless it holds whole intervals)
    n_eager = min(len(tails), mi - 1 - ((first + steps - 1) % mi)) if tails else 0
    eager = []
    if n_eager > 0:
        ens.set_graphs(False)
        eager = list(ens.advance(n_eager))
        ens.set_graphs(True)
    ntot = int(sum(len(sy["rdi"]) for sy in systems))
    gain_n = [ens.tuning(s)[2] for s in range(len(systems))]
    ms = np.array(gain_n)
    return dict(
        systems=len(systems), particles=ntot,
        n_range=[int(min(len(sy["rdi"]) for sy in systems)),
                 int(max(len(sy["rdi"]) for sy in systems))],
        meshes=sorted(set(int(sy["dim"]) for sy in systems)),
        steps_per_s=round(value, 4), particle_steps_per_s=round(value * ntot, 1),
        amortisation=how, first_step=first, steps=steps,
        step_latency_ms_p50=round(float(np.percentile(lat, 50)), 4),
        step_latency_ms_p99=round(float(np.percentile(lat, 99)), 4),
        interval_start_step_ms=round(float(np.mean([x for j, x in enumerate(lat)
                                                    if (first + j) % mi == 0])), 3),
        tail_step_ms_graph=round(float(np.mean(tails)), 4) if tails else None,
        tail_step_ms_eager=round(float(np.mean(eager)), 4) if eager else None,
        autotune_s=round(t_tune, 2),
        autotune_matvec_ms_sum={"narrow_default": round(float(ms[:, 0].sum()), 3),
                                "narrow_tuned": round(float(ms[:, 1].sum()), 3),
                                "wide_default": round(float(ms[:, 2].sum()), 3),
                                "wide_tuned": round(float(ms[:, 3].sum()), 3)},
        workload="64 systems, N log-uniform [500, 20000], phi U[0.05, 0.35], mesh nearest even "
                 "2^a3^b5^c to L/0.8 in [24, 96], p=6 rc=5 xi=0.896, FP32-mixed, block Lanczos "
                 "b=4 tol=1e-6, mob-interval=100, dt=1e-3, steric r0=2 k0=125")


def alg_bytes(stage, b, n, K, e, poly=False, nnz=0):
    """Algorithmic HBM bytes of one launch group processing b right-hand
    sides (compact meshes, weights recomputed, nothing re-read)."""
    nc = 6 if poly else 3
    if stage == "real_space":  # CSR, positions and radii, FP64 X in, Y read and written
        return 4 * nnz + 4 * (n + 1) + 4 * n * 8 + 3 * n * b * 8 * 3
    K3 = K ** 3
    spec = K * K * (K // 2 + 1) * 2 * e
    if stage == "spread":      # particle offsets/bases, FP64 forces in, meshes out
        return 3 * n * (e + 4) + 3 * n * b * 8 + nc * b * K3 * e
    if stage == "interp":      # meshes in, particle data, FP64 result written
        return nc * b * K3 * e + 3 * n * (e + 4) + 3 * n * b * 8
    if stage == "fft_xy_fwd":  # real meshes in, half spectra out
        return nc * b * (K3 * e + spec)
    if stage == "fft_xy_inv":
        return nc * b * (spec + K3 * e)
    if stage == "fft_z_influence":
        return 2 * nc * b * spec
    return 0


def lanczos_batch_cols():
    """Columns of the fused Lanczos matvec: the library runs up to 128
    columns (whole groups of b) in lockstep."""
    c = CFG
    return min(c["mob_interval"], (128 // c["b"]) * c["b"])


def read_traffic(stage):
    """HBM bytes per launch of a stage by launch width, from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary (lines 'traffic stage=<s>
    columns=<b> bytes=<B>'); empty when absent."""
    out = {}
    path = os.path.join(ROOT, "profiles", "hbm_traffic_bd_n100k.txt")
    try:
        with open(path) as fh:
            for line in fh:
                if not line.startswith("traffic "):
                    continue
                kv = dict(t.split("=", 1) for t in line.split()[1:] if "=" in t)
                if kv.get("stage") == stage and "columns" in kv and "bytes" in kv:
                    out[int(kv["columns"])] = float(kv["bytes"])
    except OSError:
        pass
    return out


def kernel_sample():
    """Short fixed kernel sample for counter runs: a calibration copy of
    2^26 doubles, then the FP32 matvec pipeline at the workload shape with
    1 and with the Lanczos batch width of columns (2 repetitions each), then
    one block-Lanczos solve of a mobility interval's columns (the Gram and
    block-product kernels at their real shapes)."""
    import stokesdt_amd as S
    c = CFG
    n = c["n"]
    L = box_size(n, c["phi"])
    ms = S.bench_copy(1 << 26, 3)
    print("copy 2^26 doubles: %.4f ms (%.1f GB/s)" % (ms, 2 * 8 * (1 << 26) / ms / 1e6),
          flush=True)
    pos = S.rsa_positions(n, L, 1.0, c["seed"])
    m = S.MobSpme(n, np.ones(n), L, c["xi"], c["rmax"], c["K"], c["p"], S.FP32)
    assert m.init()
    m.update(pos)
    for b in (1, lanczos_batch_cols()):
        t = m.stage_times(b, 2)
        print("columns %d: " % b + " ".join("%s=%.4f" % kv for kv in
                                           zip(["spread", "fft_xy_fwd", "fft_z_influence",
                                                "fft_xy_inv", "interp", "real_space", "whole"],
                                               t)), flush=True)
    nc = c["mob_interval"]
    z = S.gaussian(c["seed"], 0, 3 * n * nc).reshape(nc, 3 * n).T.copy()
    bl = S.BrwnLanczos(3 * n, c["max_iters"], c["b"], c["tol"])
    assert bl.init()
    t0 = time.perf_counter()
    y, ok = bl.compute(m, z)
    print("block Lanczos, %d columns in groups of %d: %.3f s, iterations %s" %
          (nc, c["b"], time.perf_counter() - t0, list(bl.iters())), flush=True)


def step_window(bd, steps):
    """Advances `steps` steps one at a time (each call returns after the
    device finished) and returns the wall time of every step in seconds."""
    lat = []
    for _ in range(steps):
        t = time.perf_counter()
        bd.advance(1)
        lat.append(time.perf_counter() - t)
    return lat


def amortise(lat, first_step, mi):
    """Steps/s over whole mobility intervals from a window of per-step times
    whose first step has index first_step. A step whose index is a multiple
    of mi holds the mobility update and the Lanczos solve of its interval.
    A window of whole intervals gives steps / time directly; otherwise one
    interval is composed as (interval-start steps, mean) + (mi - 1) x (other
    steps, mean)."""
    k = len(lat)
    starts = [t for j, t in enumerate(lat) if (first_step + j) % mi == 0]
    tails = [t for j, t in enumerate(lat) if (first_step + j) % mi != 0]
    if not starts:
        raise RuntimeError("the timed window holds no mobility update")
    if k % mi == 0 or not tails:
        return k / sum(lat), "whole intervals: steps / window time"
    per = float(np.mean(starts)) + (mi - 1) * float(np.mean(tails))
    return mi / per, ("one interval composed from the window: mean interval-start step + "
                      "%d x mean other step" % (mi - 1))


def run_gpu(S, pos, rdi, L, prec, steps, warmup):
    """Optimised path: the autotuner's launch choices, warmup steps, then
    untimed steps up to the next mobility-interval boundary, then `steps`
    timed steps (the first of them holds the update and the Lanczos solve of
    all mob_interval columns)."""
    c = CFG
    mi = c["mob_interval"]
    bd = S.BDSim(pos, rdi, L, c["xi"], c["rmax"], c["K"], c["p"], prec, c["max_iters"], c["b"],
                 c["tol"], c["r0"], c["k0"], c["dt"], mi, c["seed"])
    bd.init()
    tuned = bd.autotune()
    first = warmup + (-warmup) % mi
    if first:
        bd.advance(first)
    bd.set_profiling(True)
    S.lib().sdt_device_synchronize()
    lat = step_window(bd, steps)
    prof = bd.profile()
    iters = bd.last_iters()
    bd.set_profiling(False)
    value, how = amortise(lat, first, mi)
    return dict(value=value, how=how, lat=lat, first=first, prof=prof,
                iters=[int(x) for x in iters], pos=bd.positions(),
                tuning=dict(narrow=tuned[0], wide=tuned[1],
                            matvec_ms=[round(float(x), 4) for x in tuned[2]]))


POLY = dict(n=100_000, phi_max=0.3, amin=0.5, amax=1.5, K=180, p=6, seed=7)


def run_poly(S):
    """The polydisperse path at N = 100 000 (radii uniform in [0.5, 1.5],
    sequential addition at the largest radius, phi 0.3 at that radius, mesh
    180^3, spacing 0.93, whose half-spectrum planes fit LDS): stage times of the
    matvec at 1 and 100 columns, FP32-mixed and FP64, and FP32 steps/s over
    one mobility interval composed as (interval-start step) + 99 x (mean of
    9 other steps)."""
    c, q = CFG, POLY
    n = q["n"]
    L = (4.0 * np.pi * n * q["amax"] ** 3 / (3.0 * q["phi_max"])) ** (1.0 / 3.0)
    pos = S.rsa_positions(n, L, q["amax"], q["seed"])
    rdi = np.random.default_rng(q["seed"]).uniform(q["amin"], q["amax"], n)
    out = dict(workload="N=%d radii U[%.1f, %.1f] L=%.2f mesh %d^3 p=%d rc=%.0f xi=%.3f "
                        "block-Lanczos b=%d tol=%g mob-interval=%d"
                        % (n, q["amin"], q["amax"], L, q["K"], q["p"], c["rmax"], c["xi"], c["b"],
                           c["tol"], c["mob_interval"]))
    names = ["spread", "fft_xy_fwd", "fft_z_influence", "fft_xy_inv", "interp", "real_space",
             "whole"]
    for prec, tag in ((S.FP32, "fp32_mixed"), (S.FP64, "fp64")):
        m = S.MobSpme(n, rdi, L, c["xi"], c["rmax"], q["K"], q["p"], prec)
        assert m.init()
        m.update(pos)
        out[tag + "_matvec_ms"] = {str(b): dict(zip(names, [round(float(x), 4) for x in
                                                          m.stage_times(b, 3)]))
                                   for b in (1, 100)}
        del m
    mi = c["mob_interval"]
    bd = S.BDSim(pos, rdi, L, c["xi"], c["rmax"], q["K"], q["p"], S.FP32, c["max_iters"], c["b"],
                 c["tol"], c["r0"], c["k0"], c["dt"], mi, c["seed"])
    bd.init()
    lat = step_window(bd, 10)
    value, how = amortise(lat, 0, mi)
    out["fp32_mixed_steps_per_s"] = round(value, 3)
    out["interval_start_step_ms"] = round(lat[0] * 1000.0, 2)
    out["other_step_ms_mean"] = round(float(np.mean(lat[1:])) * 1000.0, 3)
    out["lanczos_iters"] = [int(x) for x in bd.last_iters()]
    return out


def run_plain(S, pos, rdi, L, steps):
    """The plain FP64 GPU path (the reference algorithm as structured on the
    CPU, on the device): library kernels are loaded on a small system first,
    then the full system is timed from step 0, whose step holds the update
    and the group-after-group Lanczos solve."""
    c = CFG
    mi = c["mob_interval"]
    nw = 2000
    Lw = box_size(nw, c["phi"])
    w = S.PlainBDSim(S.rsa_positions(nw, Lw, 1.0, 7), np.ones(nw), Lw, c["xi"], c["rmax"], 32,
                     c["p"], c["max_iters"], c["b"], c["tol"], c["r0"], c["k0"], c["dt"], mi, 7)
    w.init()
    w.advance(2)
    del w
    bd = S.PlainBDSim(pos, rdi, L, c["xi"], c["rmax"], c["K"], c["p"], c["max_iters"], c["b"],
                      c["tol"], c["r0"], c["k0"], c["dt"], mi, c["seed"])
    bd.init()
    bd.set_profiling(True)
    lat = step_window(bd, steps)
    prof = bd.profile()
    value, how = amortise(lat, 0, mi)
    return dict(value=value, how=how, lat=lat, prof=prof,
                iters=[int(x) for x in bd.last_iters()])


def displacement_error(S, O, pos, rdi, L):
    """FP32-mixed vs FP64: M f and D = M^{1/2} Z for one 8-column group at
    the start configuration (same positions, forces and Gaussian block)."""
    c = CFG
    n = rdi.size
    f = O.gaussian(11, 0, 3 * n) if O else np.random.default_rng(11).standard_normal(3 * n)
    z = S.gaussian(c["seed"], 0, 3 * n * c["b"]).reshape(c["b"], 3 * n).T.copy()
    out = {}
    for prec in (S.FP64, S.FP32):
        m = S.MobSpme(n, rdi, L, c["xi"], c["rmax"], c["K"], c["p"], prec)
        assert m.init()
        m.update(pos)
        mf = m.mul_vector(f)
        bl = S.BrwnLanczos(3 * n, c["max_iters"], c["b"], c["tol"])
        assert bl.init()
        y, ok = bl.compute(m, z)
        out[prec] = (mf, y)
    e_mf = float(np.linalg.norm(out[0][0] - out[1][0]) / np.linalg.norm(out[0][0]))
    e_d = float(np.linalg.norm(out[0][1] - out[1][1]) / np.linalg.norm(out[0][1]))
    # the dense Ewald sum (EwaldVectorKernel) on 2000 seeded particles
    rows = np.sort(np.random.default_rng(12).choice(n, EWALD_ROWS, replace=False))
    t0 = time.perf_counter()
    v_ew, xi_ew = S.rpy_ewald_rows(pos, rdi, L, EWALD_TOL, f, rows)
    t_ew = time.perf_counter() - t0
    ew = dict(rows=EWALD_ROWS, tol=EWALD_TOL, xi=round(xi_ew, 5), seconds=round(t_ew, 3))
    for prec, name in ((S.FP64, "fp64_spme"), (S.FP32, "fp32_mixed_spme")):
        mf = out[prec][0].reshape(-1, 3)[rows]
        ew["mf_err_" + name] = float(np.linalg.norm(mf - v_ew) / np.linalg.norm(v_ew))
    # the same rows from an SPME operator converged to the Ewald sum (finer
    # mesh, smaller xi, wider real space): the gap above is the workload's
    # discretisation, not the implementation
    m = S.MobSpme(n, rdi, L, ACCURATE["xi"], ACCURATE["rmax"], ACCURATE["K"], ACCURATE["p"],
                  S.FP64)
    assert m.init()
    m.update(pos)
    mf = m.mul_vector(f).reshape(-1, 3)[rows]
    ew["mf_err_converged_fp64_spme"] = float(np.linalg.norm(mf - v_ew) / np.linalg.norm(v_ew))
    ew["converged_spme"] = "mesh %d^3 p=%d xi=%.2f rc=%.1f" % (ACCURATE["K"], ACCURATE["p"],
                                                              ACCURATE["xi"], ACCURATE["rmax"])
    ew["note"] = ("the workload's mesh spacing L/128 = 0.87 at xi = 0.896 ends the reciprocal "
                  "sum where exp(-k^2/4xi^2) is still 1.8 % of its peak, so its SPME operator "
                  "(FP64 or FP32) is mf_err_fp64_spme from the Ewald sum; FP32 vs FP64 differ "
                  "by mf_err_fp32_vs_fp64")
    return e_mf, e_d, ew


def cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(O, pos, rdi, L, iters):
    """Times the FP64 C++/OpenMP oracle (a structural restatement of the
    reference: stored P/ind, 8-colour spread, per-RHS FFT loop, stored BSR,
    BlockLanczos with Y recomputed every iteration) on a bounded sample of
    the same workload: one mobility update, one single-RHS matvec, one
    steric force and one whole Lanczos group of b columns run to
    convergence. One mobility interval is composed from these: the update,
    every group (the timed group scaled by each group's iteration count and
    width, with the GPU run's counts), and mob_interval x (steric + matvec)."""
    c = CFG
    n = rdi.size
    m = O.MobSpme(n, rdi, L, c["xi"], c["rmax"], c["K"], c["p"])
    assert m.init()
    t = time.perf_counter()
    m.update(pos, rdi)
    t_update = time.perf_counter() - t
    f = np.random.default_rng(5).standard_normal(3 * n)
    t = time.perf_counter()
    m.mul_vector(f)
    t_mv = time.perf_counter() - t
    t = time.perf_counter()
    O.steric(pos, rdi, L, c["r0"], c["k0"])
    t_steric = time.perf_counter() - t
    z = O.gaussian(3, 0, 3 * n * c["b"]).reshape(c["b"], 3 * n).T.copy()
    t = time.perf_counter()
    _, it = O.lanczos(m, z, c["max_iters"], c["b"], c["tol"])
    t_group = time.perf_counter() - t
    k0 = int(it[0]) if int(it[0]) >= 0 else c["max_iters"] - 1
    mi = c["mob_interval"]
    groups = [c["b"]] * (mi // c["b"]) + ([mi % c["b"]] if mi % c["b"] else [])
    lz = sum(t_group * ((k if k >= 0 else c["max_iters"] - 1) + 1) / (k0 + 1) * g / c["b"]
             for k, g in zip(iters, groups))
    t_interval = t_update + lz + mi * (t_steric + t_mv)
    sample = ("N=100000 workload pieces timed on the host: 1 mobility update, 1 single-RHS "
              "matvec, 1 steric force, 1 block-Lanczos group (b=8) to convergence (%d "
              "iterations); one %d-step interval composed with the GPU run's iteration counts"
              % (k0 + 1, mi))
    return dict(value=mi / t_interval, unit="steps/s", cores=O.num_threads(), kind="port",
                sample=sample, model=cpu_model(),
                pieces_s=dict(update=round(t_update, 3), matvec=round(t_mv, 3),
                              steric=round(t_steric, 3), lanczos_group_b8=round(t_group, 3),
                              lanczos_group_iters=k0 + 1))


def roofline(S, r, pos, rdi, L):
    """Roofline of the dominant stage of the timed window (largest device
    time). Its launches: one single-column matvec per step and one fused
    matvec per block-Lanczos iteration over the columns still iterating; the
    timer records every launch's width, and algorithmic and measured bytes are
    affine in the width."""
    c = CFG
    n = c["n"]
    cand = ["spread", "fft_xy_fwd", "fft_z_influence", "fft_xy_inv", "interp", "real_space"]
    dom = max(cand, key=lambda k: r["prof"][k][0])
    tot_ms, launches, ncols = r["prof"][dom]
    btot = lanczos_batch_cols()
    nnz = 0
    if dom == "real_space":
        m = S.MobSpme(n, rdi, L, c["xi"], c["rmax"], c["K"], c["p"], S.FP32)
        assert m.init()
        m.update(pos)
        nnz = m.real_nnz()
        del m
    if launches == 0 or tot_ms <= 0:
        return dict(kernel=dom, bound="hbm", achieved=None, peak=HBM_PEAK_GBS, unit="GB/s",
                    frac=None, traffic=None, note="no launches of %s recorded" % dom)
    a0 = alg_bytes(dom, 0, n, c["K"], 4, nnz=nnz)
    a1 = alg_bytes(dom, 1, n, c["K"], 4, nnz=nnz) - a0
    tot_bytes = launches * a0 + a1 * ncols
    achieved = tot_bytes / (tot_ms * 1e-3) / 1e9
    traffic = None
    tr = read_traffic(dom)
    if 1 in tr and btot in tr:
        per_col = (tr[btot] - tr[1]) / (btot - 1)
        traffic = round((launches * (tr[1] - per_col) + per_col * ncols) / launches)
    return dict(kernel=dom, bound="hbm", achieved=round(achieved, 2), peak=HBM_PEAK_GBS,
                unit="GB/s", frac=round(achieved / HBM_PEAK_GBS, 5), traffic=traffic,
                launches=launches, columns=ncols, avg_launch_ms=round(tot_ms / launches, 5),
                alg_bytes_per_launch=round(tot_bytes / launches))


def stage_dict(prof):
    out = {k: dict(total_ms=round(v[0], 4), launches=v[1]) for k, v in prof.items()}
    for k, v in prof.items():
        if v[2]:
            out[k]["columns"] = v[2]
    return out


def window_info(r, mi):
    lat = np.array(r["lat"]) * 1000.0
    first = r.get("first", 0)
    st = [t for j, t in enumerate(lat) if (first + j) % mi == 0]
    tl = [t for j, t in enumerate(lat) if (first + j) % mi != 0]
    return dict(first_step=first, steps=len(lat), interval_starts=len(st),
                interval_start_step_ms=round(float(np.mean(st)), 3) if st else None,
                other_step_ms_mean=round(float(np.mean(tl)), 4) if tl else None,
                window_steps_per_s=round(len(lat) / (lat.sum() / 1000.0), 4),
                amortisation=r["how"])


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 GPU paths")
    ap.add_argument("--no-plain", action="store_true", help="skip the plain FP64 GPU path")
    ap.add_argument("--no-err", action="store_true", help="skip the FP32-vs-FP64 error check")
    ap.add_argument("--no-poly", action="store_true", help="skip the polydisperse N=100k leg")
    ap.add_argument("--no-ensemble", action="store_true", help="skip the 64-system ensemble")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving loop")
    ap.add_argument("--kernel-sample", action="store_true",
                    help="run only the short kernel sample used for counter collection")
    args = ap.parse_args()
    if args.kernel_sample:
        kernel_sample()
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: the Brownian-dynamics step of one system runs on "
              "a single MI355X here (multi-GPU sharding is design-only); run with --gpus 1",
              file=sys.stderr)
        sys.exit(2)
    if args.steps < 1:
        print("--steps must be at least 1", file=sys.stderr)
        sys.exit(2)

    import stokesdt_amd as S
    S.lib()
    c = CFG
    mi = c["mob_interval"]
    n = c["n"]
    L = box_size(n, c["phi"])
    pos = S.rsa_positions(n, L, 1.0, c["seed"])
    rdi = np.ones(n)

    note("FP32-mixed window")
    r32 = run_gpu(S, pos, rdi, L, S.FP32, args.steps, args.warmup)
    value = r32["value"]
    note("FP32-mixed: %.2f steps/s" % value)
    res = dict(metric=METRIC, value=round(value, 4), unit="steps/s", n_gpus=1, steps=args.steps,
               warmup=args.warmup, ms_per_step=round(1000.0 / value, 4),
               higher_is_better=True, scaling="weak", vs_baseline=None, dtype="fp32",
               data="synthetic (seeded RSA positions, Philox-4x32-10 Gaussian blocks)")
    res["config"] = {
        "workload": "single system N=100000 phi=0.3 (L=%.4f) mesh 128^3 p=6 rc=5 xi=0.896 "
                    "block-Lanczos b=8 tol=1e-6 mob-interval=100 dt=1e-3 steric r0=2 k0=125" % L,
        "precision": "FP32 spread/FFT/influence/interp/real-space with FP32 accumulation; "
                     "FP64 vectors, Krylov basis, Gram/QR/eigensolve",
        "parallelism": "single GPU",
    }
    res["window"] = window_info(r32, mi)
    lat = np.array(r32["lat"]) * 1000.0
    res["step_latency_ms_p50"] = round(float(np.percentile(lat, 50)), 4)
    res["step_latency_ms_p99"] = round(float(np.percentile(lat, 99)), 4)
    res["lanczos_iters_fp32"] = r32["iters"]
    res["autotune"] = dict(r32["tuning"], note="(fft_order, interp_sxt, interp_ch, "
                           "spread_stage_kb), 0 = built-in heuristic; matvec_ms = default and "
                           "tuned for 1 column, then for the Lanczos batch")
    res["stage_device_ms"] = stage_dict(r32["prof"])
    res["roofline"] = roofline(S, r32, pos, rdi, L)
    O = None
    if not args.no_cpu or not args.no_err:
        try:
            from oracle import oracle as O
            O.lib()
        except Exception as exc:   # the checker is optional for the GPU figures
            O = None
            res["oracle_error"] = str(exc)
    if not args.no_fp64:
        note("FP64 window")
        r64 = run_gpu(S, pos, rdi, L, S.FP64, args.steps, args.warmup)
        res["fp64_gpu_steps_per_s"] = round(r64["value"], 4)
        res["speedup_fp32_vs_fp64_gpu"] = round(value / r64["value"], 3)
        res["lanczos_iters_fp64"] = r64["iters"]
        d32 = r32["pos"] - pos
        d64 = r64["pos"] - pos
        res["trajectory_disp_err_fp32_vs_fp64"] = float(np.linalg.norm(d32 - d64) /
                                                        np.linalg.norm(d64))
        if not args.no_plain:
            note("plain FP64 window")
            rp = run_plain(S, pos, rdi, L, args.steps)
            res["plain_fp64_gpu"] = dict(
                steps_per_s=round(rp["value"], 4), window=window_info(rp, mi),
                lanczos_iters=rp["iters"], stage_device_ms=stage_dict(rp["prof"]),
                form="per-RHS recip loop, stored P/ind tables, FP64 global-atomic spread, "
                     "hipFFT R2C/C2R x3 each way, separate influence and interp kernels, "
                     "stored FP64 3x3 blocks + SpMV, BlockLanczos group after group "
                     "(rocSOLVER Householder QR, rocBLAS products, Y every iteration)")
            res["speedup_vs_plain_fp64_gpu"] = round(value / rp["value"], 2)
            res["north_star_speedup_target"] = 10.0
    if not args.no_poly:
        note("polydisperse N=100k")
        res["polydisperse"] = run_poly(S)
    if not args.no_ensemble:
        note("ensemble")
        ens = run_ensemble(S, args.steps, args.warmup)
        res["ensemble"] = ens
        res["ensemble_step_latency_ms_p50"] = ens["step_latency_ms_p50"]
        res["ensemble_step_latency_ms_p99"] = ens["step_latency_ms_p99"]
    if not args.no_serve:
        note("serving requests")
        reqs = serve_requests(SERVE["horizon"])
        note("serving FP32-mixed (%d requests)" % len(reqs))
        sv32 = run_serving(S, S.SERVE_FP32, reqs)
        res["serving"] = dict(
            fp32_mixed=sv32,
            workload="Poisson arrivals at 50 req/s for %.0f s (deadline %.0f s): N log-uniform "
                     "[1000, 200000], phi U[0.05, 0.3], mesh nearest even 2^a3^b5^c to L/0.8, "
                     "p=6 rc=5 xi=0.896, 100 steps dt=1e-3, block Lanczos b=8 tol=1e-6, "
                     "mob-interval=100; admission by predicted footprint against min(288 GB, "
                     "free device memory)" % (SERVE["horizon"], SERVE["deadline"]))
        if not args.no_plain:
            note("serving plain FP64")
            res["serving"]["plain_fp64"] = run_serving(S, S.SERVE_PLAIN, reqs)
    if not args.no_err:
        note("FP32-vs-FP64 errors")
        e_mf, e_d, ew = displacement_error(S, O, pos, rdi, L)
        res["mf_err_fp32_vs_fp64"] = e_mf
        res["brownian_disp_err_fp32_vs_fp64"] = e_d
        res["ewald_sampled"] = ew
    if not args.no_cpu and O is not None:
        note("CPU baseline")
        cb = cpu_baseline(O, pos, rdi, L, r32["iters"])
        res["cpu_baseline"] = dict(value=round(cb["value"], 6), unit=cb["unit"], cores=cb["cores"],
                                   kind=cb["kind"], sample=cb["sample"], model=cb["model"],
                                   pieces_s=cb["pieces_s"])
        res["gpu_vs_cpu"] = round(value / cb["value"], 1)
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
