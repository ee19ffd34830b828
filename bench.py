#!/usr/bin/env python
"""BD steps/s of the StokesDT Brownian-dynamics step on one MI355X.

Workload (the single-system configuration the headline metric is quoted on):
N = 100 000 monodisperse spheres (a = 1) at volume fraction 0.3 in a
periodic cube (L = 111.77), seeded random-sequential-addition positions,
SPME mesh 128^3, B-spline order 6, real-space cutoff 5, Ewald xi = 0.896,
block Lanczos (block 8, tol 1e-6, FP64 Krylov basis and small algebra),
mobility interval 100, steric r0 = 2, k0 = 125, dt = 1e-3. A step is one
iteration of BDSimBox::Advance (src/simbox/simbox_bd.cc:296-366); the
timed region covers `--steps` steps (default 100, one full mobility
interval, so it contains exactly one Lanczos solve of all 100 columns).

Prints one JSON line: value = FP32-mixed steps/s; also the FP64 GPU path,
the FP32-vs-FP64 displacement error, per-stage device times (HIP events on
the launching stream), the roofline of the dominant kernel and the CPU
baseline (the FP64 C++/OpenMP oracle timed on a bounded sample).
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


def box_size(n, phi):
    return (4.0 * np.pi * n / (3.0 * phi)) ** (1.0 / 3.0)


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
    1 and with the Lanczos batch width of columns (2 repetitions each)."""
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


def run_gpu(S, pos, rdi, L, prec, steps, warmup, per_step):
    c = CFG
    bd = S.BDSim(pos, rdi, L, c["xi"], c["rmax"], c["K"], c["p"], prec, c["max_iters"], c["b"],
                 c["tol"], c["r0"], c["k0"], c["dt"], c["mob_interval"], c["seed"])
    bd.init()
    if warmup:
        bd.advance(warmup)
    bd.set_profiling(True)
    S.lib().sdt_device_synchronize()
    lat = []
    t0 = time.perf_counter()
    if per_step:
        for _ in range(steps):
            t = time.perf_counter()
            bd.advance(1)
            lat.append(time.perf_counter() - t)
    else:
        bd.advance(steps)
    S.lib().sdt_device_synchronize()
    dt = time.perf_counter() - t0
    prof = bd.profile()
    iters = bd.last_iters()
    bd.set_profiling(False)
    return dict(time_s=dt, lat=lat, prof=prof, iters=[int(x) for x in iters], pos=bd.positions())


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
    return e_mf, e_d


def cpu_baseline(O, pos, rdi, L, iters):
    """Times the FP64 C++/OpenMP oracle (a structural restatement of the
    reference: stored P/ind, 8-colour spread, per-RHS FFT loop, stored BSR,
    per-iteration eigensolve) on a bounded sample of the same workload and
    composes one mobility interval from the measured pieces."""
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
    O.lanczos(m, z, 2, c["b"], 1e-30)   # exactly two block iterations
    t_two = time.perf_counter() - t
    t_iter8 = t_two / 2.0
    # one mobility interval: update, every Lanczos group (iteration counts
    # of the GPU run, iter index + 1 block matvecs each), and M_I steps of
    # steric force + one single-RHS matvec
    mi = c["mob_interval"]
    groups = [c["b"]] * (mi // c["b"]) + ([mi % c["b"]] if mi % c["b"] else [])
    lz = sum((k + 1 if k >= 0 else c["max_iters"]) * t_iter8 * g / c["b"]
             for k, g in zip(iters, groups))
    t_interval = t_update + lz + mi * (t_steric + t_mv)
    sample = ("N=100000 workload pieces timed on the host: 1 mobility update, 1 single-RHS "
              "matvec, 1 steric force, 2 block-Lanczos iterations (b=8); one 100-step interval "
              "composed with the GPU run's Lanczos iteration counts")
    return dict(value=mi / t_interval, unit="steps/s", cores=O.num_threads(), kind="port",
                sample=sample, t_update_s=t_update, t_matvec_s=t_mv, t_steric_s=t_steric,
                t_lanczos_iter_b8_s=t_iter8)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 GPU path")
    ap.add_argument("--no-err", action="store_true", help="skip the FP32-vs-FP64 error check")
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

    import stokesdt_amd as S
    S.lib()
    c = CFG
    n = c["n"]
    L = box_size(n, c["phi"])
    pos = S.rsa_positions(n, L, 1.0, c["seed"])
    rdi = np.ones(n)

    r32 = run_gpu(S, pos, rdi, L, S.FP32, args.steps, args.warmup, per_step=True)
    value = args.steps / r32["time_s"]
    res = dict(metric=METRIC, value=round(value, 4), unit="steps/s", n_gpus=1, steps=args.steps,
               warmup=args.warmup, ms_per_step=round(1000.0 * r32["time_s"] / args.steps, 4),
               higher_is_better=True, scaling="weak", vs_baseline=None, dtype="fp32",
               data="synthetic (seeded RSA positions, Philox-4x32-10 Gaussian blocks)")
    res["config"] = {
        "workload": "single system N=100000 phi=0.3 (L=%.4f) mesh 128^3 p=6 rc=5 xi=0.896 "
                    "block-Lanczos b=8 tol=1e-6 mob-interval=100 dt=1e-3 steric r0=2 k0=125" % L,
        "precision": "FP32 spread/FFT/influence/interp/real-space with FP32 accumulation; "
                     "FP64 vectors, Krylov basis, Gram/QR/eigensolve",
        "parallelism": "single GPU",
    }
    lat = np.array(r32["lat"]) * 1000.0
    res["step_latency_ms_p50"] = round(float(np.percentile(lat, 50)), 4)
    res["step_latency_ms_p99"] = round(float(np.percentile(lat, 99)), 4)
    res["lanczos_iters_fp32"] = r32["iters"]
    stage = {k: dict(total_ms=round(v[0], 4), launches=v[1]) for k, v in r32["prof"].items()}
    for k, v in r32["prof"].items():
        if v[2]:
            stage[k]["columns"] = v[2]
    res["stage_device_ms"] = stage

    # roofline of the dominant stage (largest device time in the timed
    # region). Its launches there: one single-column matvec per step and,
    # when the window holds a mobility update, one fused matvec per block
    # iteration over the columns of the groups still iterating (the timer
    # records the columns of every launch; algorithmic and measured bytes are
    # affine in the launch width).
    cand = ["spread", "fft_xy_fwd", "fft_z_influence", "fft_xy_inv", "interp", "real_space"]
    dom = max(cand, key=lambda k: r32["prof"][k][0])
    tot_ms, launches, ncols = r32["prof"][dom]
    btot = lanczos_batch_cols()
    nnz = 0
    if dom == "real_space":
        m = S.MobSpme(n, rdi, L, c["xi"], c["rmax"], c["K"], c["p"], S.FP32)
        assert m.init()
        m.update(pos)
        nnz = m.real_nnz()
        del m
    if launches > 0 and tot_ms > 0:
        a0 = alg_bytes(dom, 0, n, c["K"], 4, nnz=nnz)
        a1 = alg_bytes(dom, 1, n, c["K"], 4, nnz=nnz) - a0
        tot_bytes = launches * a0 + a1 * ncols
        achieved = tot_bytes / (tot_ms * 1e-3) / 1e9
        traffic = None
        tr = read_traffic(dom)
        if 1 in tr and btot in tr:
            per_col = (tr[btot] - tr[1]) / (btot - 1)
            traffic = round((launches * (tr[1] - per_col) + per_col * ncols) / launches)
        res["roofline"] = dict(kernel=dom, bound="hbm", achieved=round(achieved, 2),
                               peak=HBM_PEAK_GBS, unit="GB/s",
                               frac=round(achieved / HBM_PEAK_GBS, 5), traffic=traffic,
                               launches=launches, columns=ncols,
                               avg_launch_ms=round(tot_ms / launches, 5),
                               alg_bytes_per_launch=round(tot_bytes / launches))
    else:
        res["roofline"] = dict(kernel=dom, bound="hbm", achieved=None, peak=HBM_PEAK_GBS,
                               unit="GB/s", frac=None, traffic=None,
                               note="no launches of %s recorded" % dom)
    O = None
    if not args.no_cpu or not args.no_err:
        try:
            from oracle import oracle as O
            O.lib()
        except Exception as exc:   # the checker is optional for the GPU figures
            O = None
            res["oracle_error"] = str(exc)
    if not args.no_fp64:
        r64 = run_gpu(S, pos, rdi, L, S.FP64, args.steps, args.warmup, per_step=False)
        v64 = args.steps / r64["time_s"]
        res["fp64_gpu_steps_per_s"] = round(v64, 4)
        res["speedup_fp32_vs_fp64_gpu"] = round(value / v64, 3)
        res["lanczos_iters_fp64"] = r64["iters"]
        d32 = r32["pos"] - pos
        d64 = r64["pos"] - pos
        res["trajectory_disp_err_fp32_vs_fp64"] = float(np.linalg.norm(d32 - d64) /
                                                        np.linalg.norm(d64))
    if not args.no_err:
        e_mf, e_d = displacement_error(S, O, pos, rdi, L)
        res["mf_err_fp32_vs_fp64"] = e_mf
        res["brownian_disp_err_fp32_vs_fp64"] = e_d
    if not args.no_cpu and O is not None:
        cb = cpu_baseline(O, pos, rdi, L, r32["iters"])
        res["cpu_baseline"] = dict(value=round(cb["value"], 6), unit=cb["unit"], cores=cb["cores"],
                                   kind=cb["kind"], sample=cb["sample"],
                                   pieces_s=dict(update=round(cb["t_update_s"], 3),
                                                 matvec=round(cb["t_matvec_s"], 3),
                                                 steric=round(cb["t_steric_s"], 3),
                                                 lanczos_iter_b8=round(cb["t_lanczos_iter_b8_s"], 3)))
        res["gpu_vs_cpu"] = round(value / cb["value"], 1)
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
