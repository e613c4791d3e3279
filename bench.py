#!/usr/bin/env python3
"""Benchmark of the EVPFFT Green-operator loop on one MI355X.

Workload (the 128^3 configuration): 128^3 periodic voxels, 512 periodic
Voronoi grains with uniform SO(3) texture, 2 % of the volume in spherical
pores (radius 3, zero stiffness), FCC Cu (C11/C12/C44 = 168.4/121.4/75.4 GPa,
12 {111}<110> systems, n = 10, gamma0 = 1/s, Voce tau0/tau1/theta0/theta1 =
16/10/200/20 MPa), uniaxial tension along z at 1/s with free lateral
stresses, dt = 1e-3 s, Green-loop tolerance 1e-6.  Inputs are written in the
reference's file formats by the seeded generator and read by the drop-in.

A step is one loading increment (prologue, Green loop to convergence,
epilogue).  `--warmup W` increments run on the job first, then the job is
reset to its post-setup state and increments 1..K are timed, so the timed
steps are the first K increments of the 50-increment run.  Both the mixed
(FP32 FFT/Green, FP64 Newton + FP64 correction sweeps) and the all-FP64 GPU
loops run; `value` is the mixed loop's voxel-iterations per second:
    N^3 * (Green iterations) / (wall time of the K increments).

Prints one JSON line (rank 0, N = 1 only).
"""
import argparse
import dataclasses
import json
import os
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "m-themes_amd"))

from pyevp import lib, synth  # noqa: E402

METRIC = "voxel-iters/s at 128³, mixed vs FP64 GPU at 1e-6 err; adjoint/forward time"

# FP64 operations of the per-voxel Newton kernel, counted by hand from
# newton.hip (an FMA counts as 2, a refined reciprocal square root as 10):
#   per crystal voxel and launch: 12 Schmid tensors from the rotation
#     (~55 each), three basis changes, |lambda|, 12 reciprocal CRSS, two
#     back-conversions                                                   ~800
#   per Newton iteration: rate law with dt d(edot)/d(sigma) added into the
#     packed Newton matrix (12 x ~67), S0 + S (21), strain mismatch and
#     residual (168), S0 R, C0 R and the merit gradient K C0 R (3 x 72),
#     6x6 Cholesky and the two triangular solves (~250), step control (~60)
#                                                                       ~1520
#   per line-search trial: rate law without derivative (12 x ~29), strain
#     mismatch, residual, merit function, trial point                    ~540
NEWTON_FLOPS_VOXEL = 800
NEWTON_FLOPS_ITER = 1520
NEWTON_FLOPS_TRIAL = 540
# The algorithmic count the roofline figure uses: the reference's own per-voxel
# work, counted from src/evolution.c:44-310 and src/constitutive.c:5-67 --
# about 0.5 kflop of set-up per voxel and launch (LU of C_gr, basis changes)
# and 4.7 kflop per Newton iteration (rate law with its derivative, the
# Jacobian and its pivoted-LU inverse, the step and one line-search trial),
# with the pow() calls not counted.  The kernel's own, smaller count above
# (symmetric Cholesky form, precomputed inverse) is reported beside it.
ALG_FLOPS_VOXEL = 500
ALG_FLOPS_ITER = 4700
# HBM bytes of the Newton kernel per crystal voxel and launch: reads sigma,
# strain, plastic strain (3 x 48), the rotation (72), CRSS (96: forward =
# backward), phase + orientation index (8); writes sigma and edot (2 x 48).
# (The one launch per increment with the shortcut off also writes gamdot,
# 96 B.)  Gas voxels: phase (4), sigma/edot writes (96), sigma re-read (48).
NEWTON_BYTES_CRYSTAL = 3 * 48 + 72 + 96 + 8 + 2 * 48
NEWTON_BYTES_GAS = 4 + 96 + 48
# k_newton's HBM traffic per launch on the 128^3 job (launches with the
# shortcut on), measured: rocprofv3 FETCH_SIZE x 1024 x 2 + WRITE_SIZE x 1024
# (FETCH_SIZE reads half and WRITE_SIZE all of a plain copy kernel's bytes);
# 684.5 MB read + 205.5 MB written against 861 MB algorithmic.
NEWTON_TRAFFIC_128 = 8.900e8
PEAK_HBM_GBS = 8000.0
PEAK_FP64_TFS = 78.6


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def run_gpu(inp, precision, steps, warmup, law=lib.LAW_POWER, loop_batch=None, timers=(0,), every=8):
    """Timed increments; HIP-event timers `timers` (default: the Newton kernel's
    alone) in one Green iteration of every `every` -- each event pair costs the
    stream a few microseconds between kernels."""
    job = lib.GpuJob(inp, precision, law=law)
    if loop_batch is not None:
        job.set_loop_batch(loop_batch)
    if warmup > 0:
        job.run(warmup)
        job.reset()
    job.set_timing(True, groups=timers, every=every)
    job.synchronize()
    t0 = time.perf_counter()
    it, stats = job.run(steps)
    job.synchronize()
    wall = time.perf_counter() - t0
    out = dict(wall=wall, iters=[int(i) for i in it], stats=stats,
               timing={k: job.timing(k) for k in range(6)}, macro=job.macro(),
               sig=job.state()["Sig"])
    job.close()
    return out


def convergence(r, cfg):
    """Final Err_s / Err_e of every timed increment and whether the loop ended on
    IterMax instead of the tolerance (src/evolution.c:1354)."""
    es = [s["err_s"] for s in r["stats"]]
    ee = [s["err_e"] for s in r["stats"]]
    hit = [max(abs(a), abs(b)) > cfg.err for a, b in zip(es, ee)]
    return {"err_s": es, "err_e": ee, "tol": cfg.err, "iter_max": cfg.iter_max,
            "hit_iter_max": hit, "converged_increments": int(len(hit) - sum(hit))}


def run_forced(inp, precision, iters, warmup):
    """The same increments forced to given iteration counts (equal-count comparison)."""
    job = lib.GpuJob(inp, precision)
    if warmup > 0:
        job.run(warmup)
        job.reset()
    for k in iters:
        job.run(1, forced=int(k))
    out = dict(sig=job.state()["Sig"], macro=job.macro())
    job.close()
    return out


def run_adjoint(n, increments, ck):
    """Calibration workload: forward with checkpoints, then the reverse pass
    for L = sum_k (Sigma_33,k - target_k)^2 with the target from a run at
    parameters +5 %."""
    cfg = synth.config_c4()
    cfg.n = n
    cfg.grains = max(8, 128 * n ** 3 // 64 ** 3)
    cfg.increments = increments
    d = tempfile.mkdtemp(prefix="evpadj_")
    inp = synth.write_job(d, cfg)
    job = lib.GpuJob(inp, lib.FP64)
    p0 = job.get_params(0)
    job.set_params(0, p0 * 1.05)
    target, _ = job.adjoint_record(increments, ck)
    job.set_params(0, p0)
    job.synchronize()
    t0 = time.perf_counter()
    curve, stats = job.adjoint_record(increments, ck)
    job.synchronize()
    t_fwd = time.perf_counter() - t0
    seeds = np.zeros_like(curve)
    seeds[:, 2] = 2.0 * (curve[:, 2] - target[:, 2])
    t0 = time.perf_counter()
    g, st = job.adjoint_run(seeds, cg_tol=1e-10, cg_max=4000)
    t_adj = time.perf_counter() - t0
    job.close()
    return dict(workload=f"EVPFFT {n}^3, {cfg.grains} grains, {increments} increments, "
                         f"checkpoints every {ck}",
                forward_s=t_fwd, adjoint_s=t_adj, ratio=t_adj / t_fwd,
                replay_s=st["replay_seconds"], cg_iters=st["cg_iters"],
                cg_max_iters=st["cg_max_iters"], cg_rel_resid=st["cg_rel_resid"],
                green_iters=[s["iters"] for s in stats],
                loss=float(np.sum((curve[:, 2] - target[:, 2]) ** 2)),
                grad=dict(zip(lib.PARAM_NAMES, [float(x) for x in g[0]])))


def run_pf(steps, cpu_steps):
    """The DRX configuration's phase-field step: PF_XRX on 128^3 voxels with
    256 grains and 40 nuclei (squared radius 5), `steps` substeps from the
    one-entry start; device time of the substeps (HIP events on the step's
    stream)."""
    from pyevp import pf as gpf
    shape = (128, 128, 128)
    gid, rho, prm = synth.pf_inputs(shape, 256, 40, 0x5EED0005, radius2=5)
    n = int(np.prod(shape))
    p = gpf.PhaseField(shape, 60000)
    try:
        p.pf_xrx(gid, rho, 1, prm, np.zeros(60000))  # warm-up
        t0 = time.perf_counter()
        gn, _ = p.pf_xrx(gid, rho, steps, prm, np.zeros(60000))
        wall = time.perf_counter() - t0
        st = p.stats()
    finally:
        p.close()
    ms = st["ms"]
    # roofline of the voxel-update kernel: algorithmic bytes per launch = per
    # voxel its own list (4 B count + 12 B per entry), its candidate grains
    # (4 B + 4 B per grain), its turn record (12 B), the averages of its
    # entries on the update branch (12 B per entry), its new list (4 B + 12 B
    # per entry) and its reset flag (4 B); list lengths and candidates at
    # their means after the run, the update-branch share from the counts
    lists = None
    p = gpf.PhaseField(shape, 60000)
    try:
        p.pf_xrx(gid, rho, steps, prm, np.zeros(60000))
        lists = p.lists()
    finally:
        p.close()
    nbar = float(lists["n"].mean())
    f_upd = st["updated"] / (n * steps)
    per_voxel = 4 + 12 * nbar + 4 + 4 * nbar + 12 + 12 * nbar * f_upd + 4 + 12 * nbar + 4
    launch_ms = st["ms_update"] / max(st["update_launches"], 1)
    gbs = n * per_voxel / (launch_ms * 1e-3) / 1e9
    roof = {"bound": "hbm", "kernel": "k_update", "achieved": gbs, "peak": PEAK_HBM_GBS, "unit": "GB/s",
            "frac": gbs / PEAK_HBM_GBS, "traffic": None, "avg_launch_ms": launch_ms,
            "bytes_per_voxel": per_voxel, "mean_list_length": nbar, "update_branch_share": f_upd,
            "note": "integer and branch work on lists of a few entries: bound by the latency of the "
                    "neighbour-list gathers, not by HBM bandwidth; k_avg (7x7x7 averages through LDS) "
                    f"{st['ms_avg'] / max(st['update_launches'], 1):.3f} ms per launch"}
    rec = {"workload": f"PF_XRX 128^3, 256 Voronoi grains, 40 nuclei, {steps} substeps",
           "value": n * steps / (ms * 1e-3), "unit": "voxel-substeps/s", "ms_per_substep": ms / steps,
           "wall_s_with_transfers": wall, "stats": st, "roofline": roof,
           "grains_changed": int((gn != gid).sum())}
    log(f"bench: pf_xrx: {steps} substeps, {ms / steps:.3f} ms per substep, {rec['value']:.4g} voxel-substeps/s")
    if cpu_steps > 0:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        t0 = time.perf_counter()
        r = O.pf_xrx(gid, rho, cpu_steps, prm, np.zeros(60000))
        dt = time.perf_counter() - t0
        rec["cpu_baseline"] = {"value": n * cpu_steps / dt, "unit": "voxel-substeps/s", "cores": 1,
                               "kind": "port",
                               "sample": f"the same grid, first {cpu_steps} substeps ({dt:.1f} s)",
                               "status": r["status"]}
        log(f"bench: pf_xrx cpu: {rec['cpu_baseline']}")
    return rec


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(inp, n_iter_1, n_iter_all):
    """The oracle (C restatement, OpenMP over voxels) on a bounded sample of the
    same workload: the first Green iterations of increment 1, once on one
    thread (`value`) and once on every host thread the process may use."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    n_all = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    n_all = min(n_all, int(os.environ.get("OMP_NUM_THREADS", n_all)))
    out = {}
    for threads, n_iter in ((1, n_iter_1), (n_all, n_iter_all)):
        if n_iter <= 0:
            continue
        O.set_threads(threads)
        J = O.Job(inp)
        J.run(1, forced=[n_iter])
        m = J.macro()
        n = int(np.prod(J.dims))
        out[threads] = (n * n_iter / m["green_time"], n_iter, m["green_time"])
        del J
    O.set_threads(n_all)
    one = out.get(1)
    allc = out.get(n_all)
    base = one if one is not None else allc
    res = dict(value=base[0], unit="voxel-iters/s", cores=1 if one is not None else n_all, kind="port",
               cpu_model=cpu_model(),
               sample=f"128^3 job, increment 1, first {base[1]} Green iterations on "
                      f"{1 if one is not None else n_all} thread(s) ({base[2]:.1f} s)")
    if allc is not None and one is not None:
        res["all_cores"] = dict(value=allc[0], cores=n_all,
                                sample=f"first {allc[1]} Green iterations ({allc[2]:.1f} s)",
                                speedup_vs_1=allc[0] / one[0])
    return res


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=4, help="timed increments")
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=128,
                    help="grid edge (default: the 128^3 job; 256: the 256^3 job of 4096 grains)")
    ap.add_argument("--precision", choices=["both", "mixed", "fp64"], default="both")
    ap.add_argument("--cpu-iters", type=int, default=1,
                    help="Green iterations of the one-thread CPU baseline sample (0: skip)")
    ap.add_argument("--cpu-iters-all", type=int, default=3,
                    help="Green iterations of the all-threads CPU baseline sample (0: skip)")
    ap.add_argument("--adj-n", type=int, default=64, help="grid edge of the adjoint workload")
    ap.add_argument("--adj-increments", type=int, default=20, help="0: skip the adjoint")
    ap.add_argument("--adj-checkpoint", type=int, default=5)
    ap.add_argument("--dd-steps", type=int, default=2,
                    help="increments of the same job under the dislocation-density law, "
                         "reported beside the headline (0: skip)")
    ap.add_argument("--pf-steps", type=int, default=10,
                    help="phase-field substeps of the recrystallisation step at 128^3, 256 grains, "
                         "40 nuclei, reported beside the headline (0: skip)")
    ap.add_argument("--pf-cpu-steps", type=int, default=2,
                    help="substeps of the oracle's phase-field CPU sample (0: skip)")
    ap.add_argument("--pf-only", action="store_true",
                    help="profiling mode: only the phase-field step, then exit")
    ap.add_argument("--law", choices=["power", "dd"], default="power",
                    help="law of the profiling mode's job")
    ap.add_argument("--profile-iters", type=int, default=0,
                    help="profiling mode: one increment forced to this many Green iterations "
                         "per precision, then exit (for rocprofv3 counter runs)")
    ap.add_argument("--timers", choices=["newton", "all"], default="newton",
                    help="HIP-event timers in the timed region: the Newton kernel's alone "
                         "(default), or every kernel group's (adds event overhead)")
    ap.add_argument("--timer-every", type=int, default=8,
                    help="the timers record one Green iteration in this many")
    ap.add_argument("--loop-batch", type=int, default=None,
                    help="Green iterations per batch queued ahead of the device (0: the host "
                         "waits for every iteration; default: the library's, 4)")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="profiling mode: also run one plain device copy of this size (the "
                         "reference kernel for scaling FETCH_SIZE / WRITE_SIZE)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: --gpus {args.gpus} is not supported: this build runs the whole grid "
              "on one GPU (no multi-GPU path); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    if args.pf_only:
        print(json.dumps({"pf_xrx": run_pf(max(args.pf_steps, 1), args.pf_cpu_steps)}), flush=True)
        return
    if args.n == 256:
        cfg = synth.config_c3()     # 256^3, 4096 grains, no pores
    else:
        cfg = synth.config_c2()
        if args.n != 128:
            cfg.n = args.n
            cfg.grains = max(8, 512 * args.n ** 3 // 128 ** 3)
    work = tempfile.mkdtemp(prefix="evpbench_")
    t = time.perf_counter()
    inp = synth.write_job(work, cfg)
    log(f"bench: wrote {cfg.n}^3 job ({cfg.grains} grains, {cfg.pore_fraction:.0%} pores) "
        f"in {time.perf_counter() - t:.1f} s")
    N = cfg.n ** 3
    # the same job under the dislocation-density law
    inp_dd = None
    if args.dd_steps > 0 or args.law == "dd":
        cfg_dd = dataclasses.replace(cfg, dd=synth.DDLaw())
        inp_dd = synth.write_job(tempfile.mkdtemp(prefix="evpbench_dd_"), cfg_dd)
    if args.profile_iters > 0:
        for name, prec in (("fp64", lib.FP64), ("mixed", lib.MIXED)):
            if args.precision not in ("both", name):
                continue
            if args.law == "dd":
                job = lib.GpuJob(inp_dd, prec, law=lib.LAW_DD)
            else:
                job = lib.GpuJob(inp, prec)
            t0 = time.perf_counter()
            job.run(1, forced=args.profile_iters)
            job.synchronize()
            log(f"bench: profile {name}: {args.profile_iters} iterations in {time.perf_counter() - t0:.2f} s")
            job.close()
        if args.copy_probe_mib > 0:
            job = lib.GpuJob(inp, lib.FP64)
            n = args.copy_probe_mib * (1 << 20) // 8
            gbs = job.copy_probe(n, 1)
            log(f"bench: copy probe: k_copy of {args.copy_probe_mib} MiB read + "
                f"{args.copy_probe_mib} MiB written, {gbs:.0f} GB/s")
            job.close()
        print(json.dumps({"profile_iters": args.profile_iters, "grid": cfg.n}), flush=True)
        return
    res = {}
    for name, prec in (("fp64", lib.FP64), ("mixed", lib.MIXED)):
        if args.precision not in ("both", name):
            continue
        log(f"bench: {name}: warmup {args.warmup}, timing {args.steps} increments")
        r = run_gpu(inp, prec, args.steps, args.warmup, loop_batch=args.loop_batch,
                    timers=(0,) if args.timers == "newton" else tuple(range(6)), every=args.timer_every)
        its = sum(r["iters"])
        r["vips"] = N * its / r["wall"]
        log(f"bench: {name}: iterations {r['iters']} ({sum(s['fp64_iters'] for s in r['stats'])} "
            f"FP64), {r['wall']:.2f} s, {r['vips']:.4g} voxel-iters/s")
        res[name] = r
    main_key = "mixed" if "mixed" in res else "fp64"
    R = res[main_key]

    # roofline of the dominant kernel: the per-voxel Newton update.  Its HIP
    # events time a sample of the launches (one Green iteration in
    # --timer-every); the work per launch is the average over all of them
    # (one launch per Green iteration, plus the replay closing each increment)
    ms, sampled = R["timing"][0]
    avg_ms = ms / max(sampled, 1)
    launches = sum(R["iters"]) + len(R["iters"])
    nit = sum(s["newton_iters"] for s in R["stats"])
    ntr = sum(s["ls_trials"] for s in R["stats"])
    j = lib.GpuJob(inp, lib.FP64)
    g, p = j.ids()
    j.close()
    n_cryst = int(np.sum(p == 1))
    n_gas = N - n_cryst
    flops = (ALG_FLOPS_VOXEL * n_cryst * launches + ALG_FLOPS_ITER * nit) / launches
    own = (NEWTON_FLOPS_VOXEL * n_cryst * launches + NEWTON_FLOPS_ITER * nit + NEWTON_FLOPS_TRIAL * ntr) / launches
    bytes_ = NEWTON_BYTES_CRYSTAL * n_cryst + NEWTON_BYTES_GAS * n_gas
    t_s = avg_ms * 1e-3
    tf = flops / t_s / 1e12
    tf_own = own / t_s / 1e12
    gbs = bytes_ / t_s / 1e9
    # measured HBM bytes per launch of k_newton on this job (rocprofv3 FETCH_SIZE
    # and WRITE_SIZE scaled by a copy kernel's readings), 128^3 workload only
    traffic = NEWTON_TRAFFIC_128 if cfg.n == 128 else None
    per = dict(newton_iters_per_voxel_launch=nit / max(n_cryst * launches, 1),
               trials_per_voxel_launch=ntr / max(n_cryst * launches, 1),
               flops_per_launch=flops, bytes_per_launch=bytes_, launches=launches,
               timed_launches=sampled,
               kernel_count=dict(tflops=tf_own, frac=tf_own / PEAK_FP64_TFS, flops_per_launch=own))
    if tf / PEAK_FP64_TFS >= gbs / PEAK_HBM_GBS:
        roof = dict(bound="fp64-valu", achieved=tf, peak=PEAK_FP64_TFS, unit="TFLOP/s",
                    frac=tf / PEAK_FP64_TFS, traffic=traffic, kernel="k_newton",
                    avg_launch_ms=avg_ms, hbm_gbs=gbs, **per,
                    note="bound by the FP64 vector ALU (no MFMA instruction is used); peak is "
                         "the MI355X FP64 vector figure, 78.6 TFLOP/s; achieved counts the "
                         "reference's work per voxel (0.5 kflop + 4.7 kflop per Newton iteration), "
                         "kernel_count the kernel's own smaller count; avg_launch_ms from HIP "
                         "events on the launches of one Green iteration in "
                         f"{args.timer_every} over the timed region")
    else:
        roof = dict(bound="hbm", achieved=gbs, peak=PEAK_HBM_GBS, unit="GB/s",
                    frac=gbs / PEAK_HBM_GBS, traffic=traffic, kernel="k_newton",
                    avg_launch_ms=avg_ms, tflops=tf, **per)
    kern = {}
    names = ["newton", "fft_forward", "green", "fft_inverse_update", "increment_pro_epilogue",
             "anchor_refresh"]
    for k, nm in enumerate(names):
        kms, kn = R["timing"][k]
        if kn:
            kern[nm] = dict(ms_total=round(kms, 3), launches=kn, ms_avg=round(kms / max(kn, 1), 4))

    out = {
        "metric": METRIC,
        "value": R["vips"],
        "unit": "voxel-iters/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * R["wall"] / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64" if main_key == "mixed" else "f64",
        "data": "synthetic",
        "config": {"workload": f"EVPFFT {cfg.n}^3, {cfg.grains} Voronoi grains, "
                               f"{cfg.pore_fraction:.0%} pores, "
                               f"increments 1..{args.steps} of {cfg.increments}",
                   "grid": cfg.n, "grains": cfg.grains, "pores": cfg.pore_fraction,
                   "precision": main_key, "green_tol": cfg.err, "seed": hex(cfg.seed)},
        "green_iters": R["iters"],
        "convergence": convergence(R, cfg),
        "kernels_ms": kern,
        "roofline": roof,
    }
    if "fp64" in res:
        out["fp64"] = {"value": res["fp64"]["vips"], "ms_per_step": 1e3 * res["fp64"]["wall"] / args.steps,
                       "green_iters": res["fp64"]["iters"],
                       "convergence": convergence(res["fp64"], cfg)}
    if "fp64" in res and "mixed" in res:
        a, b = res["fp64"]["sig"], res["mixed"]["sig"]
        log("bench: mixed forced to the FP64 iteration counts (precision error alone)")
        f = run_forced(inp, lib.MIXED, res["fp64"]["iters"], args.warmup)

        def szz(m):
            return m["sig_avg"][2, 2]
        out["mixed_vs_fp64"] = {
            "speedup_voxel_iters": res["mixed"]["vips"] / res["fp64"]["vips"],
            "speedup_time": res["fp64"]["wall"] / res["mixed"]["wall"],
            "stress_max_rel_err_equal_iters": float(np.max(np.abs(f["sig"] - a)) / np.max(np.abs(a))),
            "sig_zz_rel_err_equal_iters": float(abs(szz(f["macro"]) - szz(res["fp64"]["macro"])) /
                                                abs(szz(res["fp64"]["macro"]))),
            "stress_max_rel_err_own_convergence": float(np.max(np.abs(a - b)) / np.max(np.abs(a))),
            "sig_zz_rel_err_own_convergence": float(abs(szz(res["mixed"]["macro"]) -
                                                        szz(res["fp64"]["macro"])) /
                                                    abs(szz(res["fp64"]["macro"]))),
        }
    if args.dd_steps > 0:
        log(f"bench: dislocation-density law, {main_key}: timing {args.dd_steps} increments")
        prec = lib.MIXED if main_key == "mixed" else lib.FP64
        try:
            r = run_gpu(inp_dd, prec, args.dd_steps, 0, law=lib.LAW_DD)
        except lib.EvpError as e:
            r = None
            out["dd_law"] = {"error": str(e)}
            log(f"bench: dd: {e}")
    if args.dd_steps > 0 and r is not None:
        vips = N * sum(r["iters"]) / r["wall"]
        log(f"bench: dd: iterations {r['iters']}, {r['wall']:.2f} s, {vips:.4g} voxel-iters/s")
        dk = {}
        for k, nm in enumerate(names):
            kms, kn = r["timing"][k]
            if kn:
                dk[nm] = dict(ms_total=round(kms, 3), launches=kn, ms_avg=round(kms / max(kn, 1), 4))
        out["dd_law"] = {"value": vips, "unit": "voxel-iters/s", "precision": main_key,
                         "steps": args.dd_steps, "warmup": 0,
                         "ms_per_step": 1e3 * r["wall"] / args.dd_steps,
                         "green_iters": r["iters"], "kernels_ms": dk,
                         "workload": f"the same {cfg.n}^3 job under StrainRate_Orowan + "
                                     "DislocationEvolution, increments 1.."
                                     f"{args.dd_steps}"}
    if args.pf_steps > 0:
        out["pf_xrx"] = run_pf(args.pf_steps, args.pf_cpu_steps)
    if args.adj_increments > 0:
        log(f"bench: adjoint: {args.adj_n}^3, {args.adj_increments} increments")
        out["adjoint"] = run_adjoint(args.adj_n, args.adj_increments, args.adj_checkpoint)
        log(f"bench: adjoint/forward = {out['adjoint']['ratio']:.3f}")
    if args.cpu_iters > 0 or args.cpu_iters_all > 0:
        log(f"bench: cpu baseline: oracle, {args.cpu_iters} Green iteration(s) on one thread, "
            f"{args.cpu_iters_all} on all")
        out["cpu_baseline"] = cpu_baseline(inp, args.cpu_iters, args.cpu_iters_all)
        log(f"bench: cpu baseline: {out['cpu_baseline']}")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
