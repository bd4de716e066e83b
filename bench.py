#!/usr/bin/env python3
"""Benchmark of the SPM time step (KAPSEL Time_evolution_hydro) on one MI355X.

Workload: 256^3 periodic grid, dx = 1, rho = eta = 1, dt = 0.05, 4000 spheres
of radius a = 5 with interface width xi = 2 (volume fraction 0.128), placed
uniformly at random (seed 12345) without overlap closer than 1.05 * 2a, WCA
repulsion (sigma = 2a, epsilon = 1, cutoff 2^(1/6) sigma), thermal noise
kBT = 0.1 from MT19937 seeded 12345, rotation on, FIX_CELL on.  Inputs are
resident in HBM before the timed region; one step is one full
Time_evolution_hydro.

Prints one JSON line: time-steps/s (whole job), the roofline of the dominant
kernel family measured with HIP events on the library's stream, and the CPU
oracle timed on a bounded sample of the same workload on this host.

--shear-rate R > 0 runs the same box under Lees-Edwards shear instead
(Time_evolution_hydro_OBL, Shear_rate R, literal FP64 schedule).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

from kapsel_old_amd import abi  # noqa: E402

FAMILIES = ["fft_z_r2c", "fft_z_c2r", "fft_y", "fft_x", "spectral", "realspace", "make_phi", "f_hydro",
            "particle"]
HBM_PEAK_GBS = 8000.0


def workload_params(n, np_, precision=abi.FP64):
    return abi.make_params(n, a=5.0, a_xi=2.0, dt=0.05, kbt=0.1, n_particles=np_, lj_truncate=abi.LJ_WCA,
                           lj_powers=abi.LJ_12_6, epsilon=1.0, rotation=1, fix_cell=(1, 1, 1), seed=12345,
                           precision=precision)


def workload_particles(n, np_):
    p = abi.new_particles(np_)
    pos = abi.random_positions(np_, (n, n, n), 1.05 * 2 * 5.0, seed=12345)
    p["x"] = pos
    p["x_nopbc"] = pos
    return p


def algorithmic_step_bytes(n, np_, stencil, w=8, n_fft=20):
    """B_step = N_FFT (8N + 16 NX NY (NZ/2+1)) (w/8) + Np S 11 w: each 3-D FFT
    reads its input and writes its output once, the gather reads 3 values and
    the scatter read-modify-writes 4 per stencil point."""
    nn = n ** 3
    return n_fft * (8 * nn + 16 * n * n * (n // 2 + 1)) * (w / 8) + np_ * stencil * 11 * w


def gpu_run(n, np_, steps, warmup, profile_steps, schedule, precision, shear_rate=0.0):
    from kapsel_old_amd.spm import SPM
    g = SPM(workload_params(n, np_, precision))
    g.set_schedule(schedule)
    g.put_particles(workload_particles(n, np_))
    g.init_fluid()
    if shear_rate > 0:
        g.set_lees_edwards(shear_rate)
    g.step(warmup)
    g.synchronize()
    print("bench: warmup done (%d steps)" % warmup, file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    g.step(steps)
    g.synchronize()
    t1 = time.perf_counter()
    elapsed = t1 - t0
    print("bench: timed %d steps in %.3f s" % (steps, elapsed), file=sys.stderr, flush=True)
    g.check_finite()
    # kernel durations: HIP events around every launch of the same step loop
    g.timing_enable(True)
    g.timing_read()
    t2 = time.perf_counter()
    g.step(profile_steps)
    g.synchronize()
    t3 = time.perf_counter()
    ms, by, cnt = g.timing_read()
    g.timing_enable(False)
    stencil = g.stencil_size()
    g.close()
    return elapsed, (t3 - t2), ms, by, cnt, stencil


def copy_probe(mbytes, reps=5):
    """Device-to-device copies of a known size through the HIP runtime (a
    plain copy kernel), used to scale rocprof's FETCH_SIZE / WRITE_SIZE."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    n = int(mbytes * (1 << 20))
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(n)) == 0 and hip.hipMalloc(C.byref(b), C.c_size_t(n)) == 0
    hip.hipMemset(a, 1, C.c_size_t(n))
    t0 = time.perf_counter()
    for _ in range(reps):
        assert hip.hipMemcpy(b, a, C.c_size_t(n), 3) == 0  # hipMemcpyDeviceToDevice
    hip.hipDeviceSynchronize()
    t1 = time.perf_counter()
    hip.hipFree(a)
    hip.hipFree(b)
    print(json.dumps({"copy_probe_bytes": n, "reps": reps, "GBps_read_plus_write": 2 * n * reps / (t1 - t0) / 1e9}),
          flush=True)


def load_traffic(schedule, precision, n, np_):
    """HBM bytes per launch of each kernel family from the committed
    rocprofv3 --pmc summary of this workload (FETCH_SIZE and WRITE_SIZE per
    dispatch, scaled by what the counters read for a plain copy of known
    size).  None when no summary for this configuration is committed."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic_%s_%s_%dcube_%d.txt" % (schedule, precision, n, np_))
    if not os.path.exists(path):
        return None, None
    out = {}
    with open(path) as fh:
        for line in fh:
            f = line.split()
            if len(f) == 4 and f[0] == "family":
                out[f[1]] = (float(f[2]), float(f[3]))
    return out, os.path.relpath(path, ROOT)


def cpu_baseline(n, np_, max_steps=2, budget_s=40.0, shear_rate=0.0):
    """The CPU oracle (single-threaded FP64 restatement of the reference
    step) on the same workload: init untimed, then up to max_steps steps."""
    from oracle.oracle import Oracle
    o = Oracle(workload_params(n, np_))
    o.put_particles(workload_particles(n, np_))
    o.init_fluid()
    if shear_rate > 0:
        o.set_lees_edwards(shear_rate)
    done = 0
    t0 = time.perf_counter()
    while done < max_steps:
        o.step(1)
        done += 1
        if time.perf_counter() - t0 > budget_s / max_steps * done and done >= 1:
            if time.perf_counter() - t0 > budget_s:
                break
    t1 = time.perf_counter()
    o.close()
    return done / (t1 - t0), done, t1 - t0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--profile-steps", type=int, default=10)
    ap.add_argument("--grid", type=int, default=256)
    ap.add_argument("--particles", type=int, default=4000)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--schedule", choices=["fused", "literal"], default="fused")
    ap.add_argument("--precision", choices=["fp64", "fp32"], default="fp64")
    ap.add_argument("--copy-probe", type=float, default=0.0, metavar="MB")
    ap.add_argument("--shear-rate", type=float, default=0.0)
    args = ap.parse_args()
    if args.copy_probe > 0:
        copy_probe(args.copy_probe)
        return
    if args.gpus != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path (one box per GPU); "
              "use --gpus 1", file=sys.stderr)
        sys.exit(2)
    n, np_ = args.grid, args.particles
    sheared = args.shear_rate > 0
    if sheared:  # the sheared step has one schedule and runs in FP64
        args.schedule, args.precision = "literal", "fp64"
    schedule = 0 if args.schedule == "fused" else 1
    precision = abi.FP64 if args.precision == "fp64" else abi.FP32
    elapsed, prof_elapsed, ms, by, cnt, stencil = gpu_run(n, np_, args.steps, args.warmup, args.profile_steps,
                                                          schedule, precision, args.shear_rate)
    wbytes = 8 if precision == abi.FP64 else 4
    value = args.steps / elapsed
    dom = int(np.argmax(ms))
    per_launch_ms = ms[dom] / max(cnt[dom], 1)
    per_launch_bytes = by[dom] / max(cnt[dom], 1)
    achieved = per_launch_bytes / (per_launch_ms * 1e-3) / 1e9
    b_step = algorithmic_step_bytes(n, np_, stencil, w=wbytes)
    traffic, traffic_src = (None, None) if sheared else load_traffic(args.schedule, args.precision, n, np_)
    fam_traffic = traffic.get(FAMILIES[dom]) if traffic else None
    families = {FAMILIES[f]: {"ms_per_step": ms[f] / args.profile_steps, "launches_per_step":
                              cnt[f] / args.profile_steps,
                              "GBps": (by[f] / (ms[f] * 1e-3) / 1e9) if ms[f] > 0 else None}
                for f in range(len(FAMILIES))}
    out = {
        "metric": "time-steps/s at 256^3 w/ 4000 spheres" + (" under Lees-Edwards shear" if sheared else ""),
        "value": value,
        "unit": "time-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if precision == abi.FP64 else "f32 fields, f64 particles",
        "data": "synthetic",
        "config": {"workload": "SPM time step (%s), %d^3 grid, %d spheres a=5 xi=2, "
                               "WCA, kBT=0.1, dt=0.05, %s fields%s"
                               % ("Time_evolution_hydro_OBL" if sheared else "Time_evolution_hydro", n, np_,
                                  args.precision.upper(),
                                  (", Lees-Edwards Shear_rate %g" % args.shear_rate) if sheared else ""),
                   "grid": [n, n, n], "n_particles": np_, "stencil_points": stencil, "precision": args.precision,
                   "schedule": args.schedule},
        "hbm_TBps_algorithmic": b_step * value / 1e12,
        "step_floor_ms_at_8TBps": b_step / 8e12 * 1e3,
        "roofline": {"bound": "hbm", "kernel": FAMILIES[dom], "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": (fam_traffic[0] + fam_traffic[1]) if fam_traffic else None,
                     "traffic_fetch": fam_traffic[0] if fam_traffic else None,
                     "traffic_write": fam_traffic[1] if fam_traffic else None,
                     "traffic_source": traffic_src if fam_traffic else None,
                     "bytes_per_launch": per_launch_bytes, "us_per_launch": per_launch_ms * 1e3},
        "kernel_families": families,
        "profiled_ms_per_step": 1e3 * prof_elapsed / args.profile_steps,
    }
    if not args.no_cpu_baseline and precision == abi.FP64:
        v, done, secs = cpu_baseline(n, np_, shear_rate=args.shear_rate)
        out["cpu_baseline"] = {"value": v, "unit": "time-steps/s", "cores": 1, "kind": "port",
                               "sample": "%d step(s) of the same %d^3/%d workload after an untimed init, "
                                         "%.1f s, oracle/liboracle.so single-threaded" % (done, n, np_, secs)}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
