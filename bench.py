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


def cpu_baseline(n, np_, samples=3, shear_rate=0.0, threads=(1,)):
    """The CPU oracle (FP64 restatement of the reference step) on the same
    workload: init and one step untimed, then `samples` single steps timed
    for each thread count; returns {threads: (median steps/s, seconds per
    sample list)}."""
    from oracle.oracle import Oracle, set_threads
    o = Oracle(workload_params(n, np_))
    o.put_particles(workload_particles(n, np_))
    o.init_fluid()
    if shear_rate > 0:
        o.set_lees_edwards(shear_rate)
    out = {}
    for th in threads:
        set_threads(th)
        o.step(1)  # untimed: first touch with this thread count
        secs = []
        for _ in range(samples):
            t0 = time.perf_counter()
            o.step(1)
            secs.append(time.perf_counter() - t0)
        out[th] = (1.0 / float(np.median(secs)), secs)
    o.close()
    return out


def one_step_calls(n, np_, steps, precision=abi.FP64):
    """Steps per second when every step is its own spm_step(1) call (the
    fused schedule's rebuild of u, phi and up is deferred to the first call
    that reads them, so a caller stepping one step per call pays it only
    when it reads the fields)."""
    from kapsel_old_amd.spm import SPM
    g = SPM(workload_params(n, np_, precision))
    g.put_particles(workload_particles(n, np_))
    g.init_fluid()
    for _ in range(3):
        g.step(1)
    g.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        g.step(1)
    g.synchronize()
    t1 = time.perf_counter()
    g.close()
    return steps / (t1 - t0)


def closure_line():
    """The learned closure at the coarse/fine sizes of its report: live
    steps/s of the FP32 64^3 step (512 spheres a = 3) with a network attached
    and of the FP64 256^3 step with the same spheres at a = 12, and the drag
    errors of the committed report (profiles/closure_report_64cube_vs_256cube.json)."""
    from kapsel_old_amd import closure as cl
    from kapsel_old_amd.spm import Closure
    pos, _, _ = cl.random_suspension(512, 64, 3.0, None, 900)
    net = Closure(1024, seed=1)
    coarse = cl.free_run_speed(64, 3.0, pos, 0.05, 1000, abi.FP32, closure=net)
    fine = cl.free_run_speed(256, 12.0, pos * 4, 0.8, 100, abi.FP64)
    coarse_g = cl.free_run_speed(64, 3.0, pos, 0.05, 1000, abi.FP32, closure=net, graph=True)
    fine_g = cl.free_run_speed(256, 12.0, pos * 4, 0.8, 100, abi.FP64, graph=True)
    out = {"workload": "512 spheres, volume fraction 0.22: FP32 64^3 a=3 + closure vs FP64 256^3 a=12",
           "steps_per_s_fp32_coarse_closure": coarse, "steps_per_s_fp64_fine": fine, "speedup": coarse / fine,
           "graph_replayed": {"steps_per_s_fp32_coarse_closure": coarse_g, "steps_per_s_fp64_fine": fine_g,
                              "speedup": coarse_g / fine_g,
                              "note": "both steps replayed from a captured graph (one-box spm_ensemble)"}}
    path = os.path.join(ROOT, "profiles", "closure_report_64cube_vs_256cube.json")
    if os.path.exists(path):
        with open(path) as fh:
            rep = json.load(fh)
        out["drag_error_median_fp64_coarse"] = rep["drag_error"]["fp64_coarse"]["median"]
        out["drag_error_median_fp32_coarse_closure"] = rep["drag_error"]["fp32_coarse_closure"]["median"]
        out["drag_error_source"] = os.path.relpath(path, ROOT)
    return out


def electrolyte_line(steps=20, warmup=3):
    """Charged colloids at the electrolyte workload: 128^3, 32 spheres
    a = 5 with Z = 50, +-1 salt with Debye length 3, Onsager 0.1, E = (0.01,
    0, 0), uniform start; steps/s of the FP64 path and of FP32 transport
    (FP64 Poisson side), each `steps` steps after `warmup`."""
    from kapsel_old_amd.spm import SPM
    n, npart, a = 128, 32, 5.0
    el = abi.make_electrolyte(n_spec=2, valency=(1.0, -1.0), onsager=(0.1, 0.1), dielectric=0.018,
                              debye_length=3.0, surface_charge=(50.0,), e_ext=(0.01, 0.0, 0.0))
    pos = abi.random_positions(npart, (n, n, n), 2.2 * a, seed=12345)
    p = abi.new_particles(npart)
    p["x"] = pos
    p["x_nopbc"] = pos
    out = {"workload": "electrolyte, 128^3 grid, 32 spheres a=5 Z=50, +-1 salt lambda_D=3, E=(0.01,0,0), "
                       "literal schedule"}
    for prec, key in ((abi.FP64, "fp64"), (abi.FP32, "fp32_transport")):
        g = SPM(abi.make_params(n, a=a, a_xi=2.0, n_particles=npart, rotation=1, dt=0.05, kbt=1.0,
                                fix_cell=(0, 0, 0), precision=prec))
        g.put_particles(p)
        g.init_fluid()
        g.set_electrolyte(el)
        g.Init_rho_ion()
        g.step(warmup)
        g.synchronize()
        t0 = time.perf_counter()
        g.step(steps)
        g.synchronize()
        out[key] = {"value": steps / (time.perf_counter() - t0), "unit": "time-steps/s"}
        g.close()
    out["fp32_over_fp64"] = out["fp32_transport"]["value"] / out["fp64"]["value"]
    return out


def ensemble_line(nbox=64, steps=200, seed=4):
    """The batched ensemble: 64 boxes with grid edges drawn from
    {32, 48, 64, 80, 96}, 1-50 spheres a = 3-6 each, WCA, kBT = 0.05, FP32
    fields, `steps` steps of every box from one graph replay per step, after
    the autotuner has chosen the lane count (its steps are untimed); the
    host-launched rate of the same ensemble beside it.  `value` is the median
    of 3 timed runs of `steps` steps."""
    from kapsel_old_amd.ensemble import build, ensemble_boxes, grid_points
    from kapsel_old_amd.spm import autotune_report
    boxes = ensemble_boxes(nbox=nbox, seed=seed)
    ens = build(boxes)
    ens.autotune(3)
    ens.step(2)
    runs = []
    for _ in range(3):  # the median of 3 timed runs: the spread is about 10%
        t0 = time.perf_counter()
        ens.step(steps)
        runs.append(time.perf_counter() - t0)
    el = float(np.median(runs))
    lanes, tune_ms, box_ms = ens.tuning()
    ens.set_graph(False)
    k = max(10, steps // 10)
    t0 = time.perf_counter()
    ens.step(k)
    el_host = (time.perf_counter() - t0) / k
    ens.close()
    pts = grid_points(boxes)
    return {"workload": "%d boxes, edges from {32,48,64,80,96}, %d spheres in all (1-50 per box, a=3-6, xi=2), "
                        "WCA, kBT=0.05, FP32 fields, %d steps" % (nbox, sum(len(p) for _, p in boxes), steps),
            "value": nbox * steps / el, "unit": "box-steps/s",
            "ensemble_steps_per_s": steps / el, "ms_per_ensemble_step": 1e3 * el / steps,
            "ms_per_ensemble_step_runs": [1e3 * r / steps for r in runs],
            "grid_point_steps_per_s": pts * steps / el, "grid_points": pts,
            "lanes": lanes, "ms_per_step_at_1_2_4_8_lanes": [float(x) for x in tune_ms],
            "sum_of_lone_box_ms": float(box_ms.sum()),
            "host_launched_ms_per_ensemble_step": 1e3 * el_host,
            "mixed_radix_tuning": _tuning_summary(autotune_report())}


def _tuning_summary(report):
    """Passes the transform autotuner timed, how many it moved off the
    default launch shape, and the mean time saved on those."""
    picks, gains = 0, []
    lines = [ln for ln in report.splitlines() if ln.strip()]
    for ln in lines:
        pick = int(ln.split("pick=")[1].split()[0])
        ms = [float(x) for x in ln.split("ms=")[1].split(",")]
        dflt = 1
        if pick != dflt and ms[dflt] > 0:
            picks += 1
            gains.append(1.0 - ms[pick] / ms[dflt])
    return {"passes_timed": len(lines), "non_default_picks": picks,
            "mean_time_saved_on_those": float(np.mean(gains)) if gains else 0.0}


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
    ap.add_argument("--no-extras", action="store_true",
                    help="only the headline line (no FP32, one-step-call, closure or ensemble objects)")
    ap.add_argument("--ensemble-only", action="store_true",
                    help="print only the batched-ensemble object")
    args = ap.parse_args()
    if args.ensemble_only:
        print(json.dumps({"ensemble": ensemble_line()}), flush=True)
        return
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
        # the 20-transform model of the step's bytes over the measured time: an
        # effective bandwidth (the fused schedule runs 14 transforms)
        "hbm_TBps_effective": b_step * value / 1e12,
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
    if not args.no_extras and not sheared and precision == abi.FP64:
        e32, p32, ms32, by32, cnt32, _ = gpu_run(n, np_, args.steps, args.warmup, args.profile_steps, schedule,
                                                abi.FP32)
        d32 = int(np.argmax(ms32))
        pl32 = ms32[d32] / max(cnt32[d32], 1)
        out["fp32"] = {"value": args.steps / e32, "ms_per_step": 1e3 * e32 / args.steps,
                       "dtype": "f32 fields, f64 particles",
                       "roofline": {"bound": "hbm", "kernel": FAMILIES[d32],
                                    "achieved": by32[d32] / max(cnt32[d32], 1) / (pl32 * 1e-3) / 1e9,
                                    "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                    "frac": by32[d32] / max(cnt32[d32], 1) / (pl32 * 1e-3) / 1e9 / HBM_PEAK_GBS,
                                    "us_per_launch": pl32 * 1e3}}
        out["one_step_per_call"] = {"value": one_step_calls(n, np_, max(20, args.steps // 2)),
                                    "unit": "time-steps/s", "note": "spm_step(1) per step, FP64 fused"}
        out["closure"] = closure_line()
        out["electrolyte"] = electrolyte_line()
        out["ensemble"] = ensemble_line()
    if not args.no_cpu_baseline and precision == abi.FP64:
        nthreads = int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count() or 1)
        base = cpu_baseline(n, np_, shear_rate=args.shear_rate, threads=(1, nthreads) if nthreads > 1 else (1,))
        v1, s1 = base[1]
        out["cpu_baseline"] = {"value": v1, "unit": "time-steps/s", "cores": 1, "kind": "port",
                               "sample": "median of %d single steps of the same %d^3/%d workload after an untimed "
                                         "init and step, oracle/liboracle.so on 1 thread (%s s)"
                                         % (len(s1), n, np_, ", ".join("%.2f" % x for x in s1))}
        if nthreads > 1:
            vn, sn = base[nthreads]
            out["cpu_baseline"]["all_cores"] = {"value": vn, "cores": nthreads,
                                                "sample": "median of %d single steps on %d OpenMP threads (%s s)"
                                                          % (len(sn), nthreads, ", ".join("%.2f" % x for x in sn))}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
