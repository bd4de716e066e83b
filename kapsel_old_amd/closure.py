"""Training pairs, training and evaluation of the learned closure
(include/spm_closure.h) with the library's own SPM step.

A pair is one suspension run twice with prescribed particle motion (the
resistance problem: positions, velocities and angular velocities held, the
step forms the hydrodynamic force and torque of Calc_f_hydro_correct_precision,
md_force.cxx:263-389): on a coarse grid (radius a grid points) and on a grid
`refine` times finer (radius refine * a), both with an interface of 2 grid
points and rho = eta = 1.  In the fine run's own units lengths are refine
times larger, times refine^2 larger and velocities refine times smaller, so
forces compare as numbers and torques after dividing the fine one by
refine.  The coarse run's closure inputs and the fine-minus-coarse force
and torque make one training sample per particle.
"""
import itertools

import numpy as np

from . import abi
from .spm import SPM, Closure


def random_suspension(n, box, a, phi, seed, gap=0.2):
    """Positions of round(phi * box^3 / (4/3 pi a^3)) spheres (n overrides
    when given) placed uniformly without overlap closer than 2a + gap,
    random unit-scale velocities and angular velocities."""
    vol = box ** 3
    if n is None:
        n = int(round(phi * vol / (4.0 / 3.0 * np.pi * a ** 3)))
    pos = abi.random_positions(n, (box, box, box), 2 * a + gap, seed=seed)
    rng = np.random.default_rng(seed + 7)
    v = rng.standard_normal((n, 3)) * 0.01
    w = rng.standard_normal((n, 3)) * 0.01 / a
    return pos, v, w


def _params(n_grid, a, n, dt, precision):
    return abi.make_params(n_grid, a=a, a_xi=2.0, dt=dt, kbt=0.0, n_particles=n, rotation=1,
                           lj_truncate=abi.LJ_NONE, fix_cell=(1, 1, 1), precision=precision)


def resistance_run(n_grid, a, pos, v, w, dt, steps, precision=abi.FP64, closure=None, with_features=False):
    """Prescribed-motion run; returns f_hydro, torque_hydro of the last step
    (and the closure inputs of the final state)."""
    n = len(pos)
    g = SPM(_params(n_grid, a, n, dt, precision))
    p = abi.new_particles(n)
    p["x"], p["x_nopbc"], p["v"], p["omega"] = pos, pos, v, w
    g.put_particles(p)
    g.init_fluid()
    g.set_particle_motion(True)
    if closure is not None:
        g.set_closure(closure)
    g.step(steps)
    q = g.get_particles()
    out = [q["f_hydro"].copy(), q["torque_hydro"].copy()]
    if with_features:
        g.Zeta_k2u()  # real-space u of the final state
        out.append(g.closure_features())
    g.close()
    return out


def pair(n_grid, a, pos, v, w, dt, steps, refine=4):
    """One coarse/fine pair: coarse inputs x [n][126] and the target
    correction t [n][6] = (F_fine - F_coarse, T_fine / refine - T_coarse)."""
    fc, tc, x = resistance_run(n_grid, a, pos, v, w, dt, steps, with_features=True)
    r = refine
    ff, tf = resistance_run(r * n_grid, r * a, pos * r, v / r, w / r ** 2, dt * r * r, steps)
    t = np.concatenate([ff - fc, tf / r - tc], axis=1)
    return x, t, np.concatenate([ff, tf / r], axis=1), np.concatenate([fc, tc], axis=1)


def fit_scaling(x, t):
    """Input shift/scale to zero mean and unit spread, output scale to the
    spread of each target component."""
    shift = x.mean(axis=0)
    sd = x.std(axis=0)
    scale = np.where(sd > 1e-12, 1.0 / np.maximum(sd, 1e-12), 1.0)
    osd = t.std(axis=0)
    oscale = np.where(osd > 1e-30, osd, 1.0)
    return shift.astype(np.float32), scale.astype(np.float32), oscale.astype(np.float32)


def train_closure(x, t, steps=2000, batch=256, lr=1e-3, seed=1, max_batch=None):
    """A Closure fitted to (x, t) on the GPU (Adam, minibatches of a fixed
    shuffle)."""
    rng = np.random.default_rng(seed)
    order = rng.permutation(len(x))
    x, t = x[order], t[order]
    net = Closure(max(max_batch or 0, batch, 64), seed=seed)
    net.set_scaling(*fit_scaling(x, t))
    net.set_data(x, t)
    loss = net.train(steps, batch, lr)
    return net, loss


def drag_error(f, f_ref):
    """Relative error of each particle's force against the judge, |f - f_ref|
    / |f_ref|, per particle."""
    return np.linalg.norm(f - f_ref, axis=1) / np.linalg.norm(f_ref, axis=1)


def dataset(configs, n_grid=64, a=3.0, dt=0.05, steps=600, refine=4, log=None):
    """Pairs for a list of (volume fraction, seed): stacked inputs, targets,
    fine and coarse (force, torque)."""
    xs, ts, fines, coarses = [], [], [], []
    for phi, seed in configs:
        pos, v, w = random_suspension(None, n_grid, a, phi, seed)
        x, t, fine, coarse = pair(n_grid, a, pos, v, w, dt, steps, refine)
        xs.append(x)
        ts.append(t)
        fines.append(fine)
        coarses.append(coarse)
        if log:
            e = drag_error(coarse[:, :3], fine[:, :3])
            log("pair phi %.2f seed %d: %d particles, coarse drag error median %.3f" % (phi, seed, len(x),
                                                                                          np.median(e)))
    return np.concatenate(xs), np.concatenate(ts), np.concatenate(fines), np.concatenate(coarses)


# the 27 sample offsets of the inputs, lexicographic with x slowest
OFFSETS = np.array([(o // 9 - 1, (o // 3) % 3 - 1, o % 3 - 1) for o in range(27)])


def cube_symmetries():
    """The 48 signed permutation matrices that map the cubic lattice (and
    the periodic box) onto itself."""
    mats = []
    for perm in itertools.permutations(range(3)):
        for signs in itertools.product((1, -1), repeat=3):
            S = np.zeros((3, 3))
            for i in range(3):
                S[i, perm[i]] = signs[i]
            mats.append(S)
    return mats


def transform(x, t, S):
    """Inputs and targets of the configuration mapped by the lattice
    symmetry S: vectors turn with S, the torque (a pseudovector) with
    det(S) S, and the sample at offset o is the old sample at S^T o."""
    n = len(x)
    lookup = {tuple(o): k for k, o in enumerate(OFFSETS)}
    idx = np.array([lookup[tuple(np.rint(S.T @ o).astype(int))] for o in OFFSETS])
    xn = np.empty_like(x)
    xn[:, 0:3] = x[:, 0:3] @ S.T
    xn[:, 3:84] = (x[:, 3:84].reshape(n, 27, 3)[:, idx, :] @ S.T).reshape(n, 81)
    xn[:, 84:111] = x[:, 84:111][:, idx]
    det = np.rint(np.linalg.det(S))
    xn[:, 111:114] = x[:, 111:114] @ S.T  # nearest neighbour: gap, velocity, angular velocity
    xn[:, 114:117] = x[:, 114:117] @ S.T
    xn[:, 117:120] = det * (x[:, 117:120] @ S.T)
    xn[:, 120:123] = x[:, 120:123] @ S.T  # second nearest: gap, velocity
    xn[:, 123:126] = x[:, 123:126] @ S.T
    tn = np.empty_like(t)
    tn[:, 0:3] = t[:, 0:3] @ S.T
    tn[:, 3:6] = det * (t[:, 3:6] @ S.T)
    return xn.astype(np.float32), tn.astype(np.float32)


def augment(x, t):
    """The training set under all 48 lattice symmetries."""
    pairs = [transform(x, t, S) for S in cube_symmetries()]
    return np.concatenate([p[0] for p in pairs]), np.concatenate([p[1] for p in pairs])


def free_run_speed(n_grid, a, pos, dt, steps, precision, closure=None, warmup=20, graph=False):
    """Steps per second of free suspensions (the production use) with the
    closure attached or not; graph=True replays the step from a captured
    graph (a one-box ensemble, include/spm_ensemble.h) instead of launching
    its kernels from the host."""
    import time
    from .spm import Ensemble
    n = len(pos)
    prm = _params(n_grid, a, n, dt, precision)
    ens = Ensemble([prm]) if graph else None
    g = ens.boxes[0] if graph else SPM(prm)
    p = abi.new_particles(n)
    p["x"], p["x_nopbc"] = pos, pos
    g.put_particles(p)
    g.init_fluid()
    if closure is not None:
        g.set_closure(closure)
    run = ens.step if graph else g.step
    run(warmup)
    g.synchronize()
    t0 = time.perf_counter()
    run(steps)
    g.synchronize()
    t1 = time.perf_counter()
    if graph:
        ens.close()
    else:
        g.close()
    return steps / (t1 - t0)


def report(seeds=8, steps=600, train_steps=40000, log=print):
    """The learned closure at the sizes of the coarse/fine comparison: pairs
    of 64^3 (a = 3) and 256^3 (a = 12) resistance runs at volume fractions
    0.05 .. 0.3 (`seeds` suspensions each), training on the GPU with the 48
    lattice symmetries, evaluation on two held-out suspensions of 512
    spheres in 64^3 (volume fraction 0.22) with the network attached to an
    FP32 coarse step, and the speed of that step against the FP64 fine step."""
    import time
    t0 = time.time()
    cfg = [(phi, 100 + k) for k in range(seeds) for phi in (0.05, 0.1, 0.15, 0.2, 0.25, 0.3)]
    x, t, fine, coarse = dataset(cfg, steps=steps)
    t_data = time.time() - t0
    log("training pairs: %d suspensions, %d particles, %.0f s" % (len(cfg), len(x), t_data))
    xa, ta = augment(x, t)
    t1 = time.time()
    net, loss = train_closure(xa, ta, steps=train_steps, batch=1024, lr=5e-4, max_batch=8192)
    t_train = time.time() - t1
    log("training: %d samples after symmetries, %d Adam steps of 1024, loss %.4f -> %.4f, %.1f s"
        % (len(xa), train_steps, loss[:20].mean(), loss[-20:].mean(), t_train))
    errs = {"fp64_coarse": [], "fp32_coarse_closure": [], "torque_fp64_coarse": [], "torque_fp32_coarse_closure": []}
    for seed in (900, 901):
        pos, v, w = random_suspension(512, 64, 3.0, None, seed)
        fc, tc = resistance_run(64, 3.0, pos, v, w, 0.05, steps)
        ff, tf = resistance_run(256, 12.0, pos * 4, v / 4, w / 16, 0.05 * 16, steps)
        fx, tx = resistance_run(64, 3.0, pos, v, w, 0.05, steps, precision=abi.FP32, closure=net)
        errs["fp64_coarse"].append(drag_error(fc, ff))
        errs["fp32_coarse_closure"].append(drag_error(fx, ff))
        errs["torque_fp64_coarse"].append(drag_error(tc, tf / 4))
        errs["torque_fp32_coarse_closure"].append(drag_error(tx, tf / 4))
    stats = {}
    for k, v in errs.items():
        e = np.concatenate(v)
        stats[k] = {"median": float(np.median(e)), "mean": float(e.mean()), "p95": float(np.percentile(e, 95))}
        log("%-28s median %.4f mean %.4f p95 %.4f" % (k, stats[k]["median"], stats[k]["mean"], stats[k]["p95"]))
    pos, _, _ = random_suspension(512, 64, 3.0, None, 900)
    speed = {"fp32_64cube_closure": free_run_speed(64, 3.0, pos, 0.05, 1000, abi.FP32, closure=net),
             "fp64_64cube": free_run_speed(64, 3.0, pos, 0.05, 1000, abi.FP64),
             "fp64_256cube_fine": free_run_speed(256, 12.0, pos * 4, 0.8, 200, abi.FP64)}
    for k, v in speed.items():
        log("%-22s %.1f steps/s" % (k, v))
    ratio = speed["fp32_64cube_closure"] / speed["fp64_256cube_fine"]
    log("FP32 coarse + closure vs FP64 fine: %.1fx the steps/s" % ratio)
    return {"pairs": len(cfg), "particles": int(len(x)), "samples_after_symmetries": int(len(xa)),
            "train_steps": train_steps, "batch": 1024, "lr": 5e-4, "loss_first": float(loss[:20].mean()),
            "loss_last": float(loss[-20:].mean()), "seconds_pairs": t_data, "seconds_training": t_train,
            "drag_error": stats, "steps_per_s": speed, "speedup_vs_fine": ratio,
            "coarse": "64^3, a = 3, xi = 2 grid points, dt = 0.05", "fine": "256^3, a = 12, xi = 2, dt = 0.8",
            "evaluation": "2 held-out suspensions of 512 spheres in 64^3 (volume fraction 0.22), "
                          "resistance problem, 600 steps"}


def mlp_workload(batch=65536, train_steps=40, minibatch=8192, forwards=10, seed=0):
    """The network's kernels alone, for profiling: `train_steps` Adam steps
    on minibatches of random samples, then `forwards` forwards of `batch`
    samples."""
    from .spm import Closure
    rng = np.random.default_rng(seed)
    x = rng.standard_normal((batch, 126)).astype(np.float32)
    t = (0.1 * rng.standard_normal((batch, 6))).astype(np.float32)
    net = Closure(batch, seed=3)
    net.set_scaling(np.zeros(126), np.ones(126), np.ones(6))
    net.set_data(x, t)
    loss = net.train(train_steps, minibatch, 1e-3)
    for _ in range(forwards):
        net.forward(x)
    return float(loss[0]), float(loss[-1])


if __name__ == "__main__":
    import argparse
    import json
    ap = argparse.ArgumentParser(description="coarse/fine closure report")
    ap.add_argument("--seeds", type=int, default=8)
    ap.add_argument("--train-steps", type=int, default=40000)
    ap.add_argument("--out", default=None)
    ap.add_argument("--mlp-workload", type=int, default=0, metavar="BATCH",
                    help="only run the network's kernels (40 training steps, 10 forwards of BATCH samples)")
    a = ap.parse_args()
    if a.mlp_workload:
        print("loss first/last", mlp_workload(a.mlp_workload), flush=True)
        raise SystemExit(0)
    r = report(seeds=a.seeds, train_steps=a.train_steps, log=lambda s: print(s, flush=True))
    print(json.dumps(r))
    if a.out:
        with open(a.out, "w") as fh:
            json.dump(r, fh, indent=1)
