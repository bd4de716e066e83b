"""Batched-ensemble workloads (include/spm_ensemble.h): independent periodic
boxes with uneven grids and particle counts, advanced together on one GPU.

`ensemble_boxes` draws the standard workload: per box, each grid edge
uniformly from {32, 48, 64, 80, 96}, a radius uniformly in [3, 6] with
interface thickness xi = 2, 1 to 50 spheres (fewer where the box cannot hold
them below the volume fraction cap) at seeded non-overlapping random
positions, WCA repulsion and kBT = 0.05.  `build` turns the draw into an
Ensemble with particles placed and the fluid initialised.
"""
import math

import numpy as np

from . import abi

EDGES = (32, 48, 64, 80, 96)


def ensemble_boxes(nbox=64, seed=4, precision=abi.FP32, kbt=0.05, dt=0.05, max_particles=50,
                   max_fraction=0.15, edges=EDGES, device=0):
    """[(spm_params, particles)] for nbox boxes drawn from `seed`."""
    rng = np.random.default_rng(seed)
    out = []
    for b in range(nbox):
        e = [int(x) for x in rng.choice(edges, 3)]
        a = float(rng.uniform(3.0, 6.0))
        n = int(rng.integers(1, max_particles + 1))
        vol = float(e[0] * e[1] * e[2])
        n = max(1, min(n, int(max_fraction * vol / (4.0 / 3.0 * math.pi * a ** 3))))
        prm = abi.make_params(e[0], e[1], e[2], a=a, a_xi=2.0, n_particles=n, kbt=kbt, dt=dt,
                              lj_truncate=abi.LJ_WCA, rotation=1, seed=seed * 1000 + b, precision=precision,
                              device=device)
        p = abi.new_particles(n)
        pos = abi.random_positions(n, e, 2.1 * a, seed=seed * 7919 + b)
        p["x"] = pos
        p["x_nopbc"] = pos
        out.append((prm, p))
    return out


def build(boxes):
    """An Ensemble of the drawn boxes, particles placed, fluid initialised."""
    from .spm import Ensemble
    ens = Ensemble([prm for prm, _ in boxes])
    for bx, (_, p) in zip(ens.boxes, boxes):
        bx.put_particles(p)
        bx.init_fluid()
    return ens


def grid_points(boxes):
    return sum(prm.nx * prm.ny * prm.nz for prm, _ in boxes)


if __name__ == "__main__":
    import argparse
    ap = argparse.ArgumentParser(description="run the standard ensemble draw")
    ap.add_argument("--boxes", type=int, default=64)
    ap.add_argument("--seed", type=int, default=4)
    ap.add_argument("--autotune", type=int, default=2, metavar="STEPS", help="steps per lane candidate (0: none)")
    ap.add_argument("--steps", type=int, default=10)
    a = ap.parse_args()
    ens = build(ensemble_boxes(nbox=a.boxes, seed=a.seed))
    if a.autotune:
        ens.autotune(a.autotune)
    ens.step(a.steps)
    lanes, ms, box_ms = ens.tuning()
    print("boxes %d lanes %d ms/step at 1,2,4,8 lanes %s" % (a.boxes, lanes, ms), flush=True)
