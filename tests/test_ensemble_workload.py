"""The batched-ensemble workload draw (kapsel_old_amd/ensemble.py), on the
CPU: grid edges from {32, 48, 64, 80, 96}, radii in [3, 6], 1 to 50
spheres below the volume-fraction cap, non-overlapping under periodic
images, reproducible from the seed."""
import math

import numpy as np

from kapsel_old_amd import abi
from kapsel_old_amd.ensemble import EDGES, ensemble_boxes, grid_points


def test_draw_is_valid_and_reproducible():
    boxes = ensemble_boxes(nbox=64, seed=4)
    again = ensemble_boxes(nbox=64, seed=4)
    assert len(boxes) == 64
    edges_seen = set()
    for (prm, p), (prm2, p2) in zip(boxes, again):
        e = (prm.nx, prm.ny, prm.nz)
        edges_seen |= set(e)
        assert set(e) <= set(EDGES)
        assert 3.0 <= prm.a <= 6.0 and prm.a_xi == 2.0
        assert 1 <= prm.n_particles <= 50 and len(p) == prm.n_particles
        assert prm.precision == abi.FP32 and prm.kbt == 0.05 and prm.lj_truncate == abi.LJ_WCA
        frac = prm.n_particles * 4.0 / 3.0 * math.pi * prm.a ** 3 / (e[0] * e[1] * e[2])
        assert prm.n_particles == 1 or frac <= 0.15 + 1e-12
        x = p["x"]
        assert (x >= 0).all() and (x < np.array(e)).all()
        for i in range(len(x)):
            d = x[i + 1:] - x[i]
            d -= np.round(d / np.array(e)) * np.array(e)
            assert (np.sqrt((d * d).sum(axis=1)) > 2.1 * prm.a).all()
        assert e == (prm2.nx, prm2.ny, prm2.nz) and np.array_equal(p["x"], p2["x"])
    assert edges_seen == set(EDGES)
    assert grid_points(boxes) == sum(prm.nx * prm.ny * prm.nz for prm, _ in boxes)
