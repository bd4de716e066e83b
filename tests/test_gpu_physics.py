"""Whole-step checks on the GPU path that do not need the oracle.

The step as a whole cannot be pinned against golden trajectories (the
reference's step is not buildable here), so this pins the coupled physics
against a closed form: Hasimoto's drag on a simple cubic array of spheres,
with the fluid inside the particle moving rigidly with it at steady state.
"""
import numpy as np
import pytest

from kapsel_old_amd import abi
from tests.helpers import particles

pytestmark = pytest.mark.gpu


def hasimoto_k(c):
    """Drag correction K(c) of a simple cubic array (Hasimoto 1959, with the
    higher terms of Sangani & Acrivos 1982), F = 6 pi eta a U K."""
    return 1.0 / (1 - 1.7601 * c ** (1 / 3) + c - 1.5593 * c ** 2 + 3.9799 * c ** (8 / 3) - 3.0734 * c ** (10 / 3))


@pytest.mark.parametrize("schedule", [0, 1], ids=["fused", "literal"])
def test_terminal_velocity_matches_hasimoto(schedule):
    from kapsel_old_amd.spm import SPM
    n, a, F, mr = 64, 5.0, 0.01, 2.0
    vol = 4.0 / 3.0 * np.pi * a ** 3
    G = -F / (vol * (mr - 1.0))
    prm = abi.make_params(n, a=a, n_particles=1, mass_ratio=(mr,), g=G, g_direction=2, rotation=1, dt=0.05)
    g = SPM(prm)
    g.set_schedule(schedule)
    g.put_particles(particles(1, (n, n, n), a, positions=np.array([[32.0, 32.0, 32.0]])))
    g.init_fluid()
    trace = []
    for _ in range(60):
        g.step(200)
        trace.append(g.get_particles()["v"][0, 2])
    c = vol / n ** 3
    u_expected = F / (6 * np.pi * a * hasimoto_k(c))
    u_end = trace[-1]
    print("terminal vz %.6e, Hasimoto %.6e, ratio %.4f; last changes %s" % (
        u_end, u_expected, u_end / u_expected, np.diff(trace[-4:])))
    assert abs(trace[-1] - trace[-2]) < 1e-3 * abs(u_end)   # converged
    assert abs(u_end / u_expected - 1.0) < 0.03
    # the fluid inside the particle moves with it (u is a spectrum after a
    # step; Zeta_k2u gives it in real space)
    phi = g.get_field("phi")[:, :, :n]
    g.Zeta_k2u()
    uz = g.get_field("u2")[:, :, :n]
    inside = phi >= 1.0
    assert inside.sum() > 100
    assert np.abs(uz[inside] / u_end - 1.0).max() < 0.02
