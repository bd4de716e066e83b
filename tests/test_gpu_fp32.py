"""The FP32 field path against the FP64 oracle.

FP32 carries the fields, FFTs, spectral operators and the phi scatter;
particles, rigid-body/quaternion integration and the force/torque sums
stay FP64.  Each test states the error of every output against the FP64
oracle run from the same state; the bounds are what single-precision
fields allow (a few 1e-6 relative per step, growing slowly with steps).
"""
import numpy as np
import pytest

from kapsel_old_amd import abi
from oracle.oracle import Oracle
from tests.helpers import particles, rel

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("schedule", [0, 1], ids=["fused", "literal"])
def test_fp32_step_errors_against_fp64_oracle(schedule):
    from kapsel_old_amd.spm import SPM
    n, npart, a = 64, 24, 5.0
    kw = dict(a=a, n_particles=npart, kbt=0.1, lj_truncate=abi.LJ_WCA, seed=12345, rotation=1)
    o = Oracle(abi.make_params(n, **kw))
    g = SPM(abi.make_params(n, precision=abi.FP32, **kw))
    g.set_schedule(schedule)
    p = particles(npart, (n, n, n), a, seed=23, min_factor=2.05, vel=0.01, omega=0.003)
    for x in (o, g):
        x.put_particles(p)
        x.init_fluid()
    report = {}
    for s in range(5):
        for x in (o, g):
            x.step(1)
        for nm in ["u0", "u1", "u2", "zeta0", "zeta1"]:
            report[nm] = rel(g.get_field(nm), o.get_field(nm))
        po, pg = o.get_particles(), g.get_particles()
        for k in ["x", "v", "omega", "f_hydro_previous", "torque_hydro_previous"]:
            report[k] = rel(pg[k], po[k])
        print("step %d: %s" % (s, ", ".join("%s %.2e" % kv for kv in report.items())), flush=True)
    for k in ["u0", "u1", "u2", "zeta0", "zeta1"]:
        assert report[k] < 1e-4, (k, report[k])
    for k in ["v", "omega", "f_hydro_previous", "torque_hydro_previous"]:
        assert report[k] < 1e-4, (k, report[k])
    assert report["x"] < 1e-9


def test_fp32_terminal_velocity():
    """Sedimenting sphere (64^3, a = 5, F = 0.01): FP32 fields reach the FP64
    terminal velocity to 1e-4 relative."""
    from kapsel_old_amd.spm import SPM
    n, a, F, mr = 64, 5.0, 0.01, 2.0
    G = -F / ((4.0 / 3.0) * np.pi * a ** 3 * (mr - 1.0))
    out = []
    for prec in (abi.FP64, abi.FP32):
        g = SPM(abi.make_params(n, a=a, n_particles=1, mass_ratio=(mr,), g=G, g_direction=2, precision=prec))
        g.put_particles(particles(1, (n, n, n), a, positions=np.array([[32.0, 32.0, 32.0]])))
        g.init_fluid()
        g.step(3000)
        out.append(g.get_particles()["v"][0, 2])
    e = abs(out[1] - out[0]) / abs(out[0])
    print("terminal vz fp64 %.8e fp32 %.8e rel %.2e" % (out[0], out[1], e))
    assert e < 1e-4
