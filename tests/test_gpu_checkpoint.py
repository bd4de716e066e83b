"""Checkpoint and resume on the GPU.

spm_save_state / spm_load_state: a run split by a save and a load into a
fresh context ends bit for bit where the unbroken run ends -- fields,
uk_dc, particles, time and the MT19937 state -- with the thermostat on, on
both schedules and precisions and for an electrolyte.  The restart UDF
(Save_Restart_udf / Force_restore_parameters, resume.cxx:10-364) resumes
a kBT = 0 run to within the renormalization of q that the reference's
restore applies.
"""
import numpy as np
import pytest

from kapsel_old_amd import abi
from tests.helpers import particles

pytestmark = pytest.mark.gpu

GRID = (32, 32, 32)


@pytest.fixture(scope="module")
def SPM():
    from kapsel_old_amd.spm import SPM as _SPM
    return _SPM


def state(g):
    s = {"zeta0": g.get_field("zeta0"), "zeta1": g.get_field("zeta1"), "uk_dc": g.get_uk_dc(),
         "time": g.get_time(), "rng": g.get_rng()}
    p = g.get_particles()
    for k in p.dtype.names:
        s["p." + k] = p[k]
    if getattr(g, "_electro", False):
        for n in range(2):
            s["conc%d" % n] = g.get_concentration(n)
    return s


def assert_same(a, b):
    assert a.keys() == b.keys()
    for k in a:
        if k == "rng":
            np.testing.assert_array_equal(a[k][0], b[k][0])
            assert a[k][1] == b[k][1]
        elif k == "time":
            assert a[k] == b[k]
        else:
            np.testing.assert_array_equal(a[k], b[k], err_msg=k)


def make(SPM, prm, p, el=None):
    g = SPM(prm)
    g.put_particles(p)
    g.init_fluid()
    if el is not None:
        g.set_electrolyte(el)
        g.Init_rho_ion()
        g._electro = True
    return g


@pytest.mark.parametrize("schedule,precision,kbt", [(0, abi.FP64, 1e-3), (1, abi.FP64, 1e-3), (0, abi.FP32, 1e-3),
                                                    (0, abi.FP64, 0.0)])
def test_state_resume_is_exact(SPM, tmp_path, schedule, precision, kbt):
    print("schedule", schedule, "precision", precision, "kbt", kbt, flush=True)
    n = 6
    prm = abi.make_params(*GRID, a=4.0, n_particles=n, dt=0.05, kbt=kbt, precision=precision,
                          lj_truncate=abi.LJ_WCA, fix_cell=(1, 0, 1), g=0.5, mass_ratio=(1.0, 2.0))
    p = particles(n, GRID, 4.0, seed=21, vel=0.01, omega=0.005, spec_count=2)
    a = make(SPM, prm, p)
    a.set_schedule(schedule)
    a.step(3)
    f = tmp_path / "state.bin"
    a.save_state(f)
    a.step(4)
    b = make(SPM, prm, particles(n, GRID, 4.0, seed=99))
    b.set_schedule(schedule)
    b.load_state(f)
    b.step(4)
    sa, sb = state(a), state(b)
    assert sa["time"][0] == 7
    assert_same(sa, sb)


@pytest.mark.parametrize("precision", [abi.FP64, abi.FP32], ids=["fp64", "fp32_transport"])
def test_state_resume_electrolyte(SPM, tmp_path, precision):
    n = 5
    prm = abi.make_params(*GRID, a=4.0, n_particles=n, dt=0.05, kbt=1.0, fix_cell=(0, 0, 0), precision=precision)
    el = abi.make_electrolyte(n_spec=2, valency=(1.0, -1.0), onsager=(0.1, 0.1), dielectric=1.0, debye_length=3.0,
                              surface_charge=(5.0,), e_ext=(0.01, 0.0, 0.0))
    p = particles(n, GRID, 4.0, seed=31, vel=0.02, omega=0.005)
    a = make(SPM, prm, p, el)
    a.step(2)
    f = tmp_path / "state.bin"
    a.save_state(f)
    a.step(2)
    b = make(SPM, prm, p, el)
    b.load_state(f)
    b.step(2)
    assert_same(state(a), state(b))


def test_state_load_rejects(SPM, tmp_path):
    from kapsel_old_amd.spm import SpmError
    n = 3
    prm = abi.make_params(*GRID, a=4.0, n_particles=n)
    p = particles(n, GRID, 4.0, seed=5, vel=0.01)
    a = make(SPM, prm, p)
    a.step(1)
    f = tmp_path / "state.bin"
    a.save_state(f)
    other = make(SPM, abi.make_params(32, 32, 64, a=4.0, n_particles=n), p)
    before = state(other)
    with pytest.raises(SpmError):
        other.load_state(f)
    assert_same(before, state(other))
    cut = tmp_path / "cut.bin"
    cut.write_bytes(f.read_bytes()[:-100])
    flip = tmp_path / "flip.bin"
    raw = bytearray(f.read_bytes())
    raw[len(raw) // 2] ^= 1
    flip.write_bytes(bytes(raw))
    b = make(SPM, prm, p)
    before = state(b)
    for bad in (cut, flip, tmp_path / "missing.bin"):
        with pytest.raises(SpmError):
            b.load_state(bad)
    assert_same(before, state(b))


def test_udf_restart_resume(SPM, tmp_path):
    """Save_Restart_udf at ts, Force_restore_parameters into a fresh context."""
    import os
    from kapsel_old_amd import udf
    golden = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "udf", "input.udf")
    with udf.UDF(golden) as u:
        prm, _, run = u.params()
    prm.lj_truncate = abi.LJ_WCA
    n = prm.n_particles
    p = particles(n, GRID, prm.a, seed=41, vel=0.01, omega=0.005, spec_count=2, min_factor=2.3)
    a = make(SPM, prm, p)
    a.step(3)
    ts, t = a.get_time()
    out = tmp_path / "restart.udf"
    udf.write_restart(golden, out, GRID, a.get_field("zeta0"), a.get_field("zeta1"), a.get_uk_dc(),
                      a.get_particles(), ts, t)
    a.step(2)
    b = make(SPM, prm, p)
    b.step(1)  # the reference runs one step before it restores (sp_3d_ns.cxx:619-628)
    with udf.UDF(out) as u:
        r = udf.read_restart(u, GRID, n, particles=b.get_particles())
    b.put_field("zeta0", r["zeta0"])
    b.put_field("zeta1", r["zeta1"])
    b.put_uk_dc(r["uk_dc"])
    b.put_particles(r["particles"])
    b.put_time(r["ts"], r["time"])
    b.step(2)
    assert a.get_time() == b.get_time()
    for k in ("zeta0", "zeta1"):
        za, zb = a.get_field(k), b.get_field(k)
        assert np.abs(za - zb).max() <= 1e-12 * np.abs(za).max(), k
    pa, pb = a.get_particles(), b.get_particles()
    for k in ("x", "v", "omega", "f_hydro_previous", "fr_previous"):
        assert np.abs(pa[k] - pb[k]).max() <= 1e-12 * max(1.0, np.abs(pa[k]).max()), k
