"""What a step reads: the oracle (the reference's algorithm on the CPU)
continues bit for bit from zeta, uk_dc, the particle records, ts/time and
the MT19937 state (plus the ion spectra and totals for an electrolyte),
put into a fresh instance.  This is the state list spm_save_state keeps;
a field missing from it would make the resumed run differ here.
"""
import numpy as np
import pytest

from kapsel_old_amd import abi
from oracle.oracle import Oracle
from tests.helpers import particles

GRID = (16, 16, 16)


def snapshot(o, ions=0):
    s = dict(zeta0=o.get_field("zeta0"), zeta1=o.get_field("zeta1"), uk_dc=o.get_uk_dc(), time=o.get_time(),
             rng=o.get_rng(), p=o.get_particles())
    for n in range(ions):
        s["conc%d" % n] = o.get_concentration(n)
    if ions:
        s["total"] = o.get_total_solute()
    return s


def restore(o, s, ions=0):
    o.put_field("zeta0", s["zeta0"])
    o.put_field("zeta1", s["zeta1"])
    o.put_uk_dc(s["uk_dc"])
    o.put_particles(s["p"])
    o.put_time(*s["time"])
    o.put_rng(*s["rng"])
    for n in range(ions):
        o.put_concentration(n, s["conc%d" % n])
    if ions:
        o.put_total_solute(s["total"])


def same(a, b):
    for k in a:
        if k == "p":
            for f in a[k].dtype.names:
                np.testing.assert_array_equal(a[k][f], b[k][f], err_msg=f)
        elif k == "rng":
            np.testing.assert_array_equal(a[k][0], b[k][0])
            assert a[k][1] == b[k][1]
        else:
            np.testing.assert_array_equal(np.asarray(a[k]), np.asarray(b[k]), err_msg=k)


@pytest.mark.parametrize("kbt", [0.0, 1e-3])
def test_oracle_resume_navier_stokes(kbt):
    prm = abi.make_params(*GRID, a=2.0, a_xi=1.0, n_particles=3, dt=0.05, kbt=kbt, lj_truncate=abi.LJ_WCA,
                          fix_cell=(1, 0, 1))
    p = particles(3, GRID, 2.0, seed=4, vel=0.01, omega=0.005)
    a = Oracle(prm)
    a.put_particles(p)
    a.init_fluid()
    a.step(2)
    mid = snapshot(a)
    a.step(2)
    b = Oracle(prm)
    b.put_particles(particles(3, GRID, 2.0, seed=8))
    b.init_fluid()
    b.step(1)  # state a resumed run would overwrite
    restore(b, mid)
    b.step(2)
    same(snapshot(a), snapshot(b))


def test_oracle_resume_electrolyte():
    prm = abi.make_params(*GRID, a=2.0, a_xi=1.0, n_particles=2, dt=0.05, kbt=1.0, fix_cell=(0, 0, 0))
    el = abi.make_electrolyte(n_spec=2, valency=(1.0, -1.0), onsager=(0.1, 0.1), dielectric=1.0, debye_length=3.0,
                              surface_charge=(2.0,), e_ext=(0.01, 0.0, 0.0))
    p = particles(2, GRID, 2.0, seed=6, vel=0.01)

    def fresh():
        o = Oracle(prm)
        o.put_particles(p)
        o.init_fluid()
        o.set_electrolyte(el)
        o.Init_rho_ion()
        return o

    a = fresh()
    a.step(1)
    mid = snapshot(a, 2)
    a.step(2)
    b = fresh()
    b.step(1)
    restore(b, mid, 2)
    b.step(2)
    same(snapshot(a, 2), snapshot(b, 2))
