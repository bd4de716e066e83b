"""Parity of the electrolyte path (SW_EQ == Electrolyte) of libspm.so with
the CPU oracle: Init_rho_ion, Rescale_solute, Make_Coulomb_force_x_on_fluid,
NSsolute_solver_Euler and whole steps, from identical state.  Tolerance
1e-10 relative, as for the rest of the FP64 path.
"""
import numpy as np
import pytest

from kapsel_old_amd import abi
from oracle.oracle import Oracle
from tests.helpers import particles, rel, smooth_field

pytestmark = pytest.mark.gpu

TOL = 1e-10


@pytest.fixture(scope="module")
def SPM():
    from kapsel_old_amd.spm import SPM as _SPM
    return _SPM


def electrolyte(n_spec=2, **kw):
    base = dict(n_spec=n_spec, valency=(1.0, -1.0) if n_spec == 2 else (-1.0,), onsager=(0.1, 0.1),
                dielectric=1.0, debye_length=3.0, surface_charge=(5.0,), e_ext=(0.01, 0.0, 0.0))
    base.update(kw)
    return abi.make_electrolyte(**base)


def pair(SPM, grid, n, a, el, seed=3, kbt=1.0, vel=0.01, omega=0.002):
    nx, ny, nz = grid
    prm = abi.make_params(nx, ny, nz, a=a, n_particles=n, rotation=1, dt=0.05, kbt=kbt, fix_cell=(0, 0, 0))
    p = particles(n, grid, a, seed=seed, vel=vel, omega=omega)
    objs = []
    for cls in (Oracle, SPM):
        o = cls(prm)
        o.put_particles(p)
        o.init_fluid()
        o.set_electrolyte(el)
        o.Init_rho_ion()
        objs.append(o)
    return objs


def perturb(objs, nz, seed):
    """The same band-limited perturbation of every ion spectrum in both."""
    o = objs[0]
    for s in range(o.n_spec):
        c = o.get_concentration(s)
        dk = o.A2a_k(smooth_field(o.shape, nz, seed=seed + s, kmax=4, scale=1e-3))
        for obj in objs:
            obj.put_concentration(s, c + dk)


def check(name, a, b, tol=TOL):
    e = rel(a, b)
    print("%s rel err %.3e" % (name, e), flush=True)
    assert e < tol, (name, e)


GRIDS = [(32, 32, 32), (32, 64, 32)]


@pytest.mark.parametrize("n_spec", [1, 2])
def test_init_rho_ion(SPM, n_spec):
    el = electrolyte(n_spec, surface_charge=(5.0,) if n_spec == 2 else (5.0,),
                     valency=(1.0, -1.0) if n_spec == 2 else (-1.0, 0.0))
    o, g = pair(SPM, (32, 32, 32), 4, 4.0, el)
    for s in range(o.n_spec):
        check("conc%d" % s, g.get_concentration(s), o.get_concentration(s))
    check("total_solute", g.get_total_solute(), o.get_total_solute())


@pytest.mark.parametrize("grid", GRIDS)
def test_rescale_solute(SPM, grid):
    objs = pair(SPM, grid, 5, 4.0, electrolyte())
    perturb(objs, grid[2], seed=40)
    for obj in objs:
        obj.Rescale_solute()
    o, g = objs
    for s in range(2):
        check("conc%d" % s, g.get_concentration(s), o.get_concentration(s))


@pytest.mark.parametrize("ac", [0, 1])
def test_make_coulomb_force(SPM, ac):
    grid = (32, 32, 32)
    el = electrolyte(ac=ac, angular_frequency=0.7, e_ext=(0.01, -0.02, 0.005))
    objs = pair(SPM, grid, 5, 4.0, el)
    perturb(objs, grid[2], seed=50)
    for obj in objs:
        obj.put_time(3, 0.15)
        obj.Make_Coulomb_force_x_on_fluid()
    o, g = objs
    for d in range(3):
        check("f%d" % d, g.get_field("f%d" % d)[:, :, :grid[2]], o.get_field("f%d" % d)[:, :, :grid[2]])


@pytest.mark.parametrize("grid", GRIDS)
def test_nssolute_solver_euler(SPM, grid):
    objs = pair(SPM, grid, 5, 4.0, electrolyte())
    o, g = objs
    nz = grid[2]
    for d in range(3):
        o.put_field("u%d" % d, smooth_field(o.shape, nz, seed=60 + d, kmax=4, scale=0.02))
    o.U2zeta_k()
    z0, z1, dc = o.get_field("zeta0"), o.get_field("zeta1"), o.get_uk_dc()
    g.put_field("zeta0", z0)
    g.put_field("zeta1", z1)
    g.put_uk_dc(dc)
    perturb(objs, nz, seed=70)
    for obj in objs:
        obj.NSsolute_solver_Euler()
    check("zeta0", g.get_field("zeta0"), o.get_field("zeta0"))
    check("zeta1", g.get_field("zeta1"), o.get_field("zeta1"))
    for s in range(2):
        check("conc%d" % s, g.get_concentration(s), o.get_concentration(s))
    assert np.abs(g.get_uk_dc() - o.get_uk_dc()).max() <= 1e-12 * max(1.0, np.abs(o.get_uk_dc()).max())


@pytest.mark.parametrize("grid,n_spec,ac", [((32, 32, 32), 2, 0), ((32, 64, 32), 2, 1), ((32, 32, 32), 1, 0)])
def test_electrolyte_steps(SPM, grid, n_spec, ac):
    el = electrolyte(n_spec, ac=ac, angular_frequency=0.7,
                     valency=(1.0, -1.0) if n_spec == 2 else (-1.0, 0.0))
    o, g = pair(SPM, grid, 6, 4.0, el, vel=0.02, omega=0.005)
    o.step(3)
    g.step(3)
    nz = grid[2]
    check("zeta0", g.get_field("zeta0"), o.get_field("zeta0"))
    check("zeta1", g.get_field("zeta1"), o.get_field("zeta1"))
    for d in range(3):  # U2zeta_k leaves u as a spectrum
        check("u%d" % d, g.get_field("u%d" % d), o.get_field("u%d" % d))
    check("phi", g.get_field("phi")[:, :, :nz], o.get_field("phi")[:, :, :nz])
    for s in range(o.n_spec):
        check("conc%d" % s, g.get_concentration(s), o.get_concentration(s))
    pa, pb = g.get_particles(), o.get_particles()
    for k in ("x", "v", "omega", "f_hydro_previous", "torque_hydro_previous", "f_hydro1", "torque_hydro1"):
        check("particle " + k, pa[k], pb[k])
    assert g.get_time() == o.get_time()


def test_electrolyte_rejects_bad_use(SPM):
    from kapsel_old_amd.spm import SpmError
    prm = abi.make_params(32, a=4.0, n_particles=2)
    g = SPM(prm)
    with pytest.raises(SpmError):
        g.NSsolute_solver_Euler()  # before spm_set_electrolyte
    with pytest.raises(SpmError):
        g.set_electrolyte(electrolyte(dielectric=0.0))
    g.set_electrolyte(electrolyte())
    with pytest.raises(SpmError):
        g.set_electrolyte(electrolyte())  # twice


@pytest.mark.parametrize("e_ext", [(0.0, 0.0, 0.0), (0.01, 0.0, 0.0)], ids=["no_field", "field"])
def test_poisson_boltzmann_start_matches_oracle(SPM, e_ext):
    """Init_rho_ion with Poisson_Boltzmann on (uniform start relaxed by
    Set_steadystate_ion_density until the free energy settles): the same
    iteration count and ion spectra as the oracle, then three steps."""
    grid, n, a = (32, 32, 32), 3, 4.0
    prm = abi.make_params(*grid, a=a, n_particles=n, rotation=1, dt=0.05, kbt=1.0, fix_cell=(0, 0, 0))
    p = particles(n, grid, a, seed=7, vel=0.01, omega=0.002)
    el = electrolyte(onsager=(1.0, 1.0), dielectric=1.3, e_ext=e_ext)
    objs, its = [], []
    for cls in (Oracle, SPM):
        o = cls(prm)
        o.put_particles(p)
        o.init_fluid()
        o.set_electrolyte(el)
        its.append(o.Init_rho_ion_PB())
        objs.append(o)
    o, g = objs
    print("iterations oracle %d gpu %d" % tuple(its), flush=True)
    assert its[0] == its[1] and 10 < its[0] < 9999
    for s in range(2):
        check("conc%d" % s, g.get_concentration(s), o.get_concentration(s))
    check("total_solute", g.get_total_solute(), o.get_total_solute())
    for step in range(3):
        o.step(1)
        g.step(1)
        for s in range(2):
            check("step %d conc%d" % (step, s), g.get_concentration(s), o.get_concentration(s))
        check("step %d zeta0" % step, g.get_field("zeta0"), o.get_field("zeta0"))


def test_poisson_boltzmann_start_rejects_unbuilt_cases(SPM):
    from kapsel_old_amd.spm import SpmError
    prm = abi.make_params(32, a=4.0, n_particles=1, rotation=1, dt=0.05, kbt=1.0, fix_cell=(0, 0, 0))
    g = SPM(prm)
    g.put_particles(particles(1, (32, 32, 32), 4.0, seed=2))
    g.init_fluid()
    g.set_electrolyte(electrolyte())
    with pytest.raises(SpmError):
        g.Init_rho_ion_PB()  # one particle: the iterative profile is not built
