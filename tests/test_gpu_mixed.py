"""Grids of 48, 80, 96, 160 and 192 points per axis (factors 2, 3, 5): the
mixed-radix transforms, alone through spm_fft3d and inside the literal SPM
step, against the oracle (whose transform for these lengths is the DFT by
its definition, pinned to numpy in tests/test_oracle_fft.py).  The reference
runs powers of two only (input.cxx:1397-1402); these grids are an extension,
and the step is the reference's algorithm on them."""
import numpy as np
import pytest

from kapsel_old_amd import abi
from oracle.oracle import Oracle
from tests.helpers import particles, rel
from tests.test_gpu_fft3d import _fft3d

pytestmark = pytest.mark.gpu

TOL = 1e-10

SHAPES = [(48, 48, 48), (80, 16, 96), (16, 96, 80), (160, 8, 16), (8, 192, 8), (8, 8, 192), (96, 80, 64)]


@pytest.mark.parametrize("shape", SHAPES)
def test_mixed_forward_matches_oracle(shape):
    nx, ny, nz = shape
    o = Oracle(abi.make_params(nx, ny, nz))
    rng = np.random.default_rng(nx + 5 * nz)
    a = np.zeros(o.shape)
    a[:, :, :nz] = rng.standard_normal((nx, ny, nz))
    want = o.A2a_k(a)
    got = _fft3d(shape, +1, a)
    e = np.abs(got - want).max() / np.abs(want).max()
    print(shape, "forward rel err %.3e" % e, flush=True)
    assert e < 1e-13


@pytest.mark.parametrize("shape", SHAPES)
def test_mixed_inverse_of_a_non_hermitian_spectrum_matches_oracle(shape):
    nx, ny, nz = shape
    o = Oracle(abi.make_params(nx, ny, nz))
    s = np.random.default_rng(ny + 3).standard_normal(o.shape)
    want = o.A_k2a(s.copy())[:, :, :nz]
    got = _fft3d(shape, -1, s)[:, :, :nz]
    e = np.abs(got - want).max() / np.abs(want).max()
    print(shape, "inverse rel err %.3e" % e, flush=True)
    assert e < 1e-13


def _step_pair(SPM, grid, n, a, kbt, seed=12345):
    nx, ny, nz = grid
    prm = abi.make_params(nx, ny, nz, a=a, n_particles=n, kbt=kbt, lj_truncate=abi.LJ_WCA, seed=seed, rotation=1)
    o, g = Oracle(prm), SPM(prm)
    p = particles(n, grid, a, seed=29, min_factor=2.05)
    for obj in (o, g):
        obj.put_particles(p)
        obj.init_fluid()
    return o, g


@pytest.mark.parametrize("grid,n,a,kbt,steps", [
    ((48, 48, 48), 6, 4.0, 0.05, 3),
    ((80, 48, 96), 10, 5.0, 0.0, 2),
    ((16, 160, 48), 3, 3.0, 0.1, 2),
])
def test_mixed_grid_steps_match_oracle(grid, n, a, kbt, steps):
    from kapsel_old_amd.spm import SPM
    nz = grid[2]
    o, g = _step_pair(SPM, grid, n, a, kbt)
    for s in range(steps):
        for obj in (o, g):
            obj.step(1)
        for nm in ("u0", "u1", "u2", "zeta0", "zeta1", "phi"):
            e = rel(g.get_field(nm), o.get_field(nm))
            print(grid, "step", s, nm, "rel err %.3e" % e, flush=True)
            assert e < TOL, (nm, e)
        pg, po = g.get_particles(), o.get_particles()
        for k in ("x", "v", "omega", "f_hydro_previous", "torque_hydro_previous"):
            e = rel(pg[k], po[k])
            assert e < TOL, (k, e)
    assert g.get_time() == o.get_time()


def test_mixed_grid_limits():
    from kapsel_old_amd.spm import SPM, SpmError
    g = SPM(abi.make_params(48, 48, 48))
    with pytest.raises(SpmError):
        g.set_schedule(0)
    with pytest.raises(SpmError):
        g.set_lees_edwards(0.01)
    g.set_schedule(1)
