"""The oracle's Lees-Edwards (oblique-coordinate) fluid step,
Time_evolution_hydro_OBL (sp_3d_ns.cxx:214-390) with no particles, pinned
by what the mathematics fixes:

- with no shear and degree_oblique = 0 the oblique metric is the Cartesian
  one, so the step must move the velocity exactly as the Navier-Stokes step
  does (which is pinned on its own), although zeta is stored with another
  choice of components;
- U_k2zeta_k_OBL followed by Zeta_k2u_k_OBL returns the metric-solenoidal
  part of a contravariant spectrum, k_i u^i = 0, at any degree_oblique;
- a fluid at rest stays at rest under shear, the measured shear rate of the
  imposed profile is the shear rate, and degree_oblique wraps at 1.
"""
import numpy as np
import pytest

from kapsel_old_amd import abi
from oracle.oracle import Oracle
from tests.helpers import rel, smooth_field

N = 16


def velocity(o, oblique):
    """Real-space velocity of the current zeta."""
    if oblique:
        o.Zeta_k2u_k_OBL()
        return [o.A_k2a(o.get_field("u%d" % d)) for d in range(3)]
    o.Zeta_k2u()
    return [o.get_field("u%d" % d) for d in range(3)]


def start(prm, shear_rate=0.0, seed=3, scale=0.3):
    """A Navier-Stokes oracle and a Lees-Edwards one holding the same
    solenoidal velocity field."""
    a, b = Oracle(prm), Oracle(prm)
    b.set_lees_edwards(shear_rate)
    for d in range(3):
        a.put_field("u%d" % d, smooth_field(a.shape, N, seed=seed + d, kmax=3, scale=scale))
    a.U2zeta_k()
    a.Zeta_k2u()
    u = [a.get_field("u%d" % d) for d in range(3)]
    a.U2zeta_k()
    for d in range(3):
        b.put_field("u%d" % d, b.A2a_k(u[d]))
    b.U_k2zeta_k_OBL()
    return a, b


@pytest.mark.parametrize("fix_cell", [(0, 0, 0), (1, 1, 1)])
def test_unsheared_oblique_step_is_the_navier_stokes_step(fix_cell):
    prm = abi.make_params(N, dt=0.05, eta=0.5, fix_cell=fix_cell)
    a, b = start(prm)
    ua0, ub0 = velocity(a, False), velocity(b, True)
    for d in range(3):
        assert rel(ub0[d][:, :, :N], ua0[d][:, :, :N]) < 1e-14
    for _ in range(3):
        # Calc_shear_rate_eff measures the mean dux/dy of the fluid itself
        # (the wrap-around pair left out), which feeds the next step's
        # shear term; hold it at zero to compare with the unsheared solver
        b.put_oblique(0.0, 0.0)
        a.step(1)
        b.step(1)
    ua, ub = velocity(a, False), velocity(b, True)
    for d in range(3):
        e = rel(ub[d][:, :, :N], ua[d][:, :, :N])
        assert e < 1e-12, (d, e)
        assert rel(ua[d][:, :, :N], ua0[d][:, :, :N]) > 1e-3  # the flow did change
    # uk_dc holds only rounding noise here (|uk_dc| ~ 1e-13 of a field of order 1)
    np.testing.assert_allclose(b.get_uk_dc(), a.get_uk_dc(), rtol=0, atol=1e-13)
    assert b.get_oblique()[0] == 0.0


@pytest.mark.parametrize("deg", [0.0, 0.3, 0.75])
def test_oblique_zeta_round_trip_is_the_metric_projection(deg):
    prm = abi.make_params(N, dt=0.05)
    o = Oracle(prm)
    o.set_lees_edwards(0.0)
    o.put_oblique(deg, 0.0)
    rng = np.random.default_rng(11)
    u = [o.A2a_k(smooth_field(o.shape, N, seed=20 + d, kmax=4)) for d in range(3)]
    for d in range(3):
        o.put_field("u%d" % d, u[d])
    o.U_k2zeta_k_OBL()
    o.Zeta_k2u_k_OBL()
    back = [o.get_field("u%d" % d) for d in range(3)]
    # the projection u - k^i (k_j u^j) / K2 with k^i = g^ij k_j, K2 = k_i k^i
    nx, ny, nz_ = o.shape
    kx = np.fft.fftfreq(nx, 1.0 / nx)[:, None, None] * 2 * np.pi / nx
    ky = np.fft.fftfreq(ny, 1.0 / ny)[None, :, None] * 2 * np.pi / ny
    kz = (np.arange(nz_) // 2)[None, None, :] * 2 * np.pi / N
    kxc, kyc = (1 + deg * deg) * kx - deg * ky, -deg * kx + ky
    k2 = kx * kx + (ky - deg * kx) ** 2 + kz * kz
    ik2 = np.where(k2 > 0, 1 / np.where(k2 > 0, k2, 1), 0)
    dot = u[0] * kx + u[1] * ky + u[2] * kz
    want = [u[0] - ik2 * dot * kxc, u[1] - ik2 * dot * kyc, u[2] - ik2 * dot * kz]
    scale = max(np.abs(w).max() for w in want)
    for d in range(3):
        diff = back[d] - want[d]
        diff[0, 0, :2] = 0  # mode 0 carries uk_dc (checked below)
        assert np.abs(diff).max() < 1e-13 * scale, d
    div = back[0] * kx + back[1] * ky + back[2] * kz
    assert np.abs(div).max() < 1e-13 * scale
    # uk_dc is the contravariant mean flow of the input
    np.testing.assert_allclose([back[d][0, 0, 0] for d in range(3)], [u[d][0, 0, 0] for d in range(3)],
                               rtol=1e-14, atol=1e-16)


def test_fluid_at_rest_under_shear():
    prm = abi.make_params(N, dt=0.1, fix_cell=(0, 0, 0))
    o = Oracle(prm)
    o.set_lees_edwards(1.0)
    degs = []
    for _ in range(14):  # degree_oblique grows by 0.1 a step and wraps at 1
        o.step(1)
        degs.append(o.get_oblique()[0])
        assert o.get_oblique()[1] == pytest.approx(1.0, rel=1e-14)
    assert not o.get_field("zeta0").any() and not o.get_field("zeta1").any()
    assert not o.get_uk_dc().any()
    want, d = [], 0.0
    for _ in range(14):  # degree_oblique += Shear_rate_eff * dt; -= 1 once it reaches 1
        d += 1.0 * 0.1
        if d >= 1.0:
            d -= 1.0
        want.append(d)
    assert degs == want and any(b < a for a, b in zip(degs, degs[1:]))  # it wrapped


def _shearing_wave(dt, shear_rate, t_end=0.4, nu=0.01, kx=1, ky=2):
    """Amplitude ratio of the vorticity of one 2-D Fourier mode after t_end,
    and exp(-nu int K2 dt) over the same steps.  In a uniform shear flow a
    single 2-D mode is an exact nonlinear solution whose vorticity is carried
    by the mean flow (the wavevector tilts, kx fixed, ky - t shear kx) and
    decays as exp(-nu int |k|^2); in oblique coordinates the mode keeps its
    indices."""
    prm = abi.make_params(N, dt=dt, eta=nu, fix_cell=(0, 0, 0))
    o = Oracle(prm)
    o.set_lees_edwards(shear_rate)
    x = np.arange(N)[:, None, None]
    y = np.arange(N)[None, :, None]
    th = 2 * np.pi * (kx * x + ky * y) / N
    u = [np.zeros(o.shape) for _ in range(3)]
    u[0][:, :, :N] = -ky * 1e-3 * np.sin(th)
    u[1][:, :, :N] = kx * 1e-3 * np.sin(th)
    for d in range(3):
        o.put_field("u%d" % d, o.A2a_k(u[d]))
    o.U_k2zeta_k_OBL()
    w = 2 * np.pi / N

    def amp():  # kx != 0: zeta = (omega_x, omega_y, omega_z)[1:], so zeta1 is omega_z
        z = o.get_field("zeta1")
        return np.hypot(z[kx, ky, 0], z[kx, ky, 1])

    a0, integ = amp(), 0.0
    for _ in range(int(round(t_end / dt))):
        deg = o.get_oblique()[0]
        integ += nu * ((w * kx) ** 2 + (w * ky - w * deg * kx) ** 2) * dt
        o.step(1)
    return amp() / a0, np.exp(-integ)


def test_shearing_wave_vorticity_is_carried_by_the_shear():
    # no shear: the slaved Euler step is exact for a single mode
    r, e = _shearing_wave(0.01, 0.0)
    assert abs(r / e - 1) < 1e-12
    # with shear the step converges to the exact decay at first order in dt
    # (the shear term of Zeta_k2advection_k_OBL is explicit)
    err = [abs(_shearing_wave(dt, 0.5)[0] / _shearing_wave(dt, 0.5)[1] - 1) for dt in (0.01, 0.005, 0.0025)]
    assert err[0] < 1e-2
    assert 0.4 < err[1] / err[0] < 0.6 and 0.4 < err[2] / err[1] < 0.6, err
