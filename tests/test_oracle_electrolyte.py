"""Pins of the oracle's electrolyte restatement (operate_electrolyte.cxx,
solute_rhs.cxx, fluid_solver.cxx:181-326) against what the physics fixes,
computed here with numpy's own FFT conventions:
  - the surface-charge shell carries exactly the particle charge and is
    centred on r = a; the surface normals are unit radial vectors;
  - the uniform salt start is electroneutral and its bulk densities give
    back the Debye length (Debye-Hueckel);
  - the Coulomb body force equals rho_e (E_ext - grad Phi) with
    -lap Phi = rho_e / eps solved independently;
  - the Nernst-Planck update equals dt * div(kBT G grad C + G z e C grad Phi
    - C u) with the impermeability projection on the shell, computed
    independently; the solvent update of a shear mode is dt nu k^2 decay.
No reference build exists for these routines (they include udfmanager.h),
so the coupled electrolyte trajectory is parity unpinned beyond them.
"""
import numpy as np
import pytest

from kapsel_old_amd import abi
from oracle.oracle import Oracle
from tests.helpers import particles, smooth_field

N = 32


def make(n=3, a=4.0, n_spec=2, q=5.0, e_ext=(0.01, -0.02, 0.005), kbt=1.0, seed=7, vel=0.0):
    prm = abi.make_params(N, a=a, n_particles=n, rotation=1, dt=0.05, kbt=kbt, fix_cell=(0, 0, 0))
    o = Oracle(prm)
    p = particles(n, (N, N, N), a, seed=seed, vel=vel)
    o.put_particles(p)
    o.init_fluid()
    el = abi.make_electrolyte(n_spec=n_spec, valency=(1.0, -1.0) if n_spec == 2 else (-1.0, 0.0),
                              onsager=(0.1, 0.2), dielectric=1.3, debye_length=3.0, surface_charge=(q,),
                              e_ext=e_ext)
    o.set_electrolyte(el)
    o.Init_rho_ion()
    return o, p, el


def spectrum_to_real(a_k):
    """Ooura half spectrum (X = sum a e^{+ikx}) -> real field with numpy."""
    nz = N
    X = a_k[:, :, 0::2] - 1j * a_k[:, :, 1::2]  # numpy's e^{-ikx} spectrum
    return np.fft.irfftn(X, s=(N, N, nz), axes=(0, 1, 2))


def wavenumbers():
    """Wavenumbers with the reference's Nyquist convention: index N/2 is
    +N/2 (Calc_KX_Ooura, fft_wrapper.h:687-699), where numpy says -N/2."""
    i = np.arange(N)
    k = 2 * np.pi / N * np.where(i > N // 2, i - N, i)
    kz = 2 * np.pi / N * np.arange(N // 2 + 1)
    return np.meshgrid(k, k, kz, indexing="ij")


def ooura_planes(F):
    """The reference inverse's reading of the kz = 0 and kz = N/2 planes
    (rdft3dsort(-1), fftsg3d.c:645-668): one member of each pair (k, -k)
    kept, the other its conjugate, self-conjugate points real.  The rule is
    pinned against the reference build in test_oracle_fft.py."""
    F = F.copy()
    n1h = n2h = N // 2
    for plane, kz in ((0, 0), (1, N // 2)):
        for i in range(N):
            for j in range(N):
                edge = j in (0, n2h)
                selfc = edge and i in (0, n1h)
                keep = ((1 <= j < n2h) or (edge and i <= n1h)) if plane == 0 else (j > n2h or (edge and i > n1h) or selfc)
                if selfc:
                    F[i, j, kz] = F[i, j, kz].real
                elif not keep:
                    F[i, j, kz] = np.conj(F[(N - i) % N, (N - j) % N, kz])
    return F


def grad(f):
    kx, ky, kz = wavenumbers()
    F = np.fft.rfftn(f)
    return [np.fft.irfftn(1j * kk * F, s=f.shape, axes=(0, 1, 2)) for kk in (kx, ky, kz)]


def min_image(d):
    return d - np.round(d / N) * N


def grid_r(xp):
    ix = np.arange(N)
    X, Y, Z = np.meshgrid(ix, ix, ix, indexing="ij")
    return [min_image(X - xp[0]), min_image(Y - xp[1]), min_image(Z - xp[2])]


def test_surface_shell_carries_the_charge_at_radius_a():
    o, p, _ = make(n=1, q=-7.0)
    _, surf = o.make_phi_qq()
    s = surf[:, :, :N]
    assert abs(s.sum() - (-7.0)) < 1e-10
    assert abs(np.abs(s).sum() - 7.0) < 1e-10
    r = np.sqrt(sum(x ** 2 for x in grid_r(p["x"][0])))
    r_mean = (s * r).sum() / s.sum()
    assert abs(r_mean - 4.0) < 0.1, r_mean
    assert np.all(s[np.abs(r - 4.0) >= 1.0 + 1e-9] == 0.0)  # shell |r - a| < xi/2


def test_surface_normals_are_unit_radial():
    o, p, _ = make(n=1)
    nrm = o.make_surface_normal()
    n = np.stack([x[:, :, :N] for x in nrm])
    rv = np.stack(grid_r(p["x"][0]))
    r = np.sqrt((rv ** 2).sum(0))
    on = np.abs(r - 4.0) < 1.0 - 1e-9
    assert on.sum() > 100
    assert np.allclose(np.sqrt((n ** 2).sum(0))[on], 1.0, atol=1e-12)
    assert np.allclose((n * rv).sum(0)[on], r[on], atol=1e-12)
    assert np.all(n[:, np.abs(r - 4.0) > 1.0 + 1e-9] == 0.0)


def test_uniform_salt_start_is_neutral_and_gives_the_debye_length():
    o, p, el = make(n=3, q=5.0)
    phi, surf = o.make_phi_qq()
    c = [spectrum_to_real(o.get_concentration(s)) for s in range(2)]
    solid = 1.0 - phi[:, :, :N]
    ion_charge = ((c[0] - c[1]) * solid).sum()
    assert abs(ion_charge + surf[:, :, :N].sum()) < 1e-9 * abs(surf[:, :, :N].sum())
    # far from the particles: bulk co-ion density n_inf with 2 n_inf e^2 / (eps kBT) = 1/lambda^2
    volume = solid.sum()
    counter = 3 * 5.0 / volume
    n_inf = c[0].mean() if False else (c[0][solid > 1 - 1e-12].mean())
    lam = np.sqrt(1.3 * 1.0 / (2 * n_inf * 1.0))
    assert abs(lam - 3.0) < 1e-12, lam
    assert abs(c[1][solid > 1 - 1e-12].mean() - (n_inf + counter)) < 1e-12


def perturbed(o, seed):
    for s in range(o.n_spec):
        d = o.A2a_k(smooth_field(o.shape, N, seed=seed + s, kmax=4, scale=1e-3))
        o.put_concentration(s, o.get_concentration(s) + d)


def charge_density(o):
    phi, surf = o.make_phi_qq()
    rho = surf[:, :, :N].copy()
    val = [1.0, -1.0]
    for s in range(o.n_spec):
        rho += val[s] * spectrum_to_real(o.get_concentration(s)) * (1.0 - phi[:, :, :N])
    return rho, phi[:, :, :N]


def poisson_grad(rho, eps):
    kx, ky, kz = wavenumbers()
    k2 = kx ** 2 + ky ** 2 + kz ** 2
    R = np.fft.rfftn(rho)
    P = np.where(k2 > 0, R / np.where(k2 > 0, k2, 1.0) / eps, 0.0)
    return [np.fft.irfftn(ooura_planes(1j * kk * P), s=rho.shape, axes=(0, 1, 2)) for kk in (kx, ky, kz)]


def test_coulomb_force_is_rho_times_field():
    e_ext = (0.01, -0.02, 0.005)
    o, p, el = make(n=3, e_ext=e_ext)
    perturbed(o, 11)
    rho, _ = charge_density(o)
    gphi = poisson_grad(rho, 1.3)
    o.Make_Coulomb_force_x_on_fluid()
    for d in range(3):
        want = rho * (e_ext[d] - gphi[d])
        got = o.get_field("f%d" % d)[:, :, :N]
        assert np.abs(got - want).max() < 1e-10 * np.abs(want).max(), d
    # a plausible slip (sign of the potential) fails
    want_bad = rho * (e_ext[0] + gphi[0])
    assert np.abs(o.get_field("f0")[:, :, :N] - want_bad).max() > 1e-3 * np.abs(want_bad).max()


def truncate_real(f):
    """2/3-rule truncation of a real field (numpy spectrum)."""
    F = np.fft.rfftn(f)
    trn = (N + 2) // 3
    i = np.fft.fftfreq(N, d=1.0 / N).astype(int)
    kz = np.arange(N // 2 + 1)
    keep = (np.abs(i)[:, None, None] < trn) & (np.abs(i)[None, :, None] < trn) & (kz[None, None, :] < trn)
    return np.fft.irfftn(np.where(keep, F, 0), s=f.shape, axes=(0, 1, 2)), keep


def test_nernst_planck_update_is_the_flux_divergence():
    """Fluid at rest: dC/dt = div(kBT G grad C + G z e C grad Phi) with the
    flux projected tangential on the particle shells."""
    o, p, el = make(n=3, kbt=1.0)
    perturbed(o, 21)
    c_before = [o.get_concentration(s) for s in range(2)]
    # the ion potential uses the truncated charge (fluid_solver.cxx:263-274)
    rho, _ = charge_density(o)
    kx, ky, kz = wavenumbers()
    k2 = kx ** 2 + ky ** 2 + kz ** 2
    trn = (N + 2) // 3
    ii = np.fft.fftfreq(N, d=1.0 / N).astype(int)
    keep = (np.abs(ii)[:, None, None] < trn) & (np.abs(ii)[None, :, None] < trn) & (np.arange(N // 2 + 1)[None, None, :] < trn)
    R = np.fft.rfftn(rho)
    P = np.where((k2 > 0) & keep, R / np.where(k2 > 0, k2, 1.0) / 1.3, 0.0)
    gphi = [np.fft.irfftn(1j * kk * P, s=rho.shape, axes=(0, 1, 2)) for kk in (kx, ky, kz)]
    e_ext = (0.01, -0.02, 0.005)
    gphi = [gphi[d] - e_ext[d] for d in range(3)]  # Add_constant_field_k: grad Phi - E_ext
    nrm = np.stack([x[:, :, :N] for x in o.make_surface_normal()])
    n2 = (nrm ** 2).sum(0)
    dt = 0.05
    o.NSsolute_solver_Euler()
    for s, (z, g) in enumerate([(1.0, 0.1), (-1.0, 0.2)]):
        c = spectrum_to_real(c_before[s])
        c, _ = truncate_real(c)
        gc = grad(1.0 * g * c)
        flux = [gc[d] + c * (g * (gphi[d] * z)) for d in range(3)]
        fn = sum(flux[d] * nrm[d] for d in range(3))
        proj = np.where(n2 > 0, fn / np.where(n2 > 0, n2, 1.0), 0.0)
        flux = [flux[d] - nrm[d] * proj for d in range(3)]
        div = sum(np.fft.irfftn(1j * kk * np.fft.rfftn(flux[d]), s=c.shape, axes=(0, 1, 2)) for d, kk in enumerate((kx, ky, kz)))
        want = np.fft.rfftn(div)
        got = (spectrum_to_real(o.get_concentration(s)) - spectrum_to_real(c_before[s])) / dt
        got_k = np.fft.rfftn(got)
        err = np.abs(np.where(keep, got_k - want, 0)).max()
        assert err < 1e-9 * np.abs(np.where(keep, want, 0)).max(), (s, err)


def test_solvent_update_of_a_shear_mode_is_viscous_decay():
    """u = (A sin(2 pi y / L), 0, 0) has no advection; NSsolute_solver_Euler
    advances its vorticity by an explicit Euler step of nu lap."""
    o, p, el = make(n=1, q=1e-9, e_ext=(0.0, 0.0, 0.0))
    A, m = 0.01, 2
    u0 = np.zeros(o.shape)
    y = np.arange(N)[None, :, None]
    u0[:, :, :N] = A * np.sin(2 * np.pi * m * y / N)
    o.put_field("u0", u0)
    o.put_field("u1", np.zeros(o.shape))
    o.put_field("u2", np.zeros(o.shape))
    o.U2zeta_k()
    z_before = [o.get_field("zeta0"), o.get_field("zeta1")]
    o.NSsolute_solver_Euler()
    k2 = (2 * np.pi * m / N) ** 2
    for d in range(2):
        want = z_before[d] * (1.0 - 0.05 * 1.0 * k2)
        assert np.abs(o.get_field("zeta%d" % d) - want).max() < 1e-12 * max(1.0, np.abs(want).max())


def _pb_system(tol, pb=True):
    prm = abi.make_params(N, a=4.0, n_particles=3, rotation=1, dt=0.05, kbt=1.0, fix_cell=(0, 0, 0))
    o = Oracle(prm)
    p = particles(3, (N, N, N), 4.0, seed=7)
    o.put_particles(p)
    o.init_fluid()
    o.set_electrolyte(abi.make_electrolyte(n_spec=2, onsager=(1.0, 1.0), dielectric=1.3, debye_length=3.0,
                                           surface_charge=(5.0,)))
    it = o.Init_rho_ion_PB(tol) if pb else o.Init_rho_ion()
    return o, it


def _boltzmann_residual(o):
    """std over the fluid of ln c+ + e Phi / kBT and ln c- - e Phi / kBT
    (zero at Poisson-Boltzmann equilibrium), and the std of e Phi / kBT,
    with Phi solved by numpy from the total charge."""
    rho, phi = charge_density(o)
    kx, ky, kz = wavenumbers()
    k2 = kx ** 2 + ky ** 2 + kz ** 2
    P = np.where(k2 > 0, np.fft.rfftn(rho) / np.where(k2 > 0, k2, 1.0) / 1.3, 0.0)
    pot = np.fft.irfftn(P, s=rho.shape, axes=(0, 1, 2))
    c = [spectrum_to_real(o.get_concentration(s)) for s in range(2)]
    m = phi == 0.0
    return (np.log(c[0][m]) + pot[m]).std(), (np.log(c[1][m]) - pot[m]).std(), pot[m].std()


def test_poisson_boltzmann_start_reaches_boltzmann_equilibrium():
    """Set_steadystate_ion_density (operate_electrolyte.cxx:374-412) pinned
    by physics: relaxed far enough, the ions obey Boltzmann's law in the
    Poisson potential of the total charge, c+- ~ exp(-+ e Phi / kBT), to
    a residual of 2% of the potential's variation (the uniform start is
    100% off); at the reference's stopping rule (free energy change below
    1e-4) the residual is already below half the uniform start's.  The
    relaxation keeps the total charge neutral."""
    o_u, _ = _pb_system(0, pb=False)
    ru = _boltzmann_residual(o_u)
    o_ref, it_ref = _pb_system(1e-4)
    rr = _boltzmann_residual(o_ref)
    o_eq, it_eq = _pb_system(1e-12)
    re = _boltzmann_residual(o_eq)
    print("iterations %d / %d; residuals uniform %.2e, reference stop %.2e, converged %.2e (pot std %.2e)"
          % (it_ref, it_eq, ru[0], rr[0], re[0], re[2]), flush=True)
    assert abs(ru[0] - ru[2]) < 1e-12 * ru[2]  # uniform: the residual is the potential itself
    assert rr[0] < 0.5 * ru[0] and rr[1] < 0.5 * ru[1]
    assert re[0] < 0.02 * re[2] and re[1] < 0.02 * re[2]
    assert 10 < it_ref < it_eq < 9999
    rho, _ = charge_density(o_eq)
    phi_qq, surf = o_eq.make_phi_qq()
    assert abs(rho.sum()) < 1e-9 * abs(surf[:, :, :N].sum())
