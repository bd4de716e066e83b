"""Pins the oracle's spectral operators against identities computed
independently with numpy (spectral derivatives in numpy's own sign
convention, mapped to the reference's exp(+i k.x) spectra).

  2/3 truncation   (fft_wrapper.h:723-766)   products of truncated fields are
                                             alias-free on the kept modes
  Zeta_k2u_k / U_k2zeta_k (fft_wrapper.cxx)  zeta = the reference's two
                                             components of curl u; round trip
  Solenoidal_uk    (operate_omega.cxx:186)   k . u = 0, idempotent
  U2advection_k    (operate_omega.cxx:14)    = curl(u x omega); 0 for ABC flow
  NS_solver_slavedEuler (fluid_solver.cxx:328) exact viscous decay of a
                                             shear mode; with advection live,
                                             the slaved-Euler closed form
                                             e w + (1 - e) curl(u x w)/(nu k^2)
"""
import numpy as np
import pytest

from kapsel_old_amd import abi
from oracle.oracle import Oracle


def np_to_ooura(X, shape):
    out = np.zeros(shape)
    out[:, :, 0::2] = X.real
    out[:, :, 1::2] = -X.imag
    return out


def ooura_to_np(a):
    return a[:, :, 0::2] - 1j * a[:, :, 1::2]


def kvecs(n):
    """numpy-convention integer wavenumbers on the rfftn half grid, with the
    reference's Nyquist choice (+N/2) for x and y."""
    k = np.arange(n)
    kk = np.where(k > n // 2, k - n, k)
    kz = np.arange(n // 2 + 1)
    return np.meshgrid(kk, kk, kz, indexing="ij")


def band_limited_solenoidal(n, kmax, seed):
    rng = np.random.default_rng(seed)
    KX, KY, KZ = kvecs(n)
    mask = (np.abs(KX) <= kmax) & (np.abs(KY) <= kmax) & (KZ <= kmax)
    U = [np.where(mask, rng.standard_normal(KX.shape) + 1j * rng.standard_normal(KX.shape), 0) for _ in range(3)]
    k2 = (KX ** 2 + KY ** 2 + KZ ** 2).astype(float)
    k2[0, 0, 0] = 1
    div = (KX * U[0] + KY * U[1] + KZ * U[2]) / k2
    U = [U[0] - div * KX, U[1] - div * KY, U[2] - div * KZ]
    u = [np.fft.irfftn(x, s=(n, n, n), axes=(0, 1, 2)) for x in U]
    return u  # real, periodic, divergence-free


def spectral_grad(f, n):
    """numpy spectral derivatives of a real field on [0, n) with L = n."""
    F = np.fft.rfftn(f)
    KX, KY, KZ = kvecs(n)
    w = 2 * np.pi / n
    return [np.fft.irfftn(1j * w * K * F, s=f.shape, axes=(0, 1, 2)) for K in (KX, KY, KZ)]


def curl(u, n):
    g = [spectral_grad(c, n) for c in u]  # g[i][j] = d u_i / d x_j
    return [g[2][1] - g[1][2], g[0][2] - g[2][0], g[1][0] - g[0][1]]


def zeta_components(w_hat_oo, KX, KY):
    """The reference's reduced vorticity (fft_wrapper.cxx:298-313): (w_y, w_z)
    if kx != 0; (w_z, w_x) if kx == 0, ky != 0; else (w_x, w_y)."""
    z0 = np.where(KX != 0, w_hat_oo[1], np.where(KY != 0, w_hat_oo[2], w_hat_oo[0]))
    z1 = np.where(KX != 0, w_hat_oo[2], np.where(KY != 0, w_hat_oo[0], w_hat_oo[1]))
    return z0, z1


def to_layout(u, shape, n):
    a = np.zeros(shape)
    a[:, :, :n] = u
    return a


def test_truncation_removes_aliasing_of_products():
    n = 16
    o = Oracle(abi.make_params(n))
    rng = np.random.default_rng(5)
    m = 2 * n

    def upsample(a):
        A = np.fft.rfftn(a)
        B = np.zeros((m, m, m // 2 + 1), dtype=complex)
        ix = np.r_[0:n // 2, m - n // 2:m]
        B[np.ix_(ix, ix, np.arange(n // 2 + 1))] = A[np.ix_(np.r_[0:n // 2, n // 2:n], np.r_[0:n // 2, n // 2:n],
                                                          np.arange(n // 2 + 1))]
        return np.fft.irfftn(B, s=(m, m, m)) * (m / n) ** 3

    fields = []
    for _ in range(2):
        a = np.zeros(o.shape)
        a[:, :, :n] = rng.standard_normal((n, n, n))
        fields.append(o.A_k2a(o.truncate(o.A2a_k(a)))[:, :, :n])
    prod = to_layout(fields[0] * fields[1], o.shape, n)
    P = o.truncate(o.A2a_k(prod))
    exact = np.fft.rfftn(upsample(fields[0]) * upsample(fields[1]))
    ix = np.r_[0:n // 2, m - n // 2:m]
    exact = exact[np.ix_(ix, ix, np.arange(n // 2 + 1))] * (n / m) ** 3
    keep = ooura_to_np(o.truncate(np.ones(o.shape))) != 0
    Pn = ooura_to_np(P)
    assert np.abs(Pn[keep] - exact[keep]).max() < 1e-10 * np.abs(exact).max()
    assert np.all(Pn[~keep] == 0)
    # the same comparison fails without the truncation of the factors
    raw = [rng.standard_normal((n, n, n)) for _ in range(2)]
    Praw = ooura_to_np(o.truncate(o.A2a_k(to_layout(raw[0] * raw[1], o.shape, n))))
    ex_raw = np.fft.rfftn(upsample(raw[0]) * upsample(raw[1]))[np.ix_(ix, ix, np.arange(n // 2 + 1))] * (n / m) ** 3
    assert np.abs(Praw[keep] - ex_raw[keep]).max() > 1e-3 * np.abs(ex_raw).max()


def test_zeta_is_the_reduced_curl_and_round_trips():
    n = 16
    o = Oracle(abi.make_params(n))
    u = band_limited_solenoidal(n, 5, seed=1)
    uk = [o.A2a_k(to_layout(c, o.shape, n)) for c in u]
    zeta, dc = o.u_k2zeta_k(uk)
    KX, KY, _ = kvecs(n)
    w_oo = [np.conj(np.fft.rfftn(c)) for c in curl(u, n)]
    z0, z1 = zeta_components(w_oo, KX, KY)
    scale = np.abs(z0).max()
    assert np.abs(ooura_to_np(zeta[0]) - np.conj(z0)).max() < 1e-11 * scale
    assert np.abs(ooura_to_np(zeta[1]) - np.conj(z1)).max() < 1e-11 * scale
    assert np.allclose(dc, [c.sum() for c in u], atol=1e-10)
    back = o.zeta_k2u_k(zeta, dc)
    for d in range(3):
        assert np.abs(back[d] - uk[d]).max() < 1e-11 * np.abs(uk[d]).max()


def test_solenoidal_projection():
    n = 16
    o = Oracle(abi.make_params(n))
    rng = np.random.default_rng(3)
    uk = [o.A2a_k(to_layout(rng.standard_normal((n, n, n)), o.shape, n)) for _ in range(3)]
    pu = o.solenoidal_uk(uk)
    KX, KY, KZ = kvecs(n)
    div = KX * ooura_to_np(pu[0]) + KY * ooura_to_np(pu[1]) + KZ * ooura_to_np(pu[2])
    assert np.abs(div).max() < 1e-10 * np.abs(ooura_to_np(pu[0])).max()
    ppu = o.solenoidal_uk(pu)
    assert max(np.abs(ppu[d] - pu[d]).max() for d in range(3)) < 1e-11 * np.abs(pu[0]).max()
    # a gradient field is removed entirely (its mean is kept: IK2(0) = 0)
    phi = band_limited_solenoidal(n, 5, seed=4)[0]  # any smooth scalar without Nyquist content
    g = [o.A2a_k(to_layout(c, o.shape, n)) for c in spectral_grad(phi, n)]
    pg = o.solenoidal_uk(g)
    assert max(np.abs(x).max() for x in pg) < 1e-10 * max(np.abs(x).max() for x in g)


def test_advection_is_curl_of_u_cross_omega():
    n = 16
    o = Oracle(abi.make_params(n))
    u = band_limited_solenoidal(n, 3, seed=7)   # products stay below the Nyquist mode
    adv = o.u2advection_k([to_layout(c, o.shape, n) for c in u])
    w = curl(u, n)
    uxw = [u[1] * w[2] - u[2] * w[1], u[2] * w[0] - u[0] * w[2], u[0] * w[1] - u[1] * w[0]]
    omega_oo = [np.conj(np.fft.rfftn(c)) for c in curl(uxw, n)]
    KX, KY, _ = kvecs(n)
    z0, z1 = zeta_components(omega_oo, KX, KY)
    # the reference multiplies by integer wavenumbers times 2 pi / L, as here
    scale = np.abs(z0).max()
    assert scale > 1e-3
    assert np.abs(ooura_to_np(adv[0]) - np.conj(z0)).max() < 1e-10 * scale
    assert np.abs(ooura_to_np(adv[1]) - np.conj(z1)).max() < 1e-10 * scale


def test_advection_vanishes_for_abc_flow():
    n = 32
    o = Oracle(abi.make_params(n))
    x = 2 * np.pi * np.arange(n) / n
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    A, B, Cc = 1.0, 0.7, 0.4
    u = [A * np.sin(Z) + Cc * np.cos(Y), B * np.sin(X) + A * np.cos(Z), Cc * np.sin(Y) + B * np.cos(X)]
    adv = o.u2advection_k([to_layout(c, o.shape, n) for c in u])
    assert max(np.abs(a).max() for a in adv) < 1e-10 * n ** 3


def test_slaved_euler_exact_shear_decay():
    n, dt, eta = 32, 0.05, 0.7
    prm = abi.make_params(n, dt=dt, eta=eta, rho=1.3)
    o = Oracle(prm)
    y = np.arange(n)
    kmode = 3
    for d in range(3):
        o.put_field("u%d" % d, np.zeros(o.shape))
    u0 = np.zeros(o.shape)
    u0[:, :, :n] = 0.2 * np.sin(2 * np.pi * kmode * y / n)[None, :, None]
    o.put_field("u0", u0)
    o.U2zeta_k()
    z0, z1 = o.get_field("zeta0"), o.get_field("zeta1")
    o.NS_solver_slavedEuler()
    nu = eta / 1.3
    decay = np.exp(-nu * (2 * np.pi * kmode / n) ** 2 * dt)
    assert np.abs(z0).max() + np.abs(z1).max() > 1.0
    assert np.abs(o.get_field("zeta0") - decay * z0).max() < 1e-12 * np.abs(z0).max() + 1e-14
    assert np.abs(o.get_field("zeta1") - decay * z1).max() < 1e-12 * (np.abs(z1).max() + 1)


def _slaved_euler_closed_form(n, dt, eta, rho, u, seed_flip=False):
    """One slaved-Euler step of the vorticity equation, written from the
    mathematics: d(omega)/dt = curl(u x omega) + nu lap(omega), integrated
    with the linear part exact over dt and the nonlinear term N held fixed,
    omega' = e omega + (1 - e) N / (nu k^2), e = exp(-nu k^2 dt); omega and N
    reduced to the reference's two components (fft_wrapper.cxx:298-313)."""
    w = curl(u, n)
    uxw = [u[1] * w[2] - u[2] * w[1], u[2] * w[0] - u[0] * w[2], u[0] * w[1] - u[1] * w[0]]
    N = curl(uxw, n)
    KX, KY, KZ = kvecs(n)
    z_w = zeta_components([np.fft.rfftn(c) for c in w], KX, KY)
    z_n = zeta_components([np.fft.rfftn(c) for c in N], KX, KY)
    k2 = (2 * np.pi / n) ** 2 * (KX ** 2 + KY ** 2 + KZ ** 2).astype(float)
    nu = eta / rho
    e = np.exp(-nu * k2 * dt)
    k2[0, 0, 0] = 1.0
    slave = np.where(KX ** 2 + KY ** 2 + KZ ** 2 > 0, (1 - e) / (nu * k2), 0.0)
    sign = -1.0 if seed_flip else 1.0
    # in numpy's exp(-ik.x) convention, as ooura_to_np returns the oracle's
    return [e * zw + sign * slave * zn for zw, zn in zip(z_w, z_n)]


def test_slaved_euler_advection_term_closed_form():
    """NS_solver_slavedEuler (fluid_solver.cxx:328-354) with the advection
    term live: a random band-limited solenoidal flow plus a mean flow
    (uk_dc), kept whole by the 2/3 rule (|k| <= 5 < TRN = 6 at 16^3), so the
    products are alias-free on the kept modes and the numpy curl(u x omega)
    is exact there.  The step must equal the closed form on every kept mode
    and leave the rest zero; the same closed form with the sign of the
    advection term flipped must not match."""
    n, dt, eta, rho = 16, 0.05, 0.7, 1.3
    o = Oracle(abi.make_params(n, dt=dt, eta=eta, rho=rho, fix_cell=(0, 0, 0)))
    mean = np.array([0.3, -0.2, 0.15])
    u = [1.5 * c + mean[d] for d, c in enumerate(band_limited_solenoidal(n, 5, seed=11))]
    for d in range(3):
        o.put_field("u%d" % d, to_layout(u[d], o.shape, n))
    o.U2zeta_k()
    assert np.allclose(o.get_uk_dc(), [c.sum() for c in u], rtol=1e-12)
    o.NS_solver_slavedEuler()
    got = [ooura_to_np(o.get_field("zeta%d" % c)) for c in range(2)]
    want = _slaved_euler_closed_form(n, dt, eta, rho, u)
    flip = _slaved_euler_closed_form(n, dt, eta, rho, u, seed_flip=True)
    keep = ooura_to_np(o.truncate(np.ones(o.shape))) != 0
    scale = max(np.abs(w[keep]).max() for w in want)
    for c in range(2):
        assert np.abs(got[c][keep] - want[c][keep]).max() < 1e-11 * scale
        assert np.all(got[c][~keep] == 0)
    # the advection term carries weight here: flipping its sign is caught
    assert max(np.abs(got[c][keep] - flip[c][keep]).max() for c in range(2)) > 1e-2 * scale
