"""Shared builders for parity tests: seeded particles and fields."""
import numpy as np

from kapsel_old_amd import abi


def particles(n, box, a, seed=1, vel=0.0, omega=0.0, spec_count=1, min_factor=2.1, positions=None):
    """n particles in the reference layout, non-overlapping at min_factor*a,
    with seeded random velocities/angular velocities of the given scale."""
    p = abi.new_particles(n)
    if positions is None:
        positions = abi.random_positions(n, box, min_factor * a, seed=seed) if n else np.zeros((0, 3))
    rng = np.random.default_rng(seed + 1000)
    p["x"] = positions
    p["x_nopbc"] = positions
    p["v"] = vel * rng.standard_normal((n, 3))
    p["omega"] = omega * rng.standard_normal((n, 3))
    p["spec"] = np.arange(n) % spec_count
    return p


def real_field(shape, nz, seed, scale=1.0):
    """Random real-space field in the [NX][NY][NZ_] layout (padding zero)."""
    rng = np.random.default_rng(seed)
    a = scale * rng.standard_normal(shape)
    a[:, :, nz:] = 0.0
    return a


def smooth_field(shape, nz, seed, kmax=3, scale=1.0):
    """Random real field made of a few low Fourier modes (well resolved)."""
    nx, ny, _ = shape
    rng = np.random.default_rng(seed)
    x = np.arange(nx)[:, None, None] / nx
    y = np.arange(ny)[None, :, None] / ny
    z = np.arange(nz)[None, None, :] / nz
    f = np.zeros((nx, ny, nz))
    for _ in range(6):
        kx, ky, kz = rng.integers(-kmax, kmax + 1, 3)
        ph = rng.uniform(0, 2 * np.pi)
        f += rng.standard_normal() * np.cos(2 * np.pi * (kx * x + ky * y + kz * z) + ph)
    out = np.zeros(shape)
    out[:, :, :nz] = scale * f
    return out


def rel(a, b):
    """max |a - b| / max |b| (absolute if b is zero)."""
    a = np.asarray(a, dtype=float)
    b = np.asarray(b, dtype=float)
    den = np.abs(b).max() if b.size else 0.0
    d = np.abs(a - b).max() if a.size else 0.0
    return d / den if den > 0 else d


def real_part(a, nz):
    return a[:, :, :nz]
