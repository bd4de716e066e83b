"""spm_fft3d: the context-free A2a_k / A_k2a (fft_wrapper.h:432-631) on a
caller's device field, against the oracle's transforms (themselves pinned to
the reference's Ooura build), with a caller workspace and with the
library's own per-call allocation."""
import ctypes as C

import numpy as np
import pytest

from kapsel_old_amd import abi
from kapsel_old_amd.spm import lib
from oracle.oracle import Oracle
from tests.hipmem import DeviceArray, synchronize

pytestmark = pytest.mark.gpu

SHAPES = [(32, 32, 32), (16, 64, 8), (64, 16, 32), (8, 8, 128)]


def _fft3d(shape, direction, host, workspace=True):
    nx, ny, nz = shape
    L = lib()
    d = DeviceArray.from_host(host)
    ws = None
    if workspace:
        nbytes = L.spm_fft_workspace_size(nx, ny, nz)
        ws = DeviceArray(nbytes)
        rc = L.spm_fft3d(nx, ny, nz, direction, d.ptr, ws.ptr, nbytes, None)
    else:
        rc = L.spm_fft3d(nx, ny, nz, direction, d.ptr, None, 0, None)
    assert rc == 0, L.spm_last_error(None)
    synchronize()
    out = d.download(host.shape)
    d.free()
    if ws is not None:
        ws.free()
    return out


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("workspace", [True, False])
def test_forward_matches_oracle(shape, workspace):
    nx, ny, nz = shape
    o = Oracle(abi.make_params(nx, ny, nz))
    rng = np.random.default_rng(nx * 7 + nz)
    a = np.zeros(o.shape)
    a[:, :, :nz] = rng.standard_normal((nx, ny, nz))
    want = o.A2a_k(a)
    got = _fft3d(shape, +1, a, workspace)
    assert np.abs(got - want).max() < 1e-13 * np.abs(want).max()


@pytest.mark.parametrize("shape", SHAPES)
def test_inverse_matches_oracle_on_a_non_hermitian_spectrum(shape):
    nx, ny, nz = shape
    o = Oracle(abi.make_params(nx, ny, nz))
    rng = np.random.default_rng(nx + 3 * ny)
    s = rng.standard_normal(o.shape)
    want = o.A_k2a(s)[:, :, :nz]
    got = _fft3d(shape, -1, s)[:, :, :nz]
    assert np.abs(got - want).max() < 1e-13 * np.abs(want).max()


def test_round_trip_and_rejections():
    shape = (32, 16, 64)
    rng = np.random.default_rng(4)
    a = np.zeros((32, 16, 66))
    a[:, :, :64] = rng.standard_normal((32, 16, 64))
    back = _fft3d(shape, -1, _fft3d(shape, +1, a))
    assert np.abs(back[:, :, :64] - a[:, :, :64]).max() < 1e-14 * np.abs(a).max()
    L = lib()
    d = DeviceArray(8 * 32 * 16 * 66)
    ws = DeviceArray(L.spm_fft_workspace_size(32, 16, 64))
    assert L.spm_fft3d(32, 16, 64, 2, d.ptr, ws.ptr, ws.nbytes, None) == abi.SPM_ERR_INVALID
    assert L.spm_fft3d(32, 16, 64, 1, d.ptr, ws.ptr, ws.nbytes - 1, None) == abi.SPM_ERR_INVALID
    assert L.spm_fft3d(32, 16, 64, 1, None, ws.ptr, ws.nbytes, None) == abi.SPM_ERR_INVALID
    synchronize()
