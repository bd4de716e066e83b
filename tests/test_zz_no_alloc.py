"""spm_fft3d_noalloc (root tables in the library's static device scratch)
against spm_fft3d with a caller workspace, through the C ABI: bit for bit,
on two calls in a row per direction, the second at a smaller size than the
first, so stale tables from the larger call would show."""
import numpy as np
import pytest

from kapsel_old_amd import abi
from kapsel_old_amd.spm import lib
from tests.hipmem import DeviceArray, synchronize

pytestmark = pytest.mark.gpu


def _bind():
    return lib()


def _host(shape, seed):
    nx, ny, nz = shape
    rng = np.random.default_rng(seed)
    a = rng.standard_normal((nx, ny, 2 * (nz // 2 + 1)))
    return a


@pytest.mark.parametrize("direction", [+1, -1])
def test_noalloc_equals_workspace_entry_point(direction):
    L = _bind()
    for call, shape in enumerate([(64, 32, 128), (16, 8, 32)]):
        nx, ny, nz = shape
        a = _host(shape, 10 * call + direction + 2)
        if direction > 0:
            a[:, :, nz:] = 0.0
        d_ws, d_na = DeviceArray.from_host(a), DeviceArray.from_host(a)
        ws = DeviceArray(L.spm_fft_workspace_size(nx, ny, nz))
        assert L.spm_fft3d(nx, ny, nz, direction, d_ws.ptr, ws.ptr, ws.nbytes, None) == 0, L.spm_last_error(None)
        assert L.spm_fft3d_noalloc(nx, ny, nz, direction, d_na.ptr, None) == 0, L.spm_last_error(None)
        synchronize()
        want, got = d_ws.download(a.shape), d_na.download(a.shape)
        assert np.abs(want).max() > 0.0
        assert np.array_equal(got, want), (call, shape)
        for b in (d_ws, d_na, ws):
            b.free()


def test_noalloc_rejects_sizes_beyond_its_scratch():
    L = _bind()
    assert L.spm_fft3d_noalloc(2048, 64, 64, 1, None, None) == abi.SPM_ERR_INVALID
    assert L.spm_fft3d_noalloc(100, 64, 64, 1, None, None) == abi.SPM_ERR_INVALID
