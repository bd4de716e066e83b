"""MI355X-native SPM time step (KAPSEL's Time_evolution_hydro) behind a C ABI.

The compute lives in libspm.so (hand-written HIP for gfx950, built by
kapsel_old_amd.build); `SPM` is its Python front end.
"""
from . import abi  # noqa: F401


def __getattr__(name):
    if name in ("SPM", "SpmError"):
        from . import spm
        return getattr(spm, name)
    raise AttributeError(name)
