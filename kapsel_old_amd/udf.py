"""UDF input/output over libspm.so's reader (include/spm_udf.h).

`UDF` offers the calls the reference's scripts make on GOURMET's
UDFManager (get, size, totalRecord, jump; Udf/sk.py:4-70), and
`params()` / `user_particles()` turn an input file into the spm_params /
spm_electrolyte / particle arrays the solver takes, as input.cxx:438-1813
and init_particle.cxx:226-393 do.  `create_output` / `append_record` write
the output file's snapshot records as Output_udf (avs_output.cxx:343-396).
"""
import ctypes as C

import numpy as np

from . import abi
from .spm import SpmError, lib as _spm_lib

_bound = False


class SpmUdfRun(C.Structure):
    """spm_udf_run: run control read with the parameters."""
    _fields_ = [("equation", C.c_int), ("auto_dt", C.c_int), ("dt_factor", C.c_double),
                ("gts", C.c_int), ("num_snap", C.c_int), ("distribution", C.c_int),
                ("poisson_boltzmann", C.c_int), ("shear_rate", C.c_double)]


class SpmRestart(C.Structure):
    """spm_restart: resume.CONTINUE.Saved_Data with caller-owned buffers."""
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("n_particles", C.c_int), ("n_ion", C.c_int),
                ("ts", C.c_int), ("time", C.c_double), ("uk_dc", C.c_double * 3), ("degree_oblique", C.c_double),
                ("zeta0", C.POINTER(C.c_double)), ("zeta1", C.POINTER(C.c_double)),
                ("concentration", C.POINTER(C.c_double)), ("particles", C.c_void_p)]


NAVIER_STOKES, LEES_EDWARDS, ELECTROLYTE = 0, 2, 3
DISTRIBUTIONS = ["NONE", "uniform_random", "random_walk", "FCC", "BCC", "user_specify"]


def lib():
    global _bound
    L = _spm_lib()
    if not _bound:
        vp, ip, dp, cp = C.c_void_p, C.c_int, C.POINTER(C.c_double), C.c_char_p
        L.spm_udf_open.argtypes = [cp, C.POINTER(vp)]
        L.spm_udf_close.argtypes = [vp]
        L.spm_udf_close.restype = None
        L.spm_udf_last_error.argtypes = []
        L.spm_udf_last_error.restype = cp
        L.spm_udf_get_double.argtypes = [vp, cp, dp]
        L.spm_udf_get_int.argtypes = [vp, cp, C.POINTER(ip)]
        L.spm_udf_get_string.argtypes = [vp, cp, C.c_char_p, ip]
        L.spm_udf_size.argtypes = [vp, cp, C.POINTER(ip)]
        L.spm_udf_get_doubles.argtypes = [vp, cp, dp, ip, C.POINTER(ip)]
        L.spm_udf_num_records.argtypes = [vp, C.POINTER(ip)]
        L.spm_udf_jump.argtypes = [vp, ip]
        L.spm_udf_params.argtypes = [vp, C.POINTER(abi.SpmParams), C.POINTER(abi.SpmElectrolyte),
                                     C.POINTER(SpmUdfRun)]
        L.spm_udf_user_particles.argtypes = [vp, vp, ip]
        L.spm_udf_init_particles.argtypes = [vp, vp, ip]
        L.spm_udf_init_particles_rng.argtypes = [vp, vp, ip, C.POINTER(C.c_uint32), C.POINTER(ip)]
        L.spm_udf_create_output.argtypes = [cp, cp]
        L.spm_udf_append_record.argtypes = [cp, ip, vp, ip]
        L.spm_udf_write_restart.argtypes = [cp, cp, C.POINTER(SpmRestart)]
        L.spm_udf_read_restart.argtypes = [vp, C.POINTER(SpmRestart)]
        _bound = True
    return L


def _check(rc, where):
    if rc:
        raise SpmError(rc, where, lib().spm_udf_last_error().decode())


class UDF:
    """A parsed UDF file (schema, data section and records)."""

    def __init__(self, path):
        h = C.c_void_p()
        _check(lib().spm_udf_open(str(path).encode(), C.byref(h)), "spm_udf_open")
        self._h = h
        self.path = str(path)

    def close(self):
        if getattr(self, "_h", None):
            lib().spm_udf_close(self._h)
            self._h = None

    __del__ = close

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # ---- UDFManager-style access ----
    def get_double(self, loc):
        v = C.c_double()
        _check(lib().spm_udf_get_double(self._h, loc.encode(), C.byref(v)), "spm_udf_get_double(%s)" % loc)
        return v.value

    def get_int(self, loc):
        v = C.c_int()
        _check(lib().spm_udf_get_int(self._h, loc.encode(), C.byref(v)), "spm_udf_get_int(%s)" % loc)
        return v.value

    def get_string(self, loc):
        buf = C.create_string_buffer(4096)
        _check(lib().spm_udf_get_string(self._h, loc.encode(), buf, len(buf)), "spm_udf_get_string(%s)" % loc)
        return buf.value.decode()

    def get_doubles(self, loc):
        n = C.c_int()
        lib().spm_udf_get_doubles(self._h, loc.encode(), None, 0, C.byref(n))
        out = np.empty(max(n.value, 1))
        _check(lib().spm_udf_get_doubles(self._h, loc.encode(), out.ctypes.data_as(C.POINTER(C.c_double)),
                                         n.value, C.byref(n)), "spm_udf_get_doubles(%s)" % loc)
        return out[:n.value]

    def get(self, loc):
        """Number, string or (for a location with []) array of numbers."""
        if "[]" in loc and not loc.endswith("[]"):
            return self.get_doubles(loc)
        try:
            return self.get_double(loc)
        except SpmError:
            return self.get_string(loc)

    def size(self, loc):
        n = C.c_int()
        _check(lib().spm_udf_size(self._h, loc.encode(), C.byref(n)), "spm_udf_size(%s)" % loc)
        return n.value

    def totalRecord(self):
        n = C.c_int()
        _check(lib().spm_udf_num_records(self._h, C.byref(n)), "spm_udf_num_records")
        return n.value

    def jump(self, record):
        _check(lib().spm_udf_jump(self._h, record), "spm_udf_jump")

    # ---- input.cxx ----
    def params(self):
        """(spm_params, spm_electrolyte or None, spm_udf_run) for this input."""
        prm, el, run = abi.SpmParams(), abi.SpmElectrolyte(), SpmUdfRun()
        _check(lib().spm_udf_params(self._h, C.byref(prm), C.byref(el), C.byref(run)), "spm_udf_params")
        return prm, (el if run.equation == ELECTROLYTE else None), run

    def user_particles(self):
        """Particle array Init_Particle builds for INIT_distribution user_specify."""
        n = self.size("switch.INIT_distribution.user_specify.Particles[]")
        p = abi.new_particles(n)
        _check(lib().spm_udf_user_particles(self._h, p.ctypes.data_as(C.c_void_p), n), "spm_udf_user_particles")
        return p


    def init_particles(self, rng=False):
        """Particle array Init_Particle builds (user_specify, uniform_random,
        BCC, FCC, random_walk); with rng=True also the MT19937 state it
        leaves, (particles, mt, mti), for SPM.put_rng."""
        n = sum(self.get_int("object_type.spherical_particle.Particle_spec[%d].Particle_number" % s)
                for s in range(self.size("object_type.spherical_particle.Particle_spec[]")))
        p = abi.new_particles(n)
        if not rng:
            _check(lib().spm_udf_init_particles(self._h, p.ctypes.data_as(C.c_void_p), n), "spm_udf_init_particles")
            return p
        mt = (C.c_uint32 * abi.MT_N)()
        mti = C.c_int()
        _check(lib().spm_udf_init_particles_rng(self._h, p.ctypes.data_as(C.c_void_p), n, mt, C.byref(mti)),
               "spm_udf_init_particles_rng")
        return p, np.array(mt, dtype=np.uint32), mti.value


def _restart_struct(shape, zeta0, zeta1, conc, particles):
    nx, ny, nz = shape
    r = SpmRestart()
    r.nx, r.ny, r.nz = nx, ny, nz
    r.n_particles = len(particles)
    r.n_ion = 0 if conc is None else len(conc)
    dp = C.POINTER(C.c_double)
    r.zeta0 = zeta0.ctypes.data_as(dp)
    r.zeta1 = zeta1.ctypes.data_as(dp)
    r.concentration = conc.ctypes.data_as(dp) if conc is not None else None
    r.particles = particles.ctypes.data_as(C.c_void_p)
    return r


def _field(a, shape):
    nx, ny, nz = shape
    return np.ascontiguousarray(a, dtype=np.float64).reshape(nx, ny, 2 * (nz // 2 + 1))


def write_restart(in_path, out_path, shape, zeta0, zeta1, uk_dc, particles, ts, time, conc=None,
                  degree_oblique=0.0):
    """Save_Restart_udf: the input with resume = CONTINUE and this state."""
    z0, z1 = _field(zeta0, shape), _field(zeta1, shape)
    cc = None if conc is None else np.ascontiguousarray(
        [_field(c, shape) for c in conc], dtype=np.float64)
    p = np.ascontiguousarray(particles, dtype=abi.PARTICLE_DTYPE)
    r = _restart_struct(shape, z0, z1, cc, p)
    r.ts, r.time, r.degree_oblique = int(ts), float(time), float(degree_oblique)
    r.uk_dc[:] = [float(x) for x in uk_dc]
    _check(lib().spm_udf_write_restart(str(in_path).encode(), str(out_path).encode(), C.byref(r)),
           "spm_udf_write_restart")


def read_restart(u, shape, n_particles, n_ion=0, particles=None):
    """Force_restore_parameters: dict of the saved state of UDF `u`."""
    nx, ny, nz = shape
    z0 = np.zeros((nx, ny, 2 * (nz // 2 + 1)))
    z1 = np.zeros_like(z0)
    cc = np.zeros((n_ion,) + z0.shape) if n_ion else None
    p = abi.new_particles(n_particles) if particles is None else np.array(particles, dtype=abi.PARTICLE_DTYPE)
    r = _restart_struct(shape, z0, z1, cc, p)
    _check(lib().spm_udf_read_restart(u._h, C.byref(r)), "spm_udf_read_restart")
    return dict(zeta0=z0, zeta1=z1, concentration=cc, particles=p, uk_dc=np.array(r.uk_dc[:]), ts=r.ts,
                time=r.time, degree_oblique=r.degree_oblique)


def create_output(in_path, out_path):
    _check(lib().spm_udf_create_output(str(in_path).encode(), str(out_path).encode()), "spm_udf_create_output")


def append_record(out_path, ts, particles):
    p = np.ascontiguousarray(particles, dtype=abi.PARTICLE_DTYPE)
    _check(lib().spm_udf_append_record(str(out_path).encode(), int(ts), p.ctypes.data_as(C.c_void_p), len(p)),
           "spm_udf_append_record")
