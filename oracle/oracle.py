"""ctypes front end of the CPU oracle (oracle/liboracle.so) and of the
reference's own Ooura FFT / MT19937 build (oracle/_ref/libref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the product package.

`Oracle` exposes the same methods as kapsel_old_amd.spm.SPM, so a parity
test drives both through one call sequence.
"""
import ctypes as C
import os

import numpy as np

from kapsel_old_amd import abi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libref.so")

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle/liboracle.so is not built (run make -C oracle)")
        L = C.CDLL(LIB_PATH)
        vp, dp, ip = C.c_void_p, C.POINTER(C.c_double), C.c_int
        L.oracle_create.restype = vp
        L.oracle_create.argtypes = [C.POINTER(abi.SpmParams)]
        L.oracle_destroy.argtypes = [vp]
        for name in ["init_fluid", "ns_solver_slaved_euler", "zeta_k2u", "md_solver_position",
                     "calc_f_hydro", "force", "md_solver_velocity", "add_random_force_thermostat",
                     "make_phi_u_particle", "make_phi_particle", "make_f_particle_dt_sole",
                     "u2zeta_k"]:
            f = getattr(L, "oracle_" + name)
            f.argtypes = [vp]
            f.restype = ip
        L.oracle_step.argtypes = [vp, ip]
        L.oracle_set_threads.argtypes = [ip]
        L.oracle_set_threads.restype = None
        L.oracle_stencil_size.argtypes = [vp]
        L.oracle_stencil.argtypes = [vp, C.POINTER(C.c_int)]
        L.oracle_put_field.argtypes = [vp, ip, dp]
        L.oracle_get_field.argtypes = [vp, ip, dp]
        L.oracle_put_particles.argtypes = [vp, vp, ip]
        L.oracle_get_particles.argtypes = [vp, vp, ip]
        L.oracle_put_uk_dc.argtypes = [vp, dp]
        L.oracle_get_uk_dc.argtypes = [vp, dp]
        L.oracle_put_time.argtypes = [vp, ip, C.c_double]
        L.oracle_get_time.argtypes = [vp, C.POINTER(C.c_int), dp]
        L.oracle_put_rng.argtypes = [vp, C.POINTER(C.c_uint32), ip]
        L.oracle_get_rng.argtypes = [vp, C.POINTER(C.c_uint32), C.POINTER(C.c_int)]
        for name in ["A2a_k", "A_k2a", "truncate"]:
            f = getattr(L, "oracle_" + name)
            f.argtypes = [vp, dp]
            f.restype = None
        PP = dp * 3
        L.oracle_zeta_k2u_k.argtypes = [vp, dp * 2, dp, PP]
        L.oracle_zeta_k2u_k.restype = None
        L.oracle_u_k2zeta_k.argtypes = [vp, PP, dp * 2, dp]
        L.oracle_u_k2zeta_k.restype = None
        L.oracle_u2advection_k.argtypes = [vp, PP, dp * 2]
        L.oracle_u2advection_k.restype = None
        L.oracle_solenoidal_uk.argtypes = [vp, PP]
        L.oracle_solenoidal_uk.restype = None
        L.oracle_phi.argtypes = [vp, C.c_double]
        L.oracle_phi.restype = C.c_double
        L.oracle_lj_lnk.argtypes = [vp, dp, ip, dp]
        L.oracle_lj_lnk.restype = None
        L.oracle_lj_allpairs.argtypes = [vp, dp, ip, dp]
        L.oracle_lj_allpairs.restype = None
        L.oracle_genrand_int32.argtypes = [vp]
        L.oracle_genrand_int32.restype = C.c_uint32
        L.oracle_genrand_real3.argtypes = [vp]
        L.oracle_genrand_real3.restype = C.c_double
        L.oracle_gauss2.argtypes = [vp, dp]
        L.oracle_gauss2.restype = None
        L.oracle_set_electrolyte.argtypes = [vp, C.POINTER(abi.SpmElectrolyte)]
        L.oracle_init_rho_ion_pb.argtypes = [vp, C.c_double, C.POINTER(ip)]
        for name in ["init_rho_ion", "nssolute_solver_euler", "rescale_solute", "make_coulomb_force"]:
            f = getattr(L, "oracle_" + name)
            f.argtypes = [vp]
            f.restype = ip
        for name in ["put_concentration", "get_concentration"]:
            getattr(L, "oracle_" + name).argtypes = [vp, ip, dp]
        for name in ["put_total_solute", "get_total_solute"]:
            getattr(L, "oracle_" + name).argtypes = [vp, dp]
        L.oracle_make_phi_qq.argtypes = [vp, dp, dp]
        L.oracle_make_phi_qq.restype = None
        L.oracle_make_surface_normal.argtypes = [vp, PP]
        L.oracle_make_surface_normal.restype = None
        L.oracle_set_lees_edwards.argtypes = [vp, C.c_double]
        L.oracle_put_oblique.argtypes = [vp, C.c_double, C.c_double]
        L.oracle_get_oblique.argtypes = [vp, dp, dp]
        L.oracle_distance0_obl.argtypes = [vp, dp, dp, dp]
        L.oracle_distance0_obl.restype = C.c_double
        L.oracle_le_shear_stress.argtypes = [vp, dp]
        for name in ["ns_solver_shear_obl", "zeta_k2u_k_obl", "u_k2zeta_k_obl", "make_phi_u_particle_obl",
                     "calc_f_hydro_obl", "md_solver_position_obl", "md_solver_velocity_obl", "force_obl",
                     "u2u_oblique", "u_oblique2u", "remap_obl"]:
            f = getattr(L, "oracle_" + name)
            f.argtypes = [vp]
            f.restype = ip
        _lib = L
    return _lib


def set_threads(n):
    """OpenMP threads of the oracle's parallel loops (line transforms,
    per-particle force sums); results do not depend on it."""
    lib().oracle_set_threads(int(n))


def _dp(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_double))


class Oracle:
    """CPU FP64 restatement of the SPM step behind the SPM call surface."""

    def __init__(self, params):
        self.params = params
        self._h = lib().oracle_create(C.byref(params))
        if not self._h:
            raise ValueError("oracle_create rejected the parameters")
        p = params
        self.shape = (p.nx, p.ny, 2 * (p.nz // 2 + 1))
        self.nfield = p.nx * p.ny * self.shape[2]
        self.n_particles = p.n_particles

    def close(self):
        if self._h:
            lib().oracle_destroy(self._h)
            self._h = None

    __del__ = close

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _call(self, name, *args):
        rc = getattr(lib(), "oracle_" + name)(self._h, *args)
        if rc:
            raise RuntimeError("oracle_%s failed (%d)" % (name, rc))

    # ---- the solver call surface (names follow the reference) ----
    def init_fluid(self): self._call("init_fluid")
    def NS_solver_slavedEuler(self): self._call("ns_solver_slaved_euler")
    def Zeta_k2u(self): self._call("zeta_k2u")
    def MD_solver_position(self): self._call("md_solver_position")
    def Calc_f_hydro_correct_precision(self): self._call("calc_f_hydro")
    def Force(self): self._call("force")
    def MD_solver_velocity(self): self._call("md_solver_velocity")
    def Add_random_force_thermostat(self): self._call("add_random_force_thermostat")
    def Make_phi_u_particle(self): self._call("make_phi_u_particle")
    def Make_phi_particle(self): self._call("make_phi_particle")
    def Make_f_particle_dt_sole(self): self._call("make_f_particle_dt_sole")
    def U2zeta_k(self): self._call("u2zeta_k")

    # electrolyte (SW_EQ == Electrolyte)
    def set_electrolyte(self, el):
        self.n_spec = el.n_spec
        self._call("set_electrolyte", C.byref(el))

    def Init_rho_ion(self): self._call("init_rho_ion")

    def Init_rho_ion_PB(self, tol=1e-4):
        """Init_rho_ion with Poisson_Boltzmann on (salt, several particles),
        relaxed until the free energy changes by less than tol (the
        reference's 1e-4); returns the iteration count."""
        it = C.c_int()
        self._call("init_rho_ion_pb", C.c_double(tol), C.byref(it))
        return it.value
    def NSsolute_solver_Euler(self): self._call("nssolute_solver_euler")
    def Rescale_solute(self): self._call("rescale_solute")
    def Make_Coulomb_force_x_on_fluid(self): self._call("make_coulomb_force")

    def get_concentration(self, n):
        a = np.zeros(self.shape)
        self._call("get_concentration", n, _dp(a))
        return a

    def put_concentration(self, n, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        self._call("put_concentration", n, _dp(a))

    def get_total_solute(self):
        t = np.zeros(2)
        self._call("get_total_solute", _dp(t))
        return t[:self.n_spec]

    def put_total_solute(self, t):
        tt = np.zeros(2)
        tt[:len(t)] = t
        self._call("put_total_solute", _dp(tt))

    def make_phi_qq(self):
        """(phi, surface charge density) of Make_phi_qq_particle."""
        phi, surf = np.zeros(self.shape), np.zeros(self.shape)
        lib().oracle_make_phi_qq(self._h, _dp(phi), _dp(surf))
        return phi, surf

    def make_surface_normal(self):
        n, ptrs = self._vec([np.zeros(self.shape) for _ in range(3)])
        lib().oracle_make_surface_normal(self._h, ptrs)
        return n

    # Lees-Edwards shear, fluid only (SW_EQ == Shear_Navier_Stokes_Lees_Edwards)
    def set_lees_edwards(self, shear_rate):
        self._call("set_lees_edwards", C.c_double(shear_rate))

    def put_oblique(self, degree_oblique, shear_rate_eff):
        self._call("put_oblique", C.c_double(degree_oblique), C.c_double(shear_rate_eff))

    def get_oblique(self):
        d, s = C.c_double(), C.c_double()
        self._call("get_oblique", C.byref(d), C.byref(s))
        return d.value, s.value

    def NS_solver_slavedEuler_Shear_OBL(self): self._call("ns_solver_shear_obl")
    def Zeta_k2u_k_OBL(self): self._call("zeta_k2u_k_obl")
    def U_k2zeta_k_OBL(self):
        """Solenoidal_uk_OBL, contra2co and U_k2zeta_k_OBL of the spectrum in u."""
        self._call("u_k2zeta_k_obl")

    # spherical particles under shear; the Cartesian velocity is u0..u2
    def Make_phi_u_particle_OBL(self): self._call("make_phi_u_particle_obl")
    def Calc_f_hydro_correct_precision_OBL(self): self._call("calc_f_hydro_obl")
    def MD_solver_position_OBL(self): self._call("md_solver_position_obl")
    def MD_solver_velocity_OBL(self): self._call("md_solver_velocity_obl")
    def Force_OBL(self): self._call("force_obl")
    def U2u_oblique(self): self._call("u2u_oblique")
    def U_oblique2u(self): self._call("u_oblique2u")
    def Remap_OBL(self): self._call("remap_obl")

    def le_shear_stress(self):
        """Mean_shear_stress's Lees-Edwards line as a dict."""
        out = np.zeros(7)
        self._call("le_shear_stress", _dp(out))
        return dict(zip(abi.LE_STRESS_KEYS, out.tolist()))

    def Distance0_OBL(self, x1, x2):
        a, b = np.array(x1, dtype=np.float64), np.array(x2, dtype=np.float64)
        x12 = np.zeros(3)
        r = lib().oracle_distance0_obl(self._h, _dp(a), _dp(b), _dp(x12))
        return r, x12

    def step(self, nsteps=1):
        self._call("step", nsteps)

    # ---- state ----
    def get_field(self, name):
        out = np.empty(self.shape)
        self._call("get_field", abi.FIELD_IDS[name], _dp(out))
        return out

    def put_field(self, name, arr):
        arr = np.ascontiguousarray(arr, dtype=np.float64).reshape(self.shape)
        self._call("put_field", abi.FIELD_IDS[name], _dp(arr))

    def get_particles(self):
        p = abi.new_particles(self.n_particles)
        self._call("get_particles", p.ctypes.data_as(C.c_void_p), self.n_particles)
        return p

    def put_particles(self, p):
        p = np.ascontiguousarray(p, dtype=abi.PARTICLE_DTYPE)
        self._call("put_particles", p.ctypes.data_as(C.c_void_p), len(p))

    def get_uk_dc(self):
        out = np.empty(3)
        self._call("get_uk_dc", _dp(out))
        return out

    def put_uk_dc(self, v):
        v = np.ascontiguousarray(v, dtype=np.float64)
        self._call("put_uk_dc", _dp(v))

    def get_time(self):
        ts, t = C.c_int(), C.c_double()
        self._call("get_time", C.byref(ts), C.byref(t))
        return ts.value, t.value

    def put_time(self, ts, t):
        self._call("put_time", ts, C.c_double(t))

    def get_rng(self):
        mt = (C.c_uint32 * abi.MT_N)()
        mti = C.c_int()
        self._call("get_rng", mt, C.byref(mti))
        return np.array(mt, dtype=np.uint32), mti.value

    def put_rng(self, mt, mti):
        arr = (C.c_uint32 * abi.MT_N)(*[int(x) for x in mt])
        self._call("put_rng", arr, mti)

    def stencil(self):
        n = lib().oracle_stencil_size(self._h)
        out = np.empty((n, 3), dtype=np.int32)
        lib().oracle_stencil(self._h, out.ctypes.data_as(C.POINTER(C.c_int)))
        return out

    # ---- building blocks on numpy arrays ----
    def A2a_k(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64).reshape(self.shape).copy()
        lib().oracle_A2a_k(self._h, _dp(a))
        return a

    def A_k2a(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64).reshape(self.shape).copy()
        lib().oracle_A_k2a(self._h, _dp(a))
        return a

    def truncate(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64).reshape(self.shape).copy()
        lib().oracle_truncate(self._h, _dp(a))
        return a

    def _vec(self, arrs):
        arrs = [np.ascontiguousarray(x, dtype=np.float64).reshape(self.shape).copy() for x in arrs]
        ptrs = (C.POINTER(C.c_double) * len(arrs))(*[_dp(x) for x in arrs])
        return arrs, ptrs

    def zeta_k2u_k(self, zeta, uk_dc):
        z, zp = self._vec(zeta)
        u, up = self._vec([np.zeros(self.shape)] * 3)
        dc = np.ascontiguousarray(uk_dc, dtype=np.float64)
        lib().oracle_zeta_k2u_k(self._h, zp, _dp(dc), up)
        return u

    def u_k2zeta_k(self, uk):
        u, up = self._vec(uk)
        z, zp = self._vec([np.zeros(self.shape)] * 2)
        dc = np.zeros(3)
        lib().oracle_u_k2zeta_k(self._h, up, zp, _dp(dc))
        return z, dc

    def u2advection_k(self, u):
        u, up = self._vec(u)
        a, ap = self._vec([np.zeros(self.shape)] * 2)
        lib().oracle_u2advection_k(self._h, up, ap)
        return a

    def solenoidal_uk(self, uk):
        u, up = self._vec(uk)
        lib().oracle_solenoidal_uk(self._h, up)
        return u

    def phi_profile(self, x):
        return lib().oracle_phi(self._h, float(x))

    def lj(self, x, lnk=True):
        x = np.ascontiguousarray(x, dtype=np.float64)
        fr = np.zeros_like(x)
        f = lib().oracle_lj_lnk if lnk else lib().oracle_lj_allpairs
        f(self._h, _dp(x), len(x), _dp(fr))
        return fr

    def genrand_int32(self):
        return lib().oracle_genrand_int32(self._h)

    def genrand_real3(self):
        return lib().oracle_genrand_real3(self._h)

    def gauss2(self):
        out = np.empty(2)
        lib().oracle_gauss2(self._h, _dp(out))
        return out


# ---------------------------------------------------------------------
# The reference's own Ooura FFT and MT19937, compiled from its sources.
# ---------------------------------------------------------------------
_ref = None


def ref_available():
    """Whether oracle/_ref/libref.so was built (checks the file only, so a
    skip condition evaluated at collection loads nothing)."""
    return os.path.exists(REF_PATH)


def ref_lib():
    """Handle on oracle/_ref/libref.so, or None where it was not built."""
    global _ref
    if _ref is None and os.path.exists(REF_PATH):
        L = C.CDLL(REF_PATH)
        ppp = C.POINTER(C.POINTER(C.POINTER(C.c_double)))
        L.rdft3d.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, ppp,
                             C.POINTER(C.c_double), C.POINTER(C.c_int), C.POINTER(C.c_double)]
        L.rdft3d.restype = None
        L.rdft3dsort.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, ppp]
        L.rdft3dsort.restype = None
        L.init_genrand.argtypes = [C.c_ulong]
        L.init_genrand.restype = None
        L.genrand_int32.restype = C.c_ulong
        L.genrand_real3.restype = C.c_double
        dp, vp = C.POINTER(C.c_double), C.c_void_p
        for name in ["ref_orientation_euler", "ref_orientation_ab2"]:
            getattr(L, name).argtypes = [vp, C.c_double]
            getattr(L, name).restype = None
        L.ref_rigid_body_rotation.argtypes = [dp, dp, dp, C.c_int]
        L.ref_rigid_body_rotation.restype = None
        L.ref_qdot.argtypes = [dp, dp, dp, C.c_int]
        L.ref_qdot.restype = None
        L.ref_qtn_normalize.argtypes = [dp]
        L.ref_qtn_normalize.restype = None
        L.ref_particle_layout.argtypes = [C.POINTER(C.c_size_t), C.c_int]
        L.ref_particle_layout.restype = C.c_int
        _ref = L
    return _ref


def _ooura_pointers(buf):
    """double*** view of a contiguous [n1][n2][n3p] array for rdft3d."""
    n1, n2, _ = buf.shape
    rows = [(C.POINTER(C.c_double) * n2)() for _ in range(n1)]
    base = buf.ctypes.data
    stride1, stride2 = buf.strides[0], buf.strides[1]
    for i in range(n1):
        for j in range(n2):
            rows[i][j] = C.cast(base + i * stride1 + j * stride2, C.POINTER(C.c_double))
    top = (C.POINTER(C.POINTER(C.c_double)) * n1)(
        *[C.cast(r, C.POINTER(C.POINTER(C.c_double))) for r in rows])
    return C.cast(top, C.POINTER(C.POINTER(C.POINTER(C.c_double)))), rows, top


def _ooura_work(nx, ny, nz):
    nt = max(nx, ny)
    n = max(nt, nz // 2)
    ip = (C.c_int * (2 + int((n + 0.5) ** 0.5) + 2))()
    t = (C.c_double * (8 * nt))()
    w = (C.c_double * (n // 2 + nz // 4 + 2))()
    ip[0] = 0
    return t, ip, w


def ref_A2a_k(a_real):
    """fft_wrapper.h:432-461 with the reference's own rdft3d/rdft3dsort:
    real field [NX][NY][NZ_] (k < NZ used) -> half spectrum."""
    L = ref_lib()
    nx, ny, nzp = a_real.shape
    nz = nzp - 2
    buf = np.zeros((nx, ny, nzp))
    buf[:, :, :nz] = a_real[:, :, :nz]
    ppp, _keep1, _keep2 = _ooura_pointers(buf)
    t, ip, w = _ooura_work(nx, ny, nz)
    L.rdft3d(nx, ny, nz, 1, ppp, t, ip, w)
    L.rdft3dsort(nx, ny, nz, 1, ppp)
    return buf


def ref_A_k2a(a_k):
    """fft_wrapper.h:534-563 with rdft3dsort/rdft3d and the 2/N scale."""
    L = ref_lib()
    nx, ny, nzp = a_k.shape
    nz = nzp - 2
    buf = np.ascontiguousarray(a_k, dtype=np.float64).copy()
    ppp, _keep1, _keep2 = _ooura_pointers(buf)
    t, ip, w = _ooura_work(nx, ny, nz)
    L.rdft3dsort(nx, ny, nz, -1, ppp)
    L.rdft3d(nx, ny, nz, -1, ppp, t, ip, w)
    return buf * (2.0 / (nx * ny * nz))
