"""Python front end of libspm.so (the C ABI declared in include/spm.h).

`SPM` mirrors the reference's solver call surface by name --
NS_solver_slavedEuler, Make_phi_u_particle, Calc_f_hydro_correct_precision,
Force, ... -- over one GPU context.  Everything runs in the HIP library;
this module only moves arguments.  It raises if libspm.so is missing or
fails: there is no CPU fallback.
"""
import ctypes as C
import os

import numpy as np

from . import abi

HERE = os.path.dirname(os.path.abspath(__file__))
# SPM_LIBRARY names another build of the library (python -m
# kapsel_old_amd.build --out PATH); the default is the in-tree one.
LIB_PATH = os.environ.get("SPM_LIBRARY") or os.path.join(HERE, "libspm.so")

SYMBOLS = [
    "spm_create", "spm_destroy", "spm_strerror", "spm_last_error", "spm_stencil_size",
    "spm_put_particles", "spm_get_particles", "spm_put_field", "spm_get_field",
    "spm_put_uk_dc", "spm_get_uk_dc", "spm_put_time", "spm_get_time",
    "spm_put_rng", "spm_get_rng",
    "spm_init_fluid", "spm_ns_solver_slaved_euler", "spm_zeta_k2u", "spm_md_solver_position",
    "spm_calc_f_hydro", "spm_force", "spm_md_solver_velocity", "spm_add_random_force_thermostat",
    "spm_make_phi_u_particle", "spm_make_f_particle_dt_sole", "spm_u2zeta_k", "spm_step",
    "spm_synchronize", "spm_check_finite", "spm_set_schedule", "spm_timing_enable", "spm_timing_read", "spm_fused_stage",
    "spm_set_electrolyte", "spm_init_rho_ion", "spm_init_rho_ion_pb", "spm_put_concentration", "spm_get_concentration",
    "spm_put_total_solute", "spm_get_total_solute", "spm_nssolute_solver_euler", "spm_rescale_solute",
    "spm_make_coulomb_force", "spm_save_state", "spm_load_state",
    "spm_set_lees_edwards", "spm_put_oblique", "spm_get_oblique", "spm_ns_solver_slaved_euler_shear_obl",
    "spm_zeta_k2u_k_obl", "spm_u_k2zeta_k_obl", "spm_make_phi_u_particle_obl", "spm_calc_f_hydro_obl",
    "spm_md_solver_position_obl", "spm_md_solver_velocity_obl",
    "spm_fft_workspace_size", "spm_fft3d", "spm_fft3d_noalloc",
    "spm_closure_create", "spm_closure_destroy", "spm_closure_init", "spm_closure_get_params",
    "spm_closure_put_params", "spm_closure_set_scaling", "spm_closure_get_scaling", "spm_closure_forward",
    "spm_closure_gradient", "spm_closure_set_data", "spm_closure_train", "spm_set_closure", "spm_closure_features",
    "spm_set_particle_motion", "spm_le_shear_stress",
    "spm_ensemble_create", "spm_ensemble_destroy", "spm_ensemble_box", "spm_ensemble_size", "spm_ensemble_step",
    "spm_ensemble_set_graph", "spm_ensemble_autotune", "spm_ensemble_tuning", "spm_autotune_report",
    "spm_ensemble_last_error",
]

SCHEDULE_FUSED, SCHEDULE_LITERAL = 0, 1

_lib = None


class SpmError(RuntimeError):
    def __init__(self, status, where, detail=""):
        self.status = status
        super().__init__("%s: %s (%d)%s" % (where, _strerror(status), status,
                                             (": " + detail) if detail else ""))


def _strerror(status):
    return lib().spm_strerror(status).decode()


def lib():
    """Load libspm.so (raises if it is not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("%s is not built; run python -m kapsel_old_amd.build" % LIB_PATH)
        L = C.CDLL(LIB_PATH)
        vp, ip, dp = C.c_void_p, C.c_int, C.POINTER(C.c_double)
        L.spm_create.argtypes = [C.POINTER(abi.SpmParams), C.POINTER(vp)]
        L.spm_destroy.argtypes = [vp]
        L.spm_strerror.argtypes = [ip]
        L.spm_strerror.restype = C.c_char_p
        L.spm_last_error.argtypes = [vp]
        L.spm_last_error.restype = C.c_char_p
        L.spm_stencil_size.argtypes = [vp]
        L.spm_put_particles.argtypes = [vp, vp, ip]
        L.spm_get_particles.argtypes = [vp, vp, ip]
        L.spm_put_field.argtypes = [vp, ip, dp]
        L.spm_get_field.argtypes = [vp, ip, dp]
        L.spm_put_uk_dc.argtypes = [vp, dp]
        L.spm_get_uk_dc.argtypes = [vp, dp]
        L.spm_put_time.argtypes = [vp, ip, C.c_double]
        L.spm_get_time.argtypes = [vp, C.POINTER(ip), dp]
        L.spm_put_rng.argtypes = [vp, C.POINTER(C.c_uint32), ip]
        L.spm_get_rng.argtypes = [vp, C.POINTER(C.c_uint32), C.POINTER(ip)]
        for name in ["spm_init_fluid", "spm_ns_solver_slaved_euler", "spm_zeta_k2u",
                     "spm_md_solver_position", "spm_calc_f_hydro", "spm_force",
                     "spm_md_solver_velocity", "spm_add_random_force_thermostat",
                     "spm_make_phi_u_particle", "spm_make_f_particle_dt_sole",
                     "spm_u2zeta_k", "spm_synchronize", "spm_check_finite"]:
            getattr(L, name).argtypes = [vp]
        L.spm_step.argtypes = [vp, ip]
        L.spm_set_schedule.argtypes = [vp, ip]
        L.spm_fused_stage.argtypes = [vp, ip]
        L.spm_timing_enable.argtypes = [vp, ip]
        L.spm_timing_read.argtypes = [vp, dp, dp, C.POINTER(C.c_long)]
        L.spm_set_electrolyte.argtypes = [vp, C.POINTER(abi.SpmElectrolyte)]
        L.spm_init_rho_ion_pb.argtypes = [vp, C.POINTER(ip)]
        for name in ["spm_init_rho_ion", "spm_nssolute_solver_euler", "spm_rescale_solute",
                     "spm_make_coulomb_force"]:
            getattr(L, name).argtypes = [vp]
        L.spm_put_concentration.argtypes = [vp, ip, dp]
        L.spm_get_concentration.argtypes = [vp, ip, dp]
        L.spm_save_state.argtypes = [vp, C.c_char_p]
        L.spm_load_state.argtypes = [vp, C.c_char_p]
        L.spm_set_lees_edwards.argtypes = [vp, C.c_double]
        L.spm_put_oblique.argtypes = [vp, C.c_double, C.c_double]
        L.spm_get_oblique.argtypes = [vp, dp, dp]
        for name in ["spm_ns_solver_slaved_euler_shear_obl", "spm_zeta_k2u_k_obl", "spm_u_k2zeta_k_obl",
                     "spm_make_phi_u_particle_obl", "spm_calc_f_hydro_obl", "spm_md_solver_position_obl",
                     "spm_md_solver_velocity_obl"]:
            getattr(L, name).argtypes = [vp]
        L.spm_fft_workspace_size.argtypes = [ip, ip, ip]
        L.spm_fft_workspace_size.restype = C.c_size_t
        L.spm_fft3d.argtypes = [ip, ip, ip, ip, vp, vp, C.c_size_t, vp]
        L.spm_fft3d_noalloc.argtypes = [ip, ip, ip, ip, vp, vp]
        fp = C.POINTER(C.c_float)
        L.spm_closure_create.argtypes = [ip, ip, C.POINTER(vp)]
        L.spm_closure_destroy.argtypes = [vp]
        L.spm_closure_init.argtypes = [vp, C.c_uint64]
        L.spm_closure_get_params.argtypes = [vp, fp]
        L.spm_closure_put_params.argtypes = [vp, fp]
        L.spm_closure_set_scaling.argtypes = [vp, fp, fp, fp]
        L.spm_closure_get_scaling.argtypes = [vp, fp, fp, fp]
        L.spm_closure_forward.argtypes = [vp, fp, fp, ip]
        L.spm_closure_gradient.argtypes = [vp, fp, fp, ip, fp, fp]
        L.spm_closure_set_data.argtypes = [vp, fp, fp, ip]
        L.spm_closure_train.argtypes = [vp, ip, ip, C.c_float, fp]
        L.spm_set_closure.argtypes = [vp, vp]
        L.spm_closure_features.argtypes = [vp, fp]
        L.spm_set_particle_motion.argtypes = [vp, ip]
        L.spm_le_shear_stress.argtypes = [vp, dp]
        L.spm_ensemble_create.argtypes = [ip, C.POINTER(abi.SpmParams), C.POINTER(vp)]
        L.spm_ensemble_destroy.argtypes = [vp]
        L.spm_ensemble_box.argtypes = [vp, ip]
        L.spm_ensemble_box.restype = vp
        L.spm_ensemble_size.argtypes = [vp]
        L.spm_ensemble_step.argtypes = [vp, ip]
        L.spm_ensemble_set_graph.argtypes = [vp, ip]
        L.spm_ensemble_autotune.argtypes = [vp, ip]
        L.spm_ensemble_tuning.argtypes = [vp, C.POINTER(ip), dp, dp]
        L.spm_autotune_report.argtypes = [C.c_char_p, ip]
        L.spm_ensemble_last_error.argtypes = [vp]
        L.spm_ensemble_last_error.restype = C.c_char_p
        L.spm_put_total_solute.argtypes = [vp, dp]
        L.spm_get_total_solute.argtypes = [vp, dp]
        for name in SYMBOLS:
            f = getattr(L, name)
            if f.restype is C.c_int or f.restype is None:
                pass
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class SPM:
    """One SPM box on one GPU (an spm_ctx)."""

    def __init__(self, params):
        self.params = params
        h = C.c_void_p()
        rc = lib().spm_create(C.byref(params), C.byref(h))
        if rc:
            raise SpmError(rc, "spm_create", lib().spm_last_error(None).decode())
        self._h = h
        p = params
        self.shape = (p.nx, p.ny, 2 * (p.nz // 2 + 1))
        self.n_particles = p.n_particles

    @classmethod
    def _wrap(cls, handle, params, owner):
        """A view of a context another object owns (an ensemble's box)."""
        self = cls.__new__(cls)
        self.params = params
        self._h = C.c_void_p(handle)
        self._owner = owner
        p = params
        self.shape = (p.nx, p.ny, 2 * (p.nz // 2 + 1))
        self.n_particles = p.n_particles
        return self

    def close(self):
        if getattr(self, "_owner", None) is not None:
            self._h = None  # the owner destroys it
            return
        if getattr(self, "_h", None):
            lib().spm_destroy(self._h)
            self._h = None

    __del__ = close

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _call(self, name, *args):
        rc = getattr(lib(), "spm_" + name)(self._h, *args)
        if rc:
            raise SpmError(rc, "spm_" + name, lib().spm_last_error(self._h).decode())

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
    def Make_f_particle_dt_sole(self): self._call("make_f_particle_dt_sole")
    def U2zeta_k(self): self._call("u2zeta_k")

    def step(self, nsteps=1):
        self._call("step", nsteps)

    def synchronize(self):
        self._call("synchronize")

    def set_schedule(self, schedule):
        """SCHEDULE_FUSED (default) or SCHEDULE_LITERAL for spm_step."""
        self._call("set_schedule", schedule)

    # ---- learned closure and particle motion (include/spm_closure.h) ----
    def set_closure(self, closure):
        """Attach a Closure (None detaches): each step adds its force and
        torque correction after Calc_f_hydro_correct_precision."""
        self._closure = closure
        self._call("set_closure", closure._h if closure is not None else None)

    def closure_features(self):
        """The closure's inputs of every particle, float32 [n][Closure.IN]."""
        x = np.zeros((self.n_particles, Closure.IN), dtype=np.float32)
        self._call("closure_features", x.ctypes.data_as(C.POINTER(C.c_float)))
        return x

    def set_particle_motion(self, prescribed):
        """False: free particles (default); True: positions, velocities and
        angular velocities held fixed while the step forms the force."""
        self._call("set_particle_motion", 1 if prescribed else 0)

    # ---- electrolyte (SW_EQ == Electrolyte) ----
    def set_electrolyte(self, el):
        self.n_spec = el.n_spec
        self._call("set_electrolyte", C.byref(el))

    def Init_rho_ion(self): self._call("init_rho_ion")

    def Init_rho_ion_PB(self):
        """Init_rho_ion with Poisson_Boltzmann on (salt, several particles):
        the uniform start relaxed to the steady state; returns the
        relaxation's iteration count."""
        it = C.c_int()
        self._call("init_rho_ion_pb", C.byref(it))
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

    # ---- Lees-Edwards shear (SW_EQ == Shear_Navier_Stokes_Lees_Edwards) ----
    def le_shear_stress(self):
        """Mean_shear_stress's Lees-Edwards line of the last step (dict:
        time, shear_rate, strain, lj_dev_stress, hydro_stress,
        hydro_stress_new, viscosity)."""
        out = np.zeros(7)
        self._call("le_shear_stress", _dp(out))
        return dict(zip(abi.LE_STRESS_KEYS, out.tolist()))

    def set_lees_edwards(self, shear_rate):
        self._call("set_lees_edwards", C.c_double(shear_rate))

    def put_oblique(self, degree_oblique, shear_rate_eff):
        self._call("put_oblique", C.c_double(degree_oblique), C.c_double(shear_rate_eff))

    def get_oblique(self):
        """(degree_oblique, Shear_rate_eff)."""
        d, s = C.c_double(), C.c_double()
        self._call("get_oblique", C.byref(d), C.byref(s))
        return d.value, s.value

    def NS_solver_slavedEuler_Shear_OBL(self): self._call("ns_solver_slaved_euler_shear_obl")
    def Make_phi_u_particle_OBL(self): self._call("make_phi_u_particle_obl")
    def Calc_f_hydro_correct_precision_OBL(self): self._call("calc_f_hydro_obl")
    def MD_solver_position_OBL(self): self._call("md_solver_position_obl")
    def MD_solver_velocity_OBL(self): self._call("md_solver_velocity_obl")
    def Zeta_k2u_k_OBL(self): self._call("zeta_k2u_k_obl")

    def U_k2zeta_k_OBL(self):
        """Solenoidal_uk_OBL, contra2co and U_k2zeta_k_OBL of the spectrum in u."""
        self._call("u_k2zeta_k_obl")

    def fused_stage(self, stage):
        """Test hook: one stage of the fused schedule (see include/spm.h)."""
        self._call("fused_stage", stage)

    def timing_enable(self, on=True):
        self._call("timing_enable", 1 if on else 0)

    def timing_read(self):
        """Per kernel family: (ms, algorithmic bytes, launches) since the last read."""
        ms = (C.c_double * 9)()
        by = (C.c_double * 9)()
        cnt = (C.c_long * 9)()
        self._call("timing_read", ms, by, cnt)
        return list(ms), list(by), list(cnt)

    def check_finite(self):
        self._call("check_finite")

    # ---- state ----
    def stencil_size(self):
        return lib().spm_stencil_size(self._h)

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

    def save_state(self, path):
        """Exact checkpoint (spm_save_state), RNG state included."""
        self._call("save_state", str(path).encode())

    def load_state(self, path):
        self._call("load_state", str(path).encode())

    def put_rng(self, mt, mti):
        arr = (C.c_uint32 * abi.MT_N)(*[int(x) for x in mt])
        self._call("put_rng", arr, mti)


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


class Closure:
    """The learned force/torque closure's network on one GPU (an
    spm_closure): 126 inputs -> 128 -> 128 -> 128 -> 6, BF16 weights on the
    matrix cores with FP32 accumulation, trained with Adam on the GPU."""

    NPARAM = 51600
    IN, OUT = 126, 6

    def __init__(self, max_batch, device=0, seed=0):
        h = C.c_void_p()
        rc = lib().spm_closure_create(device, max_batch, C.byref(h))
        if rc:
            raise SpmError(rc, "spm_closure_create", lib().spm_last_error(None).decode())
        self._h = h
        self.max_batch = max_batch
        if seed:
            self.init(seed)

    def close(self):
        if getattr(self, "_h", None):
            lib().spm_closure_destroy(self._h)
            self._h = None

    __del__ = close

    def _call(self, name, *args):
        rc = getattr(lib(), "spm_closure_" + name)(self._h, *args)
        if rc:
            raise SpmError(rc, "spm_closure_" + name, lib().spm_last_error(None).decode())

    def init(self, seed):
        self._call("init", seed)

    def get_params(self):
        p = np.zeros(self.NPARAM, dtype=np.float32)
        self._call("get_params", _fp(p))
        return p

    def put_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        self._call("put_params", _fp(p))

    def set_scaling(self, shift, scale, oscale):
        a, b, c = (np.ascontiguousarray(v, dtype=np.float32) for v in (shift, scale, oscale))
        self._call("set_scaling", _fp(a), _fp(b), _fp(c))

    def get_scaling(self):
        a, b, c = np.zeros(self.IN, np.float32), np.zeros(self.IN, np.float32), np.zeros(6, np.float32)
        self._call("get_scaling", _fp(a), _fp(b), _fp(c))
        return a, b, c

    def forward(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.zeros((x.shape[0], self.OUT), dtype=np.float32)
        self._call("forward", _fp(x), _fp(y), x.shape[0])
        return y

    def gradient(self, x, t):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        g = np.zeros(self.NPARAM, dtype=np.float32)
        loss = np.zeros(1, dtype=np.float32)
        self._call("gradient", _fp(x), _fp(t), x.shape[0], _fp(g), _fp(loss))
        return float(loss[0]), g

    def set_data(self, x, t):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        self._call("set_data", _fp(x), _fp(t), x.shape[0])

    def train(self, steps, batch, lr=1e-3):
        loss = np.zeros(steps, dtype=np.float32)
        self._call("train", steps, batch, C.c_float(lr), _fp(loss))
        return loss


class Ensemble:
    """Independent SPM boxes on one GPU advanced together (an spm_ensemble,
    include/spm_ensemble.h): one captured graph replay steps every box.
    `boxes[b]` is an SPM view of box b for particles, fields and the rest."""

    def __init__(self, params_list):
        n = len(params_list)
        arr = (abi.SpmParams * n)(*params_list)
        h = C.c_void_p()
        rc = lib().spm_ensemble_create(n, arr, C.byref(h))
        if rc:
            raise SpmError(rc, "spm_ensemble_create", lib().spm_last_error(None).decode())
        self._h = h
        self.params = list(params_list)
        self.boxes = [SPM._wrap(lib().spm_ensemble_box(h, b), params_list[b], self) for b in range(n)]

    def close(self):
        if getattr(self, "_h", None):
            for bx in self.boxes:
                bx._h = None
            lib().spm_ensemble_destroy(self._h)
            self._h = None

    __del__ = close

    def __len__(self):
        return len(self.boxes)

    def _call(self, name, *args):
        rc = getattr(lib(), "spm_ensemble_" + name)(self._h, *args)
        if rc:
            raise SpmError(rc, "spm_ensemble_" + name, lib().spm_ensemble_last_error(self._h).decode())

    def step(self, nsteps=1):
        self._call("step", nsteps)

    def set_graph(self, on):
        self._call("set_graph", 1 if on else 0)

    def autotune(self, steps_per_candidate=3):
        self._call("autotune", steps_per_candidate)

    def tuning(self):
        """(lanes in use, ms per ensemble step at 1/2/4/8 lanes, per-box ms)."""
        lanes = C.c_int()
        ms = np.zeros(4)
        box = np.zeros(len(self.boxes))
        self._call("tuning", C.byref(lanes), _dp(ms), _dp(box))
        return lanes.value, ms, box


def autotune_report():
    """The transform autotuner's choices, one line per tuned pass."""
    n = lib().spm_autotune_report(None, 0)
    buf = C.create_string_buffer(n + 1)
    lib().spm_autotune_report(buf, n + 1)
    return buf.value.decode()
