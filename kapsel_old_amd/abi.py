"""ctypes mirror of include/spm.h (the C ABI of libspm.so).

Struct layouts must match the header byte for byte; tests/test_abi.py
compiles the header with gcc and compares sizes and offsets.
"""
import ctypes as C

DIM = 3
MAX_SPECIES = 8
MT_N = 624

# status codes (enum spm_status)
SPM_OK = 0
SPM_ERR_INVALID = 1
SPM_ERR_OOM = 2
SPM_ERR_HIP = 3
SPM_ERR_NONFINITE = 4
SPM_ERR_UNSUPPORTED = 5
SPM_ERR_NO_DEVICE = 6

# field ids (enum spm_field)
FIELD_ZETA0, FIELD_ZETA1 = 0, 1
FIELD_U0, FIELD_U1, FIELD_U2 = 2, 3, 4
FIELD_UP0, FIELD_UP1, FIELD_UP2 = 5, 6, 7
FIELD_PHI = 8
FIELD_F0, FIELD_F1, FIELD_F2 = 9, 10, 11
NUM_FIELDS = 12
FIELD_IDS = {
    "zeta0": 0, "zeta1": 1, "u0": 2, "u1": 3, "u2": 4, "up0": 5, "up1": 6,
    "up2": 7, "phi": 8, "f0": 9, "f1": 10, "f2": 11,
}

LJ_NONE, LJ_FULL, LJ_WCA = -1, 0, 1
LJ_12_6, LJ_24_12, LJ_36_18 = 0, 1, 2
FP64, FP32 = 0, 1

D3 = C.c_double * DIM
D33 = (C.c_double * DIM) * DIM


class SpmParams(C.Structure):
    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
        ("dx", C.c_double),
        ("rho", C.c_double), ("eta", C.c_double),
        ("kbt", C.c_double),
        ("alpha_v", C.c_double), ("alpha_o", C.c_double),
        ("dt", C.c_double),
        ("a", C.c_double), ("a_xi", C.c_double),
        ("n_particles", C.c_int),
        ("n_species", C.c_int),
        ("mass_ratio", C.c_double * MAX_SPECIES),
        ("rotation", C.c_int),
        ("lj_truncate", C.c_int),
        ("lj_powers", C.c_int),
        ("epsilon", C.c_double),
        ("g", C.c_double),
        ("g_direction", C.c_int),
        ("fix_cell", C.c_int * DIM),
        ("seed", C.c_ulong),
        ("precision", C.c_int),
        ("device", C.c_int),
    ]


class SpmElectrolyte(C.Structure):
    """spm_electrolyte of include/spm.h (constitutive_eq.Electrolyte)."""
    _fields_ = [
        ("n_spec", C.c_int),
        ("valency", C.c_double * 2),
        ("onsager", C.c_double * 2),
        ("dielectric", C.c_double),
        ("elementary_charge", C.c_double),
        ("debye_length", C.c_double),
        ("surface_charge", C.c_double * MAX_SPECIES),
        ("external_field", C.c_int),
        ("ac", C.c_int),
        ("e_ext", C.c_double * DIM),
        ("angular_frequency", C.c_double),
    ]


def make_electrolyte(*, n_spec=2, valency=(1.0, -1.0), onsager=(1.0, 1.0), dielectric=1.0,
                     elementary_charge=1.0, debye_length=3.0, surface_charge=(50.0,), e_ext=(0.0, 0.0, 0.0),
                     external_field=None, ac=0, angular_frequency=0.0):
    """Build an SpmElectrolyte; external_field defaults to on when e_ext != 0."""
    e = SpmElectrolyte()
    e.n_spec = n_spec
    for i in range(2):
        e.valency[i] = valency[i] if i < len(valency) else 0.0
        e.onsager[i] = onsager[i] if i < len(onsager) else 0.0
    e.dielectric, e.elementary_charge, e.debye_length = dielectric, elementary_charge, debye_length
    for i, q in enumerate(surface_charge):
        e.surface_charge[i] = q
    for d in range(DIM):
        e.e_ext[d] = e_ext[d]
    e.external_field = int(any(x != 0.0 for x in e_ext)) if external_field is None else external_field
    e.ac, e.angular_frequency = ac, angular_frequency
    return e


class SpmQuaternion(C.Structure):
    _fields_ = [("s", C.c_double), ("v", D3)]


PARTICLE_VEC_FIELDS = [
    "x", "x_previous", "x_nopbc", "v", "v_old", "v_slip",
    "f_hydro", "f_hydro_previous", "f_hydro1", "f_slip", "f_slip_previous",
    "fr", "fr_previous", "omega", "omega_old", "omega_slip",
    "torque_hydro", "torque_hydro_previous", "torque_hydro1", "torque_slip",
    "torque_slip_previous", "momentum_depend_fr",
]


class SpmParticle(C.Structure):
    _fields_ = (
        [("spec", C.c_int)]
        + [(n, D3) for n in PARTICLE_VEC_FIELDS]
        + [("QR", D33), ("QR_old", D33),
           ("q", SpmQuaternion), ("q_old", SpmQuaternion),
           ("mass", C.c_double), ("mass_center", D3), ("inertia", D33),
           ("surface_mass", C.c_double), ("surface_mass_center", D3),
           ("surface_inertia", D33), ("surface_dv", D3), ("surface_dw", D3)]
    )


assert C.sizeof(SpmParticle) == 1000


def make_params(nx, ny=None, nz=None, *, dx=1.0, rho=1.0, eta=1.0, kbt=0.0,
                alpha_v=1.0, alpha_o=1.0, dt=0.05, a=5.0, a_xi=2.0,
                n_particles=0, mass_ratio=(1.0,), rotation=1,
                lj_truncate=LJ_NONE, lj_powers=LJ_12_6, epsilon=1.0, g=0.0,
                g_direction=2, fix_cell=(1, 1, 1), seed=0, precision=FP64,
                device=0):
    """Build an SpmParams with the reference's defaults for unnamed fields."""
    p = SpmParams()
    p.nx = nx
    p.ny = nx if ny is None else ny
    p.nz = nx if nz is None else nz
    p.dx, p.rho, p.eta, p.kbt = dx, rho, eta, kbt
    p.alpha_v, p.alpha_o, p.dt = alpha_v, alpha_o, dt
    p.a, p.a_xi = a, a_xi
    p.n_particles = n_particles
    p.n_species = len(mass_ratio)
    for i, m in enumerate(mass_ratio):
        p.mass_ratio[i] = m
    p.rotation = rotation
    p.lj_truncate, p.lj_powers, p.epsilon = lj_truncate, lj_powers, epsilon
    p.g, p.g_direction = g, g_direction
    for d in range(DIM):
        p.fix_cell[d] = fix_cell[d]
    p.seed = seed
    p.precision = precision
    p.device = device
    return p


import numpy as _np

PARTICLE_DTYPE = _np.dtype(SpmParticle)


def new_particles(n):
    """Zeroed particle array in the reference layout, q = identity."""
    p = _np.zeros(n, dtype=PARTICLE_DTYPE)
    p["q"]["s"] = 1.0
    p["q_old"]["s"] = 1.0
    return p


def random_positions(n, box, min_dist, seed=12345, max_tries=1_000_000):
    """Uniform non-overlapping positions in [0, box) by random sequential
    addition with rejection at min_dist under periodic images, as
    init_particle.cxx:31-48 does for DISTRIBUTION uniform_random (which uses
    glibc rand(); here numpy's PCG64 seeded with `seed`)."""
    box = _np.asarray(box, dtype=float)
    rng = _np.random.Generator(_np.random.PCG64(seed))
    pos = _np.empty((n, 3))
    k = 0
    tries = 0
    while k < n:
        cand = rng.random(3) * box
        tries += 1
        if tries > max_tries:
            raise RuntimeError("could not place particles without overlap")
        if k:
            d = pos[:k] - cand
            d -= _np.round(d / box) * box
            if (_np.einsum("ij,ij->i", d, d) <= min_dist * min_dist).any():
                continue
        pos[k] = cand
        k += 1
    return pos


# spm_le_stress (include/spm.h), in order
LE_STRESS_KEYS = ("time", "shear_rate", "strain", "lj_dev_stress", "hydro_stress", "hydro_stress_new",
                  "viscosity")
