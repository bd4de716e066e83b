"""The C ABI: header layout vs the ctypes mirror, exported symbols.
(CPU only: loads the libraries but launches nothing.)"""
import ctypes as C
import os
import re
import subprocess

import pytest

from kapsel_old_amd import abi
from oracle.oracle import ref_available, ref_lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "spm.h")
HEADERS = [os.path.join(ROOT, "include", h) for h in sorted(os.listdir(os.path.join(ROOT, "include")))
           if h.endswith(".h")]


def declared_functions():
    src = "\n".join(open(h).read() for h in HEADERS)
    return sorted(set(re.findall(r"^\s*(?:const\s+char\s*\*|int|void|size_t|spm_ctx)\s*\*?\s*(spm_\w+)\s*\(", src, re.M)))


def test_header_layout_matches_ctypes(tmp_path):
    prog = tmp_path / "layout.c"
    fields = ["spec", "x", "v", "f_hydro", "fr", "omega", "torque_hydro", "QR", "q", "q_old", "mass",
              "surface_dw"]
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "%s"' % HEADER, "int main(void){",
             'printf("%zu %zu\\n", sizeof(spm_particle), sizeof(spm_params));']
    for f in fields:
        lines.append('printf("%%zu\\n", offsetof(spm_particle, %s));' % f)
    for f in ["nx", "dx", "n_particles", "mass_ratio", "rotation", "fix_cell", "seed", "precision", "device"]:
        lines.append('printf("%%zu\\n", offsetof(spm_params, %s));' % f)
    lines.append("return 0;}")
    prog.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-o", str(exe), str(prog)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()
    vals = [int(v) for v in out]
    assert vals[0] == 1000 == C.sizeof(abi.SpmParticle)
    assert vals[1] == C.sizeof(abi.SpmParams)
    for f, v in zip(fields, vals[2:2 + len(fields)]):
        assert getattr(abi.SpmParticle, f).offset == v, f
    for f, v in zip(["nx", "dx", "n_particles", "mass_ratio", "rotation", "fix_cell", "seed", "precision",
                     "device"], vals[2 + len(fields):]):
        assert getattr(abi.SpmParams, f).offset == v, f


PARTICLE_MEMBERS = [f[0] for f in abi.SpmParticle._fields_]


@pytest.mark.skipif(not ref_available(), reason="oracle/_ref (reference variable.h build) absent")
def test_particle_layout_matches_reference_variable_h(tmp_path):
    """Every member offset of spm_particle (include/spm.h) and of the ctypes
    mirror equals offsetof() of the reference's own Particle (variable.h:27-75)
    as compiled from the reference tree into oracle/_ref."""
    L = ref_lib()
    buf = (C.c_size_t * 64)()
    n = L.ref_particle_layout(buf, 64)
    assert n == 1 + len(PARTICLE_MEMBERS)
    ref_size, ref_off = buf[0], list(buf[1:n])
    prog = tmp_path / "offs.c"
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "%s"' % HEADER, "int main(void){",
             'printf("%zu\\n", sizeof(spm_particle));']
    lines += ['printf("%%zu\\n", offsetof(spm_particle, %s));' % m for m in PARTICLE_MEMBERS]
    lines.append("return 0;}")
    prog.write_text("\n".join(lines))
    exe = tmp_path / "offs"
    subprocess.run(["gcc", "-std=c11", "-o", str(exe), str(prog)], check=True)
    hdr = [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    assert hdr[0] == ref_size == C.sizeof(abi.SpmParticle) == 1000
    for m, r, h in zip(PARTICLE_MEMBERS, ref_off, hdr[1:]):
        assert h == r, m
        assert getattr(abi.SpmParticle, m).offset == r, m


def test_library_exports_every_declared_symbol():
    from kapsel_old_amd.spm import LIB_PATH
    if not os.path.exists(LIB_PATH):
        pytest.skip("libspm.so not built")
    names = declared_functions()
    assert "spm_step" in names and "spm_create" in names and len(names) >= 30
    out = subprocess.run(["nm", "-D", "--defined-only", LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = set(re.findall(r" T (spm_\w+)$", out, re.M))
    missing = [n for n in names if n not in exported]
    assert not missing, missing
    lib = C.CDLL(LIB_PATH)
    for n in names:
        assert hasattr(lib, n)


def test_strerror_without_device():
    from kapsel_old_amd.spm import lib
    L = lib()
    assert L.spm_strerror(0) == b"ok"
    assert L.spm_strerror(abi.SPM_ERR_NONFINITE).startswith(b"non-finite")


def test_python_front_end_covers_the_call_surface():
    from kapsel_old_amd.spm import SPM
    for m in ["NS_solver_slavedEuler", "Make_phi_u_particle", "Calc_f_hydro_correct_precision", "Force",
              "MD_solver_position", "MD_solver_velocity", "Add_random_force_thermostat",
              "Make_f_particle_dt_sole", "U2zeta_k", "Zeta_k2u", "init_fluid", "step"]:
        assert callable(getattr(SPM, m))


def test_fft_workspace_size_and_size_checks_without_device():
    """spm_fft_workspace_size is host arithmetic, and spm_fft3d rejects bad
    sizes before touching a device."""
    from kapsel_old_amd.spm import lib
    L = lib()
    ws = L.spm_fft_workspace_size(256, 256, 256)
    # four 256-byte aligned root tables of 16-byte entries: 256, 256, 128, 129
    assert ws == 4096 + 4096 + 2048 + 2304
    # mixed lengths 48, 80, 96, 160, 192: tables of 48, 80, 48, 49 entries
    assert L.spm_fft_workspace_size(48, 80, 96) == 768 + 1280 + 768 + 1024
    for bad in (12, 100, 2048, 4):
        assert L.spm_fft_workspace_size(bad, 64, 64) == 0
    assert L.spm_fft3d(100, 64, 64, 1, None, None, 0, None) == abi.SPM_ERR_INVALID
    assert b"powers of two" in L.spm_last_error(None)
