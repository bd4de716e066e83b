"""The C++ command-line program (kapsel_old_amd/bin/spm, csrc/spm_main.cpp)
against the Python driver on the same inputs: byte-identical snapshot and
restart files for a fresh run, for a resume from the restart UDF and for a
resume from an exact state file; flags attached or separate, as file_get
takes them."""
import os
import subprocess

import pytest

from tests.test_gpu_run import make_input

pytestmark = pytest.mark.gpu

EXE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "kapsel_old_amd", "bin", "spm")


def run_exe(*args):
    if not os.path.exists(EXE):
        raise RuntimeError("%s is not built; run python -m kapsel_old_amd.build" % EXE)
    r = subprocess.run([EXE, *args, "--quiet"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    return r


def same_file(a, b):
    with open(a, "rb") as fa, open(b, "rb") as fb:
        return fa.read() == fb.read()


def test_program_equals_python_driver(tmp_path):
    from kapsel_old_amd.run import main
    inp = make_input(tmp_path, 3, "in3.udf")
    d = str(tmp_path)
    print("python driver, MSTEP 6", flush=True)
    assert main(["-I" + inp, "-O" + d + "/o_py.udf", "-R" + d + "/r_py.udf", "--state", d + "/s_py.bin",
                 "--quiet"]) == 0
    print("C++ program, MSTEP 6", flush=True)
    run_exe("-I" + inp, "-O" + d + "/o_cc.udf", "-R", d + "/r_cc.udf", "--state", d + "/s_cc.bin")
    assert same_file(d + "/o_py.udf", d + "/o_cc.udf")
    assert same_file(d + "/r_py.udf", d + "/r_cc.udf")
    assert same_file(d + "/s_py.bin", d + "/s_cc.bin")


def test_program_resumes_like_python_driver(tmp_path):
    from kapsel_old_amd.run import main
    part = make_input(tmp_path, 3, "in3.udf")
    full = make_input(tmp_path, 5, "in5.udf")
    d = str(tmp_path)
    run_exe("-I" + part, "-O" + d + "/o_a.udf", "-R" + d + "/r_a.udf", "--state", d + "/s_a.bin")
    text = open(d + "/r_a.udf").read()
    assert text.count("output:{2,3,") == 1
    open(d + "/r_a5.udf", "w").write(text.replace("output:{2,3,", "output:{2,5,"))
    print("resume from the restart UDF", flush=True)
    run_exe("-I" + d + "/r_a5.udf", "-O" + d + "/o_b_cc.udf", "-R" + d + "/r_b_cc.udf")
    main(["-I" + d + "/r_a5.udf", "-O" + d + "/o_b_py.udf", "-R" + d + "/r_b_py.udf", "--quiet"])
    assert same_file(d + "/r_b_cc.udf", d + "/r_b_py.udf")
    assert same_file(d + "/o_b_cc.udf", d + "/o_b_py.udf")
    print("resume from the state file", flush=True)
    run_exe("-I" + full, "-O" + d + "/o_c_cc.udf", "-R" + d + "/r_c_cc.udf", "--resume-state", d + "/s_a.bin")
    main(["-I" + full, "-O" + d + "/o_c_py.udf", "-R" + d + "/r_c_py.udf", "--resume-state", d + "/s_a.bin",
          "--quiet"])
    assert same_file(d + "/r_c_cc.udf", d + "/r_c_py.udf")


def test_program_and_driver_write_the_same_stress_lines(tmp_path):
    """--stress on Lees-Edwards input: Mean_shear_stress's header and one
    line per step (MSTEP + 1 steps), the same bytes from both drivers."""
    from kapsel_old_amd.run import main
    from tests.test_gpu_shear import _driver_input
    inp = _driver_input(tmp_path, 3, "lep6.udf", particles=True)
    d = str(tmp_path)
    run_exe("-I" + inp, "-O" + d + "/o_cc.udf", "-R" + d + "/r_cc.udf", "--stress", d + "/s_cc.txt")
    main(["-I" + inp, "-O" + d + "/o_py.udf", "-R" + d + "/r_py.udf", "--stress", d + "/s_py.txt", "--quiet"])
    assert same_file(d + "/s_cc.txt", d + "/s_py.txt")
    lines = open(d + "/s_cc.txt").read().splitlines()
    assert lines[0].startswith("#1:time 2:shear_rate 3:shear_strain_temporal")
    assert len(lines) == 1 + 7
    vals = [[float(x) for x in ln.split()] for ln in lines[1:]]
    assert all(len(v) == 7 for v in vals) and vals[-1][1] == 4.0
    assert same_file(d + "/r_cc.udf", d + "/r_py.udf")
