"""The C++ command-line program's argument handling, on the CPU (it stops
before any device call)."""
import os
import subprocess

import pytest

EXE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "kapsel_old_amd", "bin", "spm")


@pytest.mark.skipif(not os.path.exists(EXE), reason="kapsel_old_amd/bin/spm not built")
def test_usage_and_bad_flags():
    r = subprocess.run([EXE], capture_output=True, text=True)
    assert r.returncode == 1 and "usage: spm -I" in r.stderr
    r = subprocess.run([EXE, "-Ia.udf", "-Ob.udf", "-R", "c.udf", "--schedule", "sideways"],
                       capture_output=True, text=True)
    assert r.returncode == 1 and "--schedule takes" in r.stderr
    r = subprocess.run([EXE, "-Ia.udf", "-Ob.udf", "-Rc.udf", "-X"], capture_output=True, text=True)
    assert r.returncode == 1 and "unknown flag -X" in r.stderr


@pytest.mark.skipif(not os.path.exists(EXE), reason="kapsel_old_amd/bin/spm not built")
def test_missing_input_is_reported(tmp_path):
    r = subprocess.run([EXE, "-I" + str(tmp_path / "none.udf"), "-Oo.udf", "-Rr.udf"], capture_output=True, text=True)
    assert r.returncode == 1 and "spm_udf_open" in r.stderr
