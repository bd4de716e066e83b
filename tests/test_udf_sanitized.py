"""The UDF reader (kapsel_old_amd/csrc/udf.cpp, host C++) built with
AddressSanitizer and UndefinedBehaviorSanitizer on the CPU, run over every
UDF fixture and, where the reference tree is present, its 28 .udf files:
each whole file and its truncated prefixes must parse or be rejected
without an out-of-bounds access, a leak or undefined behaviour."""
import glob
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden", "udf")
REFERENCE = "/root/reference"


def _build(tmp_path):
    exe = str(tmp_path / "udf_prefixes")
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
           "-fno-omit-frame-pointer", os.path.join(ROOT, "kapsel_old_amd", "csrc", "udf.cpp"),
           os.path.join(ROOT, "tests", "sanitize", "udf_prefixes.cpp"), "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        if "sanitize" in r.stderr and ("cannot find" in r.stderr or "unrecognized" in r.stderr):
            pytest.skip("g++ has no sanitizer runtime here")
        raise AssertionError(r.stderr[-3000:])
    return exe


def _groups():
    files = sorted(glob.glob(os.path.join(GOLDEN, "*.udf")))
    groups = [(GOLDEN, files, 7)]
    if os.path.isdir(REFERENCE):
        ref = sorted(p for p in glob.glob(os.path.join(REFERENCE, "**", "*.udf"), recursive=True))
        by_dir = {}
        for p in ref:
            by_dir.setdefault(os.path.dirname(p), []).append(p)
        groups += [(d, fs, 61) for d, fs in sorted(by_dir.items())]
    return groups


def test_udf_reader_under_asan_ubsan(tmp_path):
    exe = _build(tmp_path)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")
    nfiles = 0
    for k, (src_dir, files, step) in enumerate(_groups()):
        work = tmp_path / ("g%d" % k)
        work.mkdir()
        for p in glob.glob(os.path.join(src_dir, "*.udf")):
            shutil.copy(p, work)
        r = subprocess.run([exe, str(step), str(work)] + files, capture_output=True, text=True, env=env,
                           timeout=600)
        assert r.returncode == 0, (src_dir, r.stdout[-2000:], r.stderr[-4000:])
        assert "ERROR: AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr, r.stderr[-4000:]
        nfiles += len(files)
        print(r.stdout, flush=True)
    assert nfiles >= 9
