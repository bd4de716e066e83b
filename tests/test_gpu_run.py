"""The command-line driver (python -m kapsel_old_amd.run) end to end on a
UDF input derived from the reference's Udf/input.udf (user_specify
particles, 32^3): snapshot records, the restart file, and resumption from
the restart UDF and from an exact state file.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "udf")


def make_input(tmp_path, num_snap, name):
    src = open(os.path.join(GOLDEN, "input.udf")).read()
    edits = [('"uniform_random"', '"user_specify"'),
             ('[{5,1.50000000000000,0.0,"NONE"', '[{1,1.50000000000000,0.0,"NONE"'),
             ('{5,0.50000000000000,0.0,"NONE"', '{1,0.50000000000000,0.0,"NONE"'),
             ("output:{10,100,", "output:{2,%d," % num_snap)]
    for old, new in edits:
        assert src.count(old) == 1, old
        src = src.replace(old, new)
    with open(os.path.join(GOLDEN, "define.udf")) as f:
        (tmp_path / "define.udf").write_text(f.read())
    path = tmp_path / name
    path.write_text(src)
    return str(path)


def saved(path):
    from kapsel_old_amd import udf
    with udf.UDF(path) as u:
        r = udf.read_restart(u, (32, 32, 32), 2)
    return r


def test_run_records_and_restart(tmp_path):
    from kapsel_old_amd import abi, udf
    from kapsel_old_amd.run import main
    from kapsel_old_amd.spm import SPM
    inp = make_input(tmp_path, 3, "in3.udf")
    out, res = str(tmp_path / "out.udf"), str(tmp_path / "res.udf")
    print("driver run MSTEP 6", flush=True)
    assert main(["-I" + inp, "-O" + out, "-R" + res, "-Ddefine.udf", "--quiet"]) == 0
    with udf.UDF(out) as u:
        assert u.totalRecord() == 4
        ts = []
        for r in range(4):
            u.jump(r)
            ts.append(u.get_int("t"))
        assert ts == [0, 2, 4, 6]
        u.jump(0)
        np.testing.assert_array_equal(u.get_doubles("Particles[].R.x"), [16.0, 4.0])
    r = saved(res)
    assert r["ts"] == 6
    # the same run through the API: 7 steps (ts = 0..6)
    with udf.UDF(inp) as u:
        prm, _, _ = u.params()
        p = u.user_particles()
    g = SPM(prm)
    g.put_particles(p)
    g.init_fluid()
    g.step(7)
    assert g.get_time() == (7, r["time"])
    np.testing.assert_array_equal(r["zeta0"], g.get_field("zeta0"))
    np.testing.assert_array_equal(r["particles"]["x"], g.get_particles()["x"])
    g.close()


def test_run_resume(tmp_path):
    from kapsel_old_amd.run import main
    full = make_input(tmp_path, 5, "in5.udf")
    part = make_input(tmp_path, 3, "in3.udf")
    d = str(tmp_path)
    print("uninterrupted MSTEP 10", flush=True)
    main(["-I" + full, "-O" + d + "/o_full.udf", "-R" + d + "/r_full.udf", "--quiet"])
    print("first half MSTEP 6", flush=True)
    main(["-I" + part, "-O" + d + "/o_a.udf", "-R" + d + "/r_a.udf", "--state", d + "/s_a.bin", "--quiet"])
    # the restart file, with the run extended to Num_snap 5, is the next input
    text = open(d + "/r_a.udf").read()
    assert text.count("output:{2,3,") == 1
    open(d + "/r_a5.udf", "w").write(text.replace("output:{2,3,", "output:{2,5,"))
    print("resume from restart UDF", flush=True)
    main(["-I" + d + "/r_a5.udf", "-O" + d + "/o_b.udf", "-R" + d + "/r_b.udf", "--quiet"])
    print("resume from state file", flush=True)
    main(["-I" + full, "-O" + d + "/o_c.udf", "-R" + d + "/r_c.udf", "--resume-state", d + "/s_a.bin", "--quiet"])
    ref, b, c = saved(d + "/r_full.udf"), saved(d + "/r_b.udf"), saved(d + "/r_c.udf")
    assert ref["ts"] == b["ts"] == c["ts"] == 10
    # exact from the state file
    for k in ("zeta0", "zeta1", "uk_dc"):
        np.testing.assert_array_equal(c[k], ref[k], err_msg=k)
    assert c["time"] == ref["time"]
    for k in ("x", "v", "omega", "q", "fr_previous"):
        np.testing.assert_array_equal(c["particles"][k], ref["particles"][k], err_msg=k)
    # from the restart UDF: q is renormalized on restore, as in the reference
    for k in ("zeta0", "zeta1"):
        assert np.abs(b[k] - ref[k]).max() <= 1e-12 * np.abs(ref[k]).max(), k
    for k in ("x", "v", "omega"):
        assert np.abs(b["particles"][k] - ref["particles"][k]).max() <= 1e-12 * max(
            1.0, np.abs(ref["particles"][k]).max()), k
    assert b["time"] == pytest.approx(ref["time"], rel=1e-15)
