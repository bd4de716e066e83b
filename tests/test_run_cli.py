"""The driver's flags (file_get, input.cxx:1819-1877) on the CPU: values
attached to the flag as the reference takes them (-Iinput.udf) or given
separately, the required ones enforced."""
import pytest

from kapsel_old_amd.run import parse_args


def test_attached_and_separate_values():
    a = parse_args(["-Iin.udf", "-Oout.udf", "-Rres.udf", "-Ddefine.udf"])
    assert (a.input, a.output, a.restart, a.define) == ("in.udf", "out.udf", "res.udf", "define.udf")
    b = parse_args(["-I", "in.udf", "-O", "out.udf", "-R", "res.udf", "--state", "s.bin", "--fp32"])
    assert (b.input, b.output, b.restart, b.state, b.fp32, b.schedule) == ("in.udf", "out.udf", "res.udf", "s.bin",
                                                                           True, "fused")


def test_required_flags():
    with pytest.raises(SystemExit):
        parse_args(["-Iin.udf", "-Oout.udf"])  # no -R
