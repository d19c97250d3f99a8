"""The reference driver's parameter file (pb_read_param_file; host only):
the reference defaults, "key: value" lines split at ':' and ' ', unknown
keys and keys without a value ignored (periodicBrinkman.cu:44-87)."""
import pytest

import periodicbrinkman_amd as pb

DEFAULTS = dict(Lx=1.0, Ly=1.0, Lz=1.0, d=0.6, e=0.0, alpha=1.0, M=4, boxSize=1, toPrint=0)


def test_defaults_are_the_reference_ones():
    assert pb.driver_defaults() == DEFAULTS


def test_values_override_defaults(tmp_path):
    f = tmp_path / "param.txt"
    f.write_text("Lx: 2.5\nLy:0.75\nLz 1.25\nd: 0.125\ne: 0.01\nalpha: 10\nM: 3\nboxSize: 5\ntoPrint: 1\n")
    p = pb.read_param_file(f)
    assert p == dict(Lx=2.5, Ly=0.75, Lz=1.25, d=0.125, e=0.01, alpha=10.0, M=3, boxSize=5, toPrint=1)


def test_partial_file_unknown_keys_and_empty_values(tmp_path):
    f = tmp_path / "param.txt"
    f.write_text("alpha: 3\n\nfoo: 7\nd:\n   M  :  2  \nLx\n")
    p = pb.read_param_file(f)
    want = dict(DEFAULTS, alpha=3.0, M=2)
    assert p == want


def test_missing_file_is_a_status(tmp_path):
    with pytest.raises(pb.PBError) as ei:
        pb.read_param_file(tmp_path / "nope.txt")
    assert ei.value.status == pb.PB_ERR_INVALID
