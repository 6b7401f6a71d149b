"""The restart file format and directory handling, on the CPU: nested
dictionaries of plain values and numpy arrays round-trip exactly, nothing but
plain data is accepted or executed, and the reference's delete_old rotation
(Restart.pyx:92-152) keeps the newest two dumps and the retained times."""
import os

import numpy as np
import pytest

from pycles_amd.Restart import Restart, read_file, write_file


def test_round_trip_is_exact(tmp_path):
    rng = np.random.default_rng(0)
    data = {"PV": {"values": rng.normal(size=(5, 12, 4, 8)).astype(np.float32), "nv": 5,
                   "index_name": ["u", "v", "w", "s", "qt"], "bc_type": np.array([1.0, 1.0, -1.0, 1.0, 1.0]),
                   "base": {"s": rng.normal(size=12), "qt": rng.normal(size=12)}},
            "TS": {"t": 30.000000000000004, "dt": 0.7}, "last_restart_time": 30.0,
            "counts": np.arange(7, dtype=np.int64), "flag": np.bool_(True)}
    p = str(tmp_path / "r.pcr")
    write_file(p, data)
    back = read_file(p)
    assert back["TS"] == data["TS"] and back["last_restart_time"] == 30.0 and back["flag"] is True
    assert back["PV"]["index_name"] == data["PV"]["index_name"]
    for a, b in ((back["PV"]["values"], data["PV"]["values"]), (back["counts"], data["counts"]),
                 (back["PV"]["base"]["qt"], data["PV"]["base"]["qt"]),
                 (back["PV"]["bc_type"], data["PV"]["bc_type"])):
        assert a.dtype == b.dtype and a.shape == b.shape and np.array_equal(a, b)


def test_only_plain_data(tmp_path):
    p = str(tmp_path / "r.pcr")
    with pytest.raises(TypeError):
        write_file(p, {"x": np.array([object()], dtype=object)})
    with pytest.raises(TypeError):
        write_file(p, {"x": {1, 2}})
    with open(p, "wb") as f:
        f.write(b"\x80\x04K\x01.")  # a pickle stream
    with pytest.raises(ValueError, match="not a pycles_amd restart file"):
        read_file(p)


def test_truncated_file_is_rejected(tmp_path):
    p = str(tmp_path / "r.pcr")
    write_file(p, {"a": np.ones(1000)})
    with open(p, "r+b") as f:
        f.truncate(os.path.getsize(p) - 100)
    with pytest.raises(ValueError, match="truncated"):
        read_file(p)


def test_delete_old_rotation(tmp_path):
    nl = {"meta": {"uuid": "abcdef12345", "simname": "t"}, "output": {"output_root": str(tmp_path) + "/"},
          "restart": {"frequency": 10.0, "delete_old": True, "times_retained": [10]}}
    Re = Restart(nl).initialize()
    assert Re.enabled and Re.restart_path.endswith("Output.t.12345/Restart")
    for t in (10.0, 20.0, 30.0, 40.0):
        Re.last_restart_time = t
        Re.restart_data["x"] = np.full(3, t)
        Re.write()
    names = sorted(os.listdir(Re.restart_path))
    # the newest is _new, the one before _recent, older ones removed unless retained
    assert names == ["10", "30_recent", "40_new"], names
    Re.cleanup()
    assert sorted(os.listdir(Re.restart_path)) == ["10", "30", "40"]
    assert np.array_equal(read_file(os.path.join(Re.restart_path, "40", "0.pcr"))["x"], np.full(3, 40.0))


def test_no_restart_section_writes_nothing():
    assert not Restart({}).enabled
