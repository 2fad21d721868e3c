"""KDMP v2 header parsing through the C-ABI (io/dump.zig:82-98), no GPU:
the reference's known-answer header of restart_tests.zig:71-123 (5x4x3,
t = 12.3456789, nstep 4242, frame 7) and its ReadError cases."""
import struct

import pytest

from goldens import kdmp_bytes, read_kdmp
from koral_zig_amd import hip

import numpy as np


def _header(ver=2, magic=b"KDMP"):
    return magic + struct.pack("<IIIIIdQI", ver, 5, 4, 3, 13, 12.3456789, 4242, 7)


def test_header_known_answer():
    h = hip.parse_kdmp_header(_header())
    assert h == dict(nx=5, ny=4, nz=3, nv=13, t=12.3456789, nstep=4242, out_idx=7)
    assert len(_header()) == 44


@pytest.mark.parametrize("data,err", [(b"KDMX" + _header()[4:], "BadMagic"), (_header(ver=1), "BadVersion"),
                                      (_header()[:43], "Truncated")])
def test_header_rejects(data, err):
    with pytest.raises(RuntimeError, match=err):
        hip.parse_kdmp_header(data)


def test_python_restatement_round_trip():
    p = np.arange(3 * 4 * 5 * 13, dtype=float).reshape(3, 4, 5, 13) * 0.25
    b = kdmp_bytes(p, 1.5, 9, 2)
    d = read_kdmp(b)
    assert (d["nx"], d["ny"], d["nz"], d["nv"], d["t"], d["nstep"], d["out_idx"]) == (5, 4, 3, 13, 1.5, 9, 2)
    assert np.array_equal(d["p"], p)
    assert hip.parse_kdmp_header(b)["nstep"] == 9
