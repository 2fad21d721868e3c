"""The C-ABI libraries load and export every entry point their headers
declare; argument checks that run before any device work reject what this
build does not implement.  No GPU needed."""
import ctypes as C
import re
from pathlib import Path

import pytest

from koral_zig_amd import abi_types as T
from koral_zig_amd import hip, host

ROOT = Path(__file__).resolve().parent.parent


def declared(header):
    text = (ROOT / "include" / header).read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(kz_[a-z0-9_]+)\s*\(", text)))


@pytest.mark.parametrize("header,libfn", [("kz_step.h", hip.lib), ("kz_cellops.h", hip.lib),
                                          ("kz_host.h", host.lib)])
def test_exports(header, libfn):
    names = declared(header)
    assert len(names) >= 10
    L = libfn()
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_binding_covers_header():
    names = set(declared("kz_step.h")) | set(declared("kz_cellops.h"))
    assert names == set(hip._SIGS), names ^ set(hip._SIGS)


def test_struct_sizes():
    """ctypes mirrors match the C layout the headers fix."""
    assert C.sizeof(T.kz_floor_params) == 7 * 8 + 4 * 4
    assert C.sizeof(T.kz_rad_params) == 6 * 8
    assert C.sizeof(T.kz_implicit_params) == 15 * 8 + 8 * 4
    assert C.sizeof(T.kz_opac_params) == 26 * 8 + 8 * 4
    assert C.sizeof(T.kz_counters) == 5 * 8 + 5 * 8
    assert C.sizeof(T.kz_grid) == 4 * 4 + 6 * 8


@pytest.mark.parametrize("field", ["radviscosity", "dynamo", "reduceorderatbh", "reduceorderafterfixup"])
def test_create_rejects_unsupported(field):
    grid = T.grid_puffy(16, 12)
    tab = host.metric_tables(grid)
    phys = T.physics_puffy()
    setattr(phys, field, 1)
    with pytest.raises(RuntimeError, match="Unsupported|-7"):
        hip.Sim(grid, phys, tab)


def test_create_rejects_3d():
    grid = T.grid_puffy(16, 12, 4)
    tab = host.metric_tables(grid)
    with pytest.raises(RuntimeError, match="Unsupported|-7"):
        hip.Sim(grid, T.physics_puffy(), tab)
