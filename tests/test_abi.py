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
                                          ("kz_implicit.h", hip.lib), ("kz_host.h", host.lib)])
def test_exports(header, libfn):
    names = declared(header)
    assert len(names) >= 5
    L = libfn()
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_binding_covers_header():
    names = set(declared("kz_step.h")) | set(declared("kz_cellops.h")) | set(declared("kz_implicit.h"))
    assert names == set(hip._SIGS), names ^ set(hip._SIGS)


def test_struct_sizes():
    """ctypes mirrors match the C layout the headers fix."""
    assert C.sizeof(T.kz_floor_params) == 7 * 8 + 4 * 4
    assert C.sizeof(T.kz_rad_params) == 6 * 8
    assert C.sizeof(T.kz_implicit_params) == 15 * 8 + 8 * 4
    assert C.sizeof(T.kz_opac_params) == 26 * 8 + 8 * 4
    assert C.sizeof(T.kz_counters) == 5 * 8 + 5 * 8
    assert C.sizeof(T.kz_grid) == 4 * 4 + 6 * 8


@pytest.mark.parametrize("frame", [-1, 2])
def test_create_rejects_unknown_floor_frame(frame):
    """koral/solve/invert.zig:35: the b^2/rho floor frame is drift (0) or
    ZAMO (1)."""
    grid = T.grid_puffy(16, 12)
    tab = host.metric_tables(grid)
    phys = T.physics_puffy()
    phys.floors.b2rhofloorframe = frame
    with pytest.raises(RuntimeError, match="InvalidConfig"):
        hip.Sim(grid, phys, tab)


def test_create_rejects_collapsed_grid():
    grid = T.grid_puffy(16, 12)
    tab = host.metric_tables(grid)
    grid.ny = 1
    with pytest.raises(RuntimeError, match="Unsupported|-7"):
        hip.Sim(grid, T.physics_puffy(), tab)


def test_create_rejects_radvisc_without_opacities():
    """sim.zig:319-363: radiative viscosity needs the opacities."""
    grid = T.grid_puffy(16, 12)
    tab = host.metric_tables(grid)
    phys = T.physics_puffy(radviscosity=True)
    phys.has_opac = 0
    with pytest.raises(RuntimeError, match="InvalidConfig"):
        hip.Sim(grid, phys, tab)


def test_metric_tables_symmetric_bitwise():
    """The device keeps g and the Christoffels as upper triangles; MetricCache
    (and the host library restating it) builds them symmetric bit for bit."""
    import numpy as np
    for nx, ny in ((32, 30), (96, 90)):
        grid = T.grid_puffy(nx, ny)
        t = host.metric_tables(grid)
        kr = t["kris"].reshape(-1, 4, 4, 4).view(np.uint64)
        assert np.array_equal(kr, kr.transpose(0, 1, 3, 2))
        for k in ("g", "gb0", "gb1", "gb2"):
            a = t[k].reshape(-1, 4, 5)[:, :, :4].view(np.uint64)
            assert np.array_equal(a, a.transpose(0, 2, 1)), k


def test_create_rejects_asymmetric_metric():
    """An asymmetric block is rejected before any device work."""
    grid = T.grid_puffy(16, 12)
    tab = host.metric_tables(grid)
    g = tab["g"].copy()
    g[20 * 50 + 1] = g[20 * 50 + 5] * (1 + 1e-15) + 1e-300  # g[0][1] != g[1][0] in cell 50
    with pytest.raises(RuntimeError, match="InvalidConfig"):
        hip.Sim(grid, T.physics_puffy(), dict(tab, g=g))
