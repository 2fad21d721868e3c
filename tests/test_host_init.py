"""Caller-side host library (koral-zig_amd/host) pinned against the
reference's metric goldens and its own Zig self-golden post-init snapshot."""
import numpy as np
import pytest

import kzo
from goldens import DevTracker, read_kgld, read_kslf
from koral_zig_amd import abi_types as T
from koral_zig_amd import host

COORDS = {1: 1, 2: 2, 4: 0, 10: 3}  # golden ids (golden.zig:284-292) -> kz_host ids


def test_metric_points_golden():
    """metric_golden_tests.zig:41-103 — gcov 1e-13; bl/ks/mink gcon/gdet 1e-13,
    dlgdet/Christoffels 1e-12; MKS2-derived 1e-8 (C's own export spread)."""
    inp, out = read_kgld("metric/metric_points.kgld", 5, 101)
    trk = {k: DevTracker() for k in ("gcov", "gcon", "gcon_m", "gdet", "gdet_m", "dl", "dl_m", "kr", "kr_m", "gtt")}
    for i, r in enumerate(inp):
        c = COORDS[int(r[0])]
        m = c == 3
        o = host.metric_point(c, r[1:5])
        trk["gcov"].add(out[i, 0:16], o[0:16], i)
        trk["gcon_m" if m else "gcon"].add(out[i, 16:32], o[16:32], i)
        trk["gdet_m" if m else "gdet"].add(out[i, 32], o[32], i)
        trk["dl_m" if m else "dl"].add(out[i, 33:36], o[33:36], i)
        trk["kr_m" if m else "kr"].add(out[i, 36:100], o[36:100], i)
        trk["gtt"].add(out[i, 100], o[100], i)
    for k, b in dict(gcov=1e-13, gcon=1e-13, gcon_m=1e-8, gdet=1e-13, gdet_m=1e-8, dl=1e-12, dl_m=1e-8, kr=1e-12,
                     kr_m=1e-8, gtt=1e-13).items():
        assert trk[k].ok(b), (k, trk[k].max_dev)


def test_dxdx_golden():
    """metric_golden_tests.zig:105-127 — MKS2<->BL Jacobians (the pairs the
    PUFFY boundary uses), 1e-12."""
    inp, out = read_kgld("metric/coco_dxdx.kgld", 6, 20)
    n = 0
    for i, r in enumerate(inp):
        fr, to = int(r[0]), int(r[1])
        if (fr, to) == (10, 1):
            j = host.dxdx(0, r[2:6])
        elif (fr, to) == (1, 10):
            j = host.dxdx(1, r[2:6])
        else:
            continue
        n += 1
        assert DevTracker().add(out[i, 4:20], j.ravel()).ok(1e-12), i
    assert n > 0


def test_krzysie_grid_golden():
    """metric_golden_tests.zig:129-171 — sqrt(-g)-corrected Christoffels on the
    384x360 PUFFY grid (grid centres exact to 1e-14; Christoffels 1e-8)."""
    inp, out = read_kgld("metric/krzysie_grid.kgld", 2, 66)
    grid = T.grid_puffy(384, 360)
    t = host.metric_tables(grid)
    (sz, sy, sx), (ngz, ngy, ngx) = T.padded_shape(grid)
    kris = t["kris"].reshape(sz, sy, sx, 64)
    dx = (grid.maxx - grid.minx) / grid.nx
    dy = (grid.maxy - grid.miny) / grid.ny
    trk = DevTracker()
    for i, r in enumerate(inp):
        ix, iy = int(r[0]), int(r[1])
        xc = 0.5 * ((grid.minx + ix * dx) + (grid.minx + (ix + 1) * dx))
        yc = 0.5 * ((grid.miny + iy * dy) + (grid.miny + (iy + 1) * dy))
        assert DevTracker().add(out[i, 0:2], [xc, yc]).ok(1e-14)
        trk.add(out[i, 2:66], kris[ngz, iy + ngy, ix + ngx], i)
    assert trk.ok(1e-8), trk.max_dev


@pytest.fixture(scope="module")
def fast_init():
    s = read_kslf("selfgolden/puffy_fast.kslf")
    grid = T.grid_puffy(s["nx"], s["ny"])
    phys = T.physics_puffy()
    tab = host.metric_tables(grid)
    sim = kzo.Sim(grid, phys, tab)
    host.init_puffy(sim, grid, tab, phys)
    return s, sim


def test_init_vs_selfgolden(fast_init):
    """The whole init (selfscenarios.zig:99-104: initAll + initTimestepGuess)
    against the Zig self-golden post-init snapshot of PUFFY 32x30, ghosts
    included.  Bits are not reproducible without Zig's libm; the torus
    quadrature and transcendentals agree to <= 1e-10 field-relative."""
    s, sim = fast_init
    p, u, f = sim.get_state()
    f0 = s["fields"][0]
    assert f0["step"] == 0
    pr = f0["p"].reshape(p.shape)
    ur = f0["u"].reshape(u.shape)
    for a, b in ((p, pr), (u, ur)):
        scale = np.maximum(np.abs(b).max(axis=(0, 1, 2)), 1e-300)
        dev = (np.abs(a - b) / scale).max()
        assert dev < 1e-10, dev
    assert (f == f0["flags"].reshape(f.shape)).all()
    c = sim.counters()
    assert abs(c["tstepdenmax"] - s["scalars"][0, 3]) <= 1e-12 * s["scalars"][0, 3]


def test_perturbation_is_seeded_and_bounded():
    """The seeded init perturbation: reproducible, |xi| < 1, B untouched."""
    grid = T.grid_puffy(8, 6)
    (sz, sy, sx), _ = T.padded_shape(grid)
    p = np.ones((sz, sy, sx, T.NV))
    a = host.perturb(grid, p.copy(), 0x4B5A4D493335355F, 1e-4)
    b = host.perturb(grid, p.copy(), 0x4B5A4D493335355F, 1e-4)
    c = host.perturb(grid, p.copy(), 0x1234, 1e-4)
    assert np.array_equal(a, b) and not np.array_equal(a, c)
    dom = a[:, 3:-3, 3:-3]
    assert (dom[..., T.B1:T.B3 + 1] == 1.0).all()
    assert (np.abs(dom[..., [0, 1, 2, 3, 4, 9, 10, 11, 12]] - 1.0) < 1e-4).all()
    xs = np.array([host.uniform(7, i, 3) for i in range(20000)])
    assert xs.min() >= -1.0 and xs.max() < 1.0 and abs(xs.mean()) < 0.02


def test_calc_b_from_a_golden():
    """step_golden_tests.zig:340-380 — calc_BfromA (corner average + curl,
    overwrite) on a periodic 32x32 Minkowski grid vs C, 1e-14."""
    inp, out = read_kgld("flux/bfroma.kgld", 5, 3)
    tp = 2.0 * np.pi
    grid = T.kz_grid(nx=32, ny=32, nz=1, ng=3, minx=0.0, maxx=tp, miny=0.0, maxy=tp, minz=0.0, maxz=1.0)
    (sz, sy, sx), (ngz, ngy, ngx) = T.padded_shape(grid)
    a = np.zeros((32, 32, 3))
    for r in inp:
        a[int(r[1]), int(r[0])] = r[2:5]
    # periodic ghosts in both dimensions (bc_x = bc_y = periodic)
    iy = (np.arange(sy) - ngy) % 32
    ix = (np.arange(sx) - ngx) % 32
    p = np.zeros((sz, sy, sx, T.NV))
    p[0, :, :, 6:9] = a[iy][:, ix]
    g20 = np.zeros((sz * sy * sx, 20))
    g20[:, 19] = 1.0  # Minkowski: sqrt(-g) = 1
    p = host.calc_b_from_a(grid, {"g": g20.ravel()}, p)
    got = np.array([p[0, int(r[1]) + ngy, int(r[0]) + ngx, 6:9] for r in inp])
    t = DevTracker().add(out, got)
    assert t.ok(1e-14), t.max_dev
