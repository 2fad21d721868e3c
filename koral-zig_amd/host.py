"""Binding of libkz_host.so — the caller-side tables and initial state.

Restates the once-per-run host work of koral-zig (MetricCache.init, the
PUFFY initializer, calc_BfromA, the boundary tables) so a run can be set up
for the per-step kernels.  See include/kz_host.h for the entry points.
"""
import ctypes as C
from pathlib import Path

import numpy as np

from . import abi_types as T

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "lib" / "libkz_host.so"
_lib = None

MP_PUFFY = dict(a=0.0, mksr0=0.1, mksh0=0.9)  # Physics.mp (puffy.zig:87)


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} missing: run __graft_entry__.build()")
        L = C.CDLL(str(LIB_PATH))
        D = C.POINTER(C.c_double)
        G = C.POINTER(T.kz_grid)
        L.kz_host_cell_count.restype = C.c_long
        L.kz_host_cell_count.argtypes = [G]
        L.kz_host_face_count.restype = C.c_long
        L.kz_host_face_count.argtypes = [G, C.c_int]
        L.kz_host_metric_tables.argtypes = [G, C.c_double, C.c_double, C.c_double] + [D] * 9
        L.kz_host_metric_point.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, D, D]
        L.kz_host_geom_batch.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, C.c_int64, D, D]
        L.kz_host_dxdx.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, D, D]
        L.kz_host_bc_xhi_tables.argtypes = [G, C.c_double, C.c_double, C.c_double, D]
        L.kz_host_puffy_prepinit.argtypes = [G, C.c_double, C.c_double, C.c_double, C.c_double,
                                             C.POINTER(T.kz_opac_params), C.POINTER(T.kz_rad_params),
                                             C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
                                             D, D, D]
        L.kz_host_calc_b_from_a.argtypes = [G, D, D]
        L.kz_host_uniform.restype = C.c_double
        L.kz_host_uniform.argtypes = [C.c_uint64, C.c_uint64, C.c_uint64]
        L.kz_host_perturb.argtypes = [G, C.c_uint64, C.c_double, C.c_double, C.c_int, D]
        L.kz_host_geometry_bl.argtypes = [G, C.c_double, C.c_double, C.c_double, C.c_long, C.c_long, C.c_long,
                                          C.c_void_p]
        L.kz_host_bl_tables.argtypes = [G, C.c_double, C.c_double, C.c_double, D]
        L.kz_host_radius_bl.argtypes = [G, C.c_double, C.c_double, C.c_double, D]
        L.kz_host_shell_widths_bl.argtypes = [G, C.c_double, C.c_double, C.c_double, C.c_long, D, D]
        L.kz_host_reduce_order_ix_max.restype = C.c_int32
        L.kz_host_reduce_order_ix_max.argtypes = [G, C.c_double, C.c_double, C.c_double]
        _lib = L
    return _lib


def metric_tables(grid, mp=MP_PUFFY):
    """MetricCache arrays (+ the xhi BC tables) for `grid`, reference layouts."""
    L = lib()
    nc = L.kz_host_cell_count(C.byref(grid))
    t = dict(g=np.zeros(nc * 20), gcon=np.zeros(nc * 20), kris=np.zeros(nc * 64))
    for d in range(3):
        nf = L.kz_host_face_count(C.byref(grid), d)
        t[f"gb{d}"] = np.zeros(nf * 20)
        t[f"gconb{d}"] = np.zeros(nf * 20)
    L.kz_host_metric_tables(C.byref(grid), mp["a"], mp["mksr0"], mp["mksh0"], T.ptr(t["g"]), T.ptr(t["gcon"]),
                            T.ptr(t["kris"]), T.ptr(t["gb0"]), T.ptr(t["gconb0"]), T.ptr(t["gb1"]),
                            T.ptr(t["gconb1"]), T.ptr(t["gb2"]), T.ptr(t["gconb2"]))
    t["bcx"] = np.zeros(grid.nz * grid.ny * grid.ng * 66)
    L.kz_host_bc_xhi_tables(C.byref(grid), mp["a"], mp["mksr0"], mp["mksh0"], T.ptr(t["bcx"]))
    t["mp"] = dict(mp)
    return t


def metric_point(coords, x4, mp=MP_PUFFY):
    out = np.zeros(101)
    x = np.ascontiguousarray(x4, np.float64)
    lib().kz_host_metric_point(coords, mp["a"], mp["mksr0"], mp["mksh0"], T.ptr(x), T.ptr(out))
    return out


def geom_batch(coords, x4, mp=MP_PUFFY):
    """kz_geom records (GEOM_DOUBLES each) at the points x4[n, 4]."""
    x = np.ascontiguousarray(x4, np.float64)
    n = len(x)
    out = np.zeros((n, T.GEOM_DOUBLES))
    lib().kz_host_geom_batch(coords, mp["a"], mp["mksr0"], mp["mksh0"], n, T.ptr(x), T.ptr(out))
    return out


def dxdx(direction, x4, mp=MP_PUFFY):
    out = np.zeros(16)
    x = np.ascontiguousarray(x4, np.float64)
    lib().kz_host_dxdx(direction, mp["a"], mp["mksr0"], mp["mksh0"], T.ptr(x), T.ptr(out))
    return out.reshape(4, 4)


# PUFFY atmosphere / torus constants (puffy.zig:84-106 Physics defaults)
PUFFY_INIT = dict(mass=10.0, rhoatmmin=1.0e-24, atm_tgas=1.0e10, atm_trad_init=3.0e5, lt_kappa=6.0e1,
                  maxbeta=1.0 / 20.0)


def puffy_prepinit(grid, tables, phys, init=PUFFY_INIT):
    """Domain primitives after prepinit+init (A_phi in B3, ENTR = S(rho,u))."""
    (sz, sy, sx), _ = T.padded_shape(grid)
    p = np.zeros((sz, sy, sx, T.NV))
    mp = tables["mp"]
    lib().kz_host_puffy_prepinit(C.byref(grid), mp["a"], mp["mksr0"], mp["mksh0"], phys.gam,
                                 C.byref(phys.opac), C.byref(phys.rad), init["mass"], init["rhoatmmin"],
                                 init["atm_tgas"], init["atm_trad_init"], init["lt_kappa"],
                                 T.ptr(tables["g"]), T.ptr(tables["gcon"]), T.ptr(p))
    return p


def calc_b_from_a(grid, tables, p):
    p = np.ascontiguousarray(p)
    rc = lib().kz_host_calc_b_from_a(C.byref(grid), T.ptr(tables["g"]), T.ptr(p))
    if rc:
        raise RuntimeError(T.ERRORS.get(rc, rc))
    return p


def perturb(grid, p, seed, amp=1.0e-4, gam=5.0 / 3.0, rho_u_only=False):
    p = np.ascontiguousarray(p)
    lib().kz_host_perturb(C.byref(grid), seed, amp, gam, int(rho_u_only), T.ptr(p))
    return p


def uniform(seed, cell, var):
    return lib().kz_host_uniform(seed, cell, var)


def geometry_bl(grid, ix, iy, iz=0, mp=MP_PUFFY):
    out = np.zeros(47)
    lib().kz_host_geometry_bl(C.byref(grid), mp["a"], mp["mksr0"], mp["mksh0"], ix, iy, iz,
                              out.ctypes.data_as(C.c_void_p))
    return out


def init_puffy(engine, grid, tables, phys, init=PUFFY_INIT, seed=None, amp=1.0e-4, rho_u_only=False, p0=None):
    """The ko.c init sequence (puffy.zig:873-884 initAllWith) driven through
    an engine (the HIP sim handle, or the CPU restatement in tests):
    prepinit+init -> set_bc -> calc_BfromA -> set_bc -> postinit (beta norm)
    -> initTimestepGuess; with `seed`, the seeded 1e-4 perturbation of every
    non-B primitive (of rho and u only with rho_u_only) is applied after
    postinit, followed by p2u + set_bc.  `p0` is puffy_prepinit's result when
    the caller computed it already.
    Returns the beta-normalisation factor."""
    p = puffy_prepinit(grid, tables, phys, init) if p0 is None else p0
    engine.set_state(p=p)
    engine.p2u_domain()
    engine.set_bc(True)
    p, _, _ = engine.get_state()
    p = calc_b_from_a(grid, tables, np.ascontiguousarray(p))
    engine.set_state(p=p)
    engine.p2u_domain()
    engine.set_bc(True)
    fac = engine.postinit(init["maxbeta"])
    if seed is not None:
        p, _, _ = engine.get_state()
        p = perturb(grid, np.ascontiguousarray(p), seed, amp, phys.gam, rho_u_only)
        engine.set_state(p=p)
        engine.p2u_domain()
        engine.set_bc(False)
    engine.init_timestep_guess()
    return fac


BL_REC = 52  # gBL[16] GBL[16] alpha gdet r theta J_mks2->bl[16]


def bl_tables(grid, mp=MP_PUFFY):
    """The BL sidecar (MetricCache.blGeom / jacMy2Bl) per padded cell."""
    L = lib()
    nc = L.kz_host_cell_count(C.byref(grid))
    out = np.zeros(nc * BL_REC)
    L.kz_host_bl_tables(C.byref(grid), mp["a"], mp["mksr0"], mp["mksh0"], T.ptr(out))
    return out


def radius_bl(grid, mp=MP_PUFFY):
    """BL radius of each radial index (scalars.zig:47-50 radiusBL)."""
    r = np.zeros(grid.nx)
    lib().kz_host_radius_bl(C.byref(grid), mp["a"], mp["mksr0"], mp["mksh0"], T.ptr(r))
    return r


def reduce_order_ix_max(grid, mp=MP_PUFFY):
    """sim.zig:481-495: the largest radial index (ghosts included) whose cell
    centre lies inside the BL horizon, for kz_physics.reduce_order_ix_max
    (reduceorderatbh); INT32_MIN when there is none."""
    return int(lib().kz_host_reduce_order_ix_max(C.byref(grid), mp["a"], mp["mksr0"], mp["mksh0"]))


def shell_widths_bl(grid, ix, mp=MP_PUFFY):
    """BL theta and phi widths of the cells of radial shell ix, [nz][ny] each
    (scalars.zig:188-204)."""
    dth = np.zeros((grid.nz, grid.ny))
    dph = np.zeros((grid.nz, grid.ny))
    lib().kz_host_shell_widths_bl(C.byref(grid), mp["a"], mp["mksr0"], mp["mksh0"], int(ix), T.ptr(dth), T.ptr(dph))
    return dth, dph
