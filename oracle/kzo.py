"""ctypes binding of the CPU restatement (oracle/libkzoracle.so).

TEST INFRASTRUCTURE: imported only by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker.  The product never loads it.
"""
import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "libkzoracle.so"


def build(force=False):
    """Compile the restatement with its Makefile (gcc, -ffp-contract=off)."""
    if force or not LIB_PATH.exists():
        subprocess.run(["make", "-s", "-B" if force else "-s", "-C", str(HERE)], check=True)


def _pkg_types():
    import importlib.util
    import sys
    name = "koral_zig_amd"
    if name not in sys.modules:
        root = HERE.parent / "koral-zig_amd"
        spec = importlib.util.spec_from_file_location(name, root / "__init__.py",
                                                      submodule_search_locations=[str(root)])
        mod = importlib.util.module_from_spec(spec)
        sys.modules[name] = mod
        spec.loader.exec_module(mod)
    from koral_zig_amd import abi_types
    return abi_types


T = _pkg_types()
_D = C.POINTER(C.c_double)
_I = C.POINTER(C.c_int32)
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        L = C.CDLL(str(LIB_PATH))
        vp = C.c_void_p
        L.kzo_sim_create.restype = vp
        L.kzo_sim_create.argtypes = [C.POINTER(T.kz_grid), C.POINTER(T.kz_physics)] + [_D] * 9 + [_D, _I]
        for n in ("kzo_sim_destroy", "kzo_sim_p2u_domain", "kzo_sim_init_timestep_guess", "kzo_flux_ct"):
            getattr(L, n).argtypes = [vp]
            getattr(L, n).restype = None
        L.kzo_sim_set_state.argtypes = [vp, _D, _D]
        L.kzo_sim_set_state.restype = None
        L.kzo_sim_get_state.argtypes = [vp, _D, _D, _I]
        L.kzo_sim_get_state.restype = None
        L.kzo_sim_get_scal.argtypes = [vp, _D]
        L.kzo_sim_get_scal.restype = None
        L.kzo_sim_calc_divb.argtypes = [vp, _D]
        L.kzo_sim_calc_divb.restype = None
        L.kzo_sim_get_flb.argtypes = [vp, C.c_int, _D]
        L.kzo_sim_get_flb.restype = None
        L.kzo_sim_set_flb.argtypes = [vp, C.c_int, _D]
        L.kzo_sim_set_flb.restype = None
        L.kzo_sim_set_bl_tables.argtypes = [vp, _D]
        L.kzo_sim_set_bl_tables.restype = None
        L.kzo_sim_calc_rij_visc_total.argtypes = [vp, C.c_double]
        L.kzo_sim_get_rijvisc.argtypes = [vp, _D]
        L.kzo_sim_get_rijvisc.restype = None
        L.kzo_sim_apply_dynamo.argtypes = [vp, C.c_double]
        L.kzo_sim_get_scaleth.argtypes = [vp, _D]
        L.kzo_sim_get_scaleth.restype = None
        L.kzo_step_function.argtypes = [C.c_double, C.c_double]
        L.kzo_step_function.restype = C.c_double
        L.kzo_visc_coeff.argtypes = [C.c_double] * 6
        L.kzo_visc_coeff.restype = C.c_double
        L.kzo_shear_from_gradients.argtypes = [_D, _D, _D, _D, _D, _D, _D]
        L.kzo_shear_from_gradients.restype = None
        L.kzo_dynamo_delta_a.argtypes = [C.POINTER(T.kz_dynamo_params)] + [C.c_double] * 12 + [_D]
        L.kzo_dynamo_delta_a.restype = C.c_double
        L.kzo_damp_bphi.argtypes = [C.c_double] * 7
        L.kzo_damp_bphi.restype = C.c_double
        L.kzo_sim_shear_lab.argtypes = [vp, C.c_long, C.c_long, C.c_long, _D]
        L.kzo_sim_get_dyn_a.argtypes = [vp, _D]
        L.kzo_sim_get_dyn_a.restype = None
        L.kzo_sim_mimic_dynamo.argtypes = [vp, C.c_double]
        L.kzo_r_horizon_bl.argtypes = [C.c_double]
        L.kzo_r_horizon_bl.restype = C.c_double
        L.kzo_r_isco_bl.argtypes = [C.c_double]
        L.kzo_r_isco_bl.restype = C.c_double
        L.kzo_sim_scalars.argtypes = [vp, C.c_long, C.c_long, C.c_long, vp, vp, vp, vp]
        L.kzo_sim_counters.argtypes = [vp, C.POINTER(T.kz_counters)]
        L.kzo_sim_counters.restype = None
        L.kzo_sim_set_time.argtypes = [vp, C.c_double, C.c_double, C.c_double]
        L.kzo_sim_set_time.restype = None
        L.kzo_sim_set_bc.argtypes = [vp, C.c_int]
        L.kzo_sim_calc_u2p.argtypes = [vp]
        L.kzo_sim_op_explicit.argtypes = [vp, C.c_double]
        L.kzo_sim_op_implicit.argtypes = [vp, C.c_double]
        L.kzo_sim_do_correct.argtypes = [vp]
        L.kzo_sim_step.argtypes = [vp, _D]
        L.kzo_sim_postinit.argtypes = [vp, C.c_double]
        L.kzo_sim_postinit.restype = C.c_double
        _lib = L
    return _lib


def set_threads(n):
    """Threads of the oracle sim's cell-parallel passes (bitwise the serial result)."""
    lib().kzo_set_threads(int(n))


def _geoms(g):
    g = np.ascontiguousarray(g, dtype=np.float64)
    assert g.ndim == 2 and g.shape[1] == T.GEOM_DOUBLES
    return g


def p2u(pp, geom, gam):
    pp = np.ascontiguousarray(pp, np.float64)
    g = _geoms(geom)
    uu = np.zeros_like(pp)
    lib().kzo_p2u_batch(C.c_int64(len(pp)), T.ptr(pp), T.ptr(g), C.c_double(gam), T.ptr(uu))
    return uu


def u2p_solver_w(uu, guess, geom, etype, gam):
    uu = np.ascontiguousarray(uu, np.float64)
    pp = np.array(guess, np.float64, copy=True, order="C")
    ret = np.zeros(len(uu), np.int32)
    lib().kzo_u2p_solver_w_batch(C.c_int64(len(uu)), T.ptr(uu), T.ptr(pp), T.ptr(_geoms(geom)),
                                 C.c_int32(etype), C.c_double(gam), T.ptr(ret, C.c_int32))
    return ret, pp


def u2p_mhd(uu, guess, geom, gam, floors):
    uu = np.ascontiguousarray(uu, np.float64)
    pp = np.array(guess, np.float64, copy=True, order="C")
    res = np.zeros(len(uu), T.U2P_RESULT_DTYPE)
    lib().kzo_u2p_mhd_batch(C.c_int64(len(uu)), T.ptr(uu), T.ptr(pp), T.ptr(_geoms(geom)), C.c_double(gam),
                            C.byref(floors), res.ctypes.data_as(C.c_void_p))
    return res, pp


def check_floors_mhd(pp, geom, gam, floors):
    pp = np.array(pp, np.float64, copy=True, order="C")
    ret = np.zeros(len(pp), np.int32)
    ng = C.c_uint64(0)
    lib().kzo_check_floors_mhd_batch(C.c_int64(len(pp)), T.ptr(pp), T.ptr(_geoms(geom)), C.c_double(gam),
                                     C.byref(floors), T.ptr(ret, C.c_int32), C.byref(ng))
    return ret, pp, ng.value


def u2p_rad(uu, guess, geom, rp):
    uu = np.ascontiguousarray(uu, np.float64)
    pp = np.array(guess, np.float64, copy=True, order="C")
    res = np.zeros(len(uu), T.U2PRAD_RESULT_DTYPE)
    lib().kzo_u2p_rad_batch(C.c_int64(len(uu)), T.ptr(uu), T.ptr(pp), T.ptr(_geoms(geom)), C.byref(rp),
                            res.ctypes.data_as(C.c_void_p))
    return res, pp


def check_floors_rad(pp, geom, rp):
    pp = np.array(pp, np.float64, copy=True, order="C")
    ret = np.zeros(len(pp), np.int32)
    lib().kzo_check_floors_rad_batch(C.c_int64(len(pp)), T.ptr(pp), T.ptr(_geoms(geom)), C.byref(rp),
                                     T.ptr(ret, C.c_int32))
    return ret, pp


def solve_implicit(pp, geom, dt, gam, rp, op, ip):
    pp = np.array(pp, np.float64, copy=True, order="C")
    uu = np.zeros_like(pp)
    dt = np.ascontiguousarray(np.broadcast_to(np.asarray(dt, np.float64), (len(pp),)))
    res = np.zeros(len(pp), T.IMPLICIT_RESULT_DTYPE)
    lib().kzo_solve_implicit_lab_batch(C.c_int64(len(pp)), T.ptr(uu), T.ptr(pp), T.ptr(_geoms(geom)), T.ptr(dt),
                                       C.c_double(gam), C.byref(rp), C.byref(op), C.byref(ip),
                                       res.ctypes.data_as(C.c_void_p))
    return res, pp, uu


def reconstruct(order, theta, u5, dx5):
    n = len(u5)
    order = np.ascontiguousarray(np.broadcast_to(np.asarray(order, np.int32), (n,)))
    theta = np.ascontiguousarray(np.broadcast_to(np.asarray(theta, np.float64), (n,)))
    l = np.zeros(n)
    r = np.zeros(n)
    lib().kzo_reconstruct_batch(C.c_int64(n), T.ptr(order, C.c_int32), T.ptr(theta),
                                T.ptr(np.ascontiguousarray(u5, np.float64)),
                                T.ptr(np.ascontiguousarray(dx5, np.float64)), T.ptr(l), T.ptr(r))
    return l, r


def laxf(fl, fr, ul, ur, ag):
    a = [np.ascontiguousarray(x, np.float64) for x in (fl, fr, ul, ur, ag)]
    out = np.zeros_like(a[0])
    lib().kzo_laxf_batch(C.c_int64(len(out)), *[T.ptr(x) for x in a], T.ptr(out))
    return out


def flux_prime(pp, idim, geom, gam):
    pp = np.ascontiguousarray(pp, np.float64)
    n = len(pp)
    idim = np.ascontiguousarray(np.broadcast_to(np.asarray(idim, np.int32), (n,)))
    ff = np.zeros_like(pp)
    err = np.zeros(n, np.int32)
    lib().kzo_flux_prime_batch(C.c_int64(n), T.ptr(pp), T.ptr(idim, C.c_int32), T.ptr(_geoms(geom)),
                               C.c_double(gam), T.ptr(ff), T.ptr(err, C.c_int32))
    return ff, err


def gas_wavespeeds(pp, geom, gam, active=(1, 1, 0)):
    pp = np.ascontiguousarray(pp, np.float64)
    a = np.zeros((len(pp), 6))
    act = np.array(active, np.int32)
    lib().kzo_gas_wavespeeds_batch(C.c_int64(len(pp)), T.ptr(pp), T.ptr(_geoms(geom)), C.c_double(gam),
                                   T.ptr(act, C.c_int32), T.ptr(a))
    return a


def rad_wavespeeds(pp, geom, tau3, active=(1, 1, 0)):
    pp = np.ascontiguousarray(pp, np.float64)
    a = np.zeros((len(pp), 12))
    act = np.array(active, np.int32)
    lib().kzo_rad_wavespeeds_batch(C.c_int64(len(pp)), T.ptr(pp), T.ptr(_geoms(geom)),
                                   T.ptr(np.ascontiguousarray(tau3, np.float64)), T.ptr(act, C.c_int32), T.ptr(a))
    return a


def rij(pp, geom):
    pp = np.ascontiguousarray(pp, np.float64)
    out = np.zeros((len(pp), 21))
    lib().kzo_rij_batch(C.c_int64(len(pp)), T.ptr(pp), T.ptr(_geoms(geom)), T.ptr(out))
    return out


def tij(pp, geom, gam):
    pp = np.ascontiguousarray(pp, np.float64)
    out = np.zeros((len(pp), 33))
    lib().kzo_tij_batch(C.c_int64(len(pp)), T.ptr(pp), T.ptr(_geoms(geom)), C.c_double(gam), T.ptr(out))
    return out


def radstate(pp, geom, gam, op, full=1):
    pp = np.ascontiguousarray(pp, np.float64)
    out = np.zeros((len(pp), 14))
    lib().kzo_radstate_batch(C.c_int64(len(pp)), T.ptr(pp), T.ptr(_geoms(geom)), C.c_double(gam), C.byref(op),
                             C.c_int32(full), T.ptr(out))
    return out


def opac_entry(pp, geom, gam, op):
    pp = np.ascontiguousarray(pp, np.float64)
    out = np.zeros((len(pp), 4))
    lib().kzo_opac_entry_batch(C.c_int64(len(pp)), T.ptr(pp), T.ptr(_geoms(geom)), C.c_double(gam), C.byref(op),
                               T.ptr(out))
    return out


def gi(pp, geom, gam, op):
    pp = np.ascontiguousarray(pp, np.float64)
    out = np.zeros((len(pp), 8))
    slim = np.zeros((len(pp), 8))
    lib().kzo_gi_batch(C.c_int64(len(pp)), T.ptr(pp), T.ptr(_geoms(geom)), C.c_double(gam), C.byref(op),
                       T.ptr(out), T.ptr(slim))
    return out, slim


def invert4(a16):
    a16 = np.ascontiguousarray(a16, np.float64).reshape(-1, 16)
    out = np.zeros_like(a16)
    lib().kzo_invert4_batch(C.c_int64(len(a16)), T.ptr(a16), T.ptr(out))
    return out


def invert4f(a16):
    a16 = np.ascontiguousarray(a16, np.float32).reshape(-1, 16)
    out = np.zeros_like(a16)
    lib().kzo_invert4f_batch(C.c_int64(len(a16)), T.ptr(a16, C.c_float), T.ptr(out, C.c_float))
    return out


MATH_FN = {"exp": 0, "log": 1, "pow": 2, "cbrt": 3, "log1p": 4, "sqrt": 5, "div": 6, "pow_from_log": 7}


def math_fn(name, x, y=None):
    x = np.ascontiguousarray(x, np.float64)
    yy = None if y is None else np.ascontiguousarray(np.broadcast_to(np.asarray(y, np.float64), x.shape))
    out = np.zeros_like(x)
    lib().kzo_math_batch(C.c_int64(x.size), C.c_int32(MATH_FN[name]), T.ptr(x), T.ptr(yy), T.ptr(out))
    return out


class Sim:
    """The restated Sim(config.puffy) over caller-built metric tables.

    `tables` is the dict from koral_zig_amd.host.metric_tables(); arrays are
    referenced, not copied, so the dict must outlive the Sim.
    """

    def __init__(self, grid, phys, tables):
        self.grid = grid
        self.phys = phys
        self.tables = tables
        err = C.c_int32(0)
        t = tables
        self.h = lib().kzo_sim_create(C.byref(grid), C.byref(phys), T.ptr(t["g"]), T.ptr(t["gcon"]),
                                      T.ptr(t["kris"]), T.ptr(t["gb0"]), T.ptr(t["gconb0"]), T.ptr(t["gb1"]),
                                      T.ptr(t["gconb1"]), T.ptr(t["gb2"]), T.ptr(t["gconb2"]), T.ptr(t["bcx"]),
                                      C.byref(err))
        if not self.h:
            raise RuntimeError(f"kzo_sim_create failed: {T.ERRORS.get(err.value, err.value)}")
        (sz, sy, sx), _ = T.padded_shape(grid)
        self.shape = (sz, sy, sx)
        self.ncell = sz * sy * sx

    def __del__(self):
        if getattr(self, "h", None):
            lib().kzo_sim_destroy(self.h)
            self.h = None

    def _chk(self, rc, what):
        if rc != 0:
            raise RuntimeError(f"oracle {what}: {T.ERRORS.get(rc, rc)}")

    def set_state(self, p=None, u=None):
        p = None if p is None else np.ascontiguousarray(p, np.float64)
        u = None if u is None else np.ascontiguousarray(u, np.float64)
        lib().kzo_sim_set_state(self.h, T.ptr(p), T.ptr(u))

    def get_state(self):
        p = np.zeros(self.shape + (T.NV,))
        u = np.zeros(self.shape + (T.NV,))
        f = np.zeros(self.shape + (T.NFLAGS,), np.int32)
        lib().kzo_sim_get_state(self.h, T.ptr(p), T.ptr(u), T.ptr(f, C.c_int32))
        return p, u, f

    def face_shape(self, dim):
        (sz, sy, sx) = self.shape
        return (sz + (dim == 2), sy + (dim == 1), sx + (dim == 0), T.NV)

    def get_flb(self, dim):
        f = np.zeros(self.face_shape(dim))
        lib().kzo_sim_get_flb(self.h, dim, T.ptr(f))
        return f

    def set_flb(self, dim, f):
        f = np.ascontiguousarray(f, np.float64)
        assert f.shape == self.face_shape(dim)
        lib().kzo_sim_set_flb(self.h, dim, T.ptr(f))

    def flux_ct(self):
        lib().kzo_flux_ct(self.h)

    def set_bl_tables(self, bl):
        """BL sidecar for radiative viscosity / dynamo (kept referenced)."""
        self._bl = np.ascontiguousarray(bl, np.float64)
        lib().kzo_sim_set_bl_tables(self.h, T.ptr(self._bl))

    def calc_rij_visc_total(self, global_dt):
        self._chk(lib().kzo_sim_calc_rij_visc_total(self.h, global_dt), "calc_rij_visc_total")
        r = np.zeros(self.shape + (16,))
        lib().kzo_sim_get_rijvisc(self.h, T.ptr(r))
        return r

    def shear_lab(self, ix, iy, iz=0):
        out = np.zeros(16)
        lib().kzo_sim_shear_lab(self.h, ix, iy, iz, T.ptr(out))
        return out.reshape(4, 4)

    def mimic_dynamo(self, dt):
        """Scale height, setBc, mimicDynamo (no final calcU2p); returns dyn_a."""
        self._chk(lib().kzo_sim_mimic_dynamo(self.h, dt), "mimic_dynamo")
        a = np.zeros(self.shape)
        lib().kzo_sim_get_dyn_a(self.h, T.ptr(a))
        return a

    def apply_dynamo(self, dt):
        self._chk(lib().kzo_sim_apply_dynamo(self.h, dt), "apply_dynamo")
        h = np.zeros(self.grid.nx)
        lib().kzo_sim_get_scaleth(self.h, T.ptr(h))
        return h

    def calc_divb(self):
        """ct.zig:336-381 calcDivB at every domain cell, [nz][ny][nx]."""
        g = self.grid
        out = np.zeros((g.nz, g.ny, g.nx))
        lib().kzo_sim_calc_divb(self.h, T.ptr(out))
        return out

    def get_scal(self):
        s = np.zeros(self.shape + (20,))
        lib().kzo_sim_get_scal(self.h, T.ptr(s))
        return s

    def scalars(self, dt, ix_h, ix_l, ix_s, dth_bl, dph_bl):
        """One scalars.dat row (dump.zig:383-411 scalarRow on one rank)."""
        dth = np.ascontiguousarray(dth_bl, np.float64)
        dph = np.ascontiguousarray(dph_bl, np.float64)
        out = np.zeros(6)
        cnt = np.zeros(2, np.uint64)
        self._chk(lib().kzo_sim_scalars(self.h, int(ix_h), int(ix_l), int(ix_s), dth.ctypes.data, dph.ctypes.data,
                                        out.ctypes.data, cnt.ctypes.data), "kzo_sim_scalars")
        c = self.counters()
        return {"t": c["t"], "dt": float(dt), "nstep": c["nstep"], "mass": out[0], "mdot": -out[1],
                "radlum": out[2], "totallum": out[3], "scaleheight": out[4], "max_pmag_ptot": out[5],
                "n_hd_fixup": int(cnt[0]), "n_radimp_fail": c["n_radimp_failures"], "n_nan": int(cnt[1]),
                "n_floorguard": c["n_guard_recoveries"]}

    def counters(self):
        c = T.kz_counters()
        lib().kzo_sim_counters(self.h, C.byref(c))
        return {n: getattr(c, n) for n, _ in T.kz_counters._fields_}

    def set_time(self, t, tstepdenmax, tstepdenmin):
        lib().kzo_sim_set_time(self.h, t, tstepdenmax, tstepdenmin)

    def p2u_domain(self):
        lib().kzo_sim_p2u_domain(self.h)

    def set_bc(self, ifinit=False):
        self._chk(lib().kzo_sim_set_bc(self.h, int(ifinit)), "set_bc")

    def calc_u2p(self):
        self._chk(lib().kzo_sim_calc_u2p(self.h), "calc_u2p")

    def op_explicit(self, dt):
        self._chk(lib().kzo_sim_op_explicit(self.h, dt), "op_explicit")

    def op_implicit(self, dt):
        self._chk(lib().kzo_sim_op_implicit(self.h, dt), "op_implicit")

    def do_correct(self):
        self._chk(lib().kzo_sim_do_correct(self.h), "do_correct")

    def init_timestep_guess(self):
        lib().kzo_sim_init_timestep_guess(self.h)

    def postinit(self, maxbeta):
        return lib().kzo_sim_postinit(self.h, maxbeta)

    def step(self, dt=None):
        if dt is None:
            self._chk(lib().kzo_sim_step(self.h, None), "step")
        else:
            d = C.c_double(dt)
            self._chk(lib().kzo_sim_step(self.h, C.pointer(d)), "step")


def true_residual(pp, pp00, dt, geom, gam, op, wp, we, wf):
    """implicit_golden_tests.zig:25-40 trueResidual over n records."""
    n = len(pp)
    pp = np.ascontiguousarray(pp, np.float64)
    pp00 = np.ascontiguousarray(pp00, np.float64)
    dt = np.ascontiguousarray(np.broadcast_to(np.asarray(dt, np.float64), (n,)))
    wp, we, wf = [np.ascontiguousarray(np.broadcast_to(np.asarray(x, np.int32), (n,))) for x in (wp, we, wf)]
    err = np.zeros(n)
    lib().kzo_true_residual_batch(C.c_int64(n), T.ptr(pp), T.ptr(pp00), T.ptr(dt), T.ptr(_geoms(geom)),
                                  C.c_double(gam), C.byref(op), T.ptr(wp, C.c_int32), T.ptr(we, C.c_int32),
                                  T.ptr(wf, C.c_int32), T.ptr(err))
    return err


WP_RAD, WP_MHD = 1, 2
WE_ENERGY, WE_ENTROPY = 0, 1
WF_LAB, WF_FF = 0, 1


def solve_4d_prim(pp00, geom, dt, gam, rp, op, ip, wp, we=WE_ENERGY, wf=WF_LAB):
    """One rung of the implicit solve from pp00 (implicit.zig solve4dPrim,
    no warm start, no ladder): (ok, iters, pp, internal uu)."""
    n = len(pp00)
    pp00 = np.ascontiguousarray(pp00, np.float64)
    dt = np.ascontiguousarray(np.broadcast_to(np.asarray(dt, np.float64), (n,)))
    pp = pp00.copy()
    uu = np.zeros_like(pp00)
    ok = np.zeros(n, np.int32)
    it = np.zeros(n, np.int32)
    lib().kzo_solve_4d_prim_batch(C.c_int64(n), T.ptr(pp00), T.ptr(_geoms(geom)), T.ptr(dt), C.c_double(gam),
                                  C.byref(rp), C.byref(op), C.byref(ip), C.c_int32(wp), C.c_int32(we),
                                  C.c_int32(wf), T.ptr(pp), T.ptr(uu), T.ptr(ok, C.c_int32), T.ptr(it, C.c_int32))
    return ok, it, pp, uu


def step_function(x, x9):
    return lib().kzo_step_function(x, x9)


def visc_coeff(chi, mindx, r_bl, rhor, alpha, global_dt):
    return lib().kzo_visc_coeff(chi, mindx, r_bl, rhor, alpha, global_dt)


def shear_from_gradients(du, du2, ucon, ucov, geom, kr):
    """radvisc.zig shearFromGradients over one point; geom a kz_geom row."""
    out = np.zeros(16)
    a = [np.ascontiguousarray(x, np.float64).ravel() for x in (du, du2, ucon, ucov, geom, kr)]
    lib().kzo_shear_from_gradients(*[T.ptr(x) for x in a], T.ptr(out))
    return out.reshape(4, 4)


def dynamo_delta_a(dp, a, dt, rhor, risco, r, th, bphi, bsq, angle, prermhd, gg33, scaleth):
    """dynamo.zig dynamoDeltaA: returns (aphi, new bphi)."""
    b = np.zeros(1)
    aphi = lib().kzo_dynamo_delta_a(C.byref(dp), a, dt, rhor, risco, r, th, bphi, bsq, angle, prermhd, gg33,
                                    scaleth, T.ptr(b))
    return aphi, b[0]


def r_horizon_bl(a):
    return lib().kzo_r_horizon_bl(a)


def r_isco_bl(a):
    return lib().kzo_r_isco_bl(a)


def damp_bphi(alphabeta, facradius, faczh, dt_over_pk, beta, betasat, bphi):
    return lib().kzo_damp_bphi(alphabeta, facradius, faczh, dt_over_pk, beta, betasat, bphi)
