"""Binding of libkz_hip.so — the per-step hot path on the GPU.

`Sim` mirrors the reference's Sim(config.puffy) methods (koral/sim.zig:
93-1900) over the C-ABI of include/kz_step.h; the *_batch functions expose
the per-point device physics of include/kz_cellops.h.  There is no CPU
fallback: every call runs the HIP kernels, and a missing library raises.
"""
import ctypes as C
from pathlib import Path

import numpy as np

from . import abi_types as T

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "lib" / "libkz_hip.so"
_lib = None

D = C.POINTER(C.c_double)
I32 = C.POINTER(C.c_int32)
GEOM = C.POINTER(C.c_double)  # kz_geom records passed as packed doubles

# every entry point the headers declare: name -> (restype, argtypes)
_SIGS = {
    # kz_step.h
    "kz_sim_create": (C.c_void_p, [C.POINTER(T.kz_grid), C.POINTER(T.kz_physics)] + [D] * 10 + [I32]),
    "kz_sim_destroy": (None, [C.c_void_p]),
    "kz_sim_set_state": (C.c_int, [C.c_void_p, D, D]),
    "kz_sim_get_state": (C.c_int, [C.c_void_p, D, D, I32]),
    "kz_sim_p2u_domain": (C.c_int, [C.c_void_p]),
    "kz_sim_calc_divb": (C.c_int, [C.c_void_p, D]),
    "kz_sim_set_bc": (C.c_int, [C.c_void_p, C.c_int]),
    "kz_sim_postinit": (C.c_int, [C.c_void_p, C.c_double, D]),
    "kz_sim_init_timestep_guess": (C.c_int, [C.c_void_p]),
    "kz_sim_set_time": (C.c_int, [C.c_void_p, C.c_double, C.c_double, C.c_double]),
    "kz_calc_u2p": (C.c_int, [C.c_void_p]),
    "kz_op_explicit": (C.c_int, [C.c_void_p, C.c_double]),
    "kz_op_implicit": (C.c_int, [C.c_void_p, C.c_double]),
    "kz_do_correct": (C.c_int, [C.c_void_p]),
    "kz_step": (C.c_int, [C.c_void_p, D]),
    "kz_step_many": (C.c_int, [C.c_void_p, C.c_int]),
    "kz_sim_sync": (C.c_int, [C.c_void_p]),
    "kz_sim_stream": (C.c_void_p, [C.c_void_p]),
    "kz_sim_capture_steps": (C.c_int, [C.c_void_p, C.c_int]),
    "kz_sim_replay": (C.c_int, [C.c_void_p]),
    "kz_sim_query": (C.c_int, [C.c_void_p]),
    "kz_group_create": (C.c_int, [C.POINTER(C.c_void_p)]),
    "kz_group_destroy": (None, [C.c_void_p]),
    "kz_group_set_runs": (C.c_int, [C.c_void_p, C.POINTER(C.c_void_p), C.c_int]),
    "kz_group_set_implicit_launch": (C.c_int, [C.c_void_p, C.c_int]),
    "kz_group_step_many": (C.c_int, [C.c_void_p, C.c_int]),
    "kz_group_capture_steps": (C.c_int, [C.c_void_p, C.c_int]),
    "kz_group_replay": (C.c_int, [C.c_void_p]),
    "kz_group_query": (C.c_int, [C.c_void_p]),
    "kz_group_sync": (C.c_int, [C.c_void_p]),
    "kz_sim_set_implicit_launch": (C.c_int, [C.c_void_p, C.c_int]),
    "kz_sim_device_bytes": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint64)]),
    "kz_sim_scalars": (C.c_int, [C.c_void_p, C.c_double, C.c_long, C.c_long, C.c_long, C.c_void_p, C.c_void_p,
                                 C.POINTER(T.kz_scalar_row)]),
    "kz_sim_counters": (C.c_int, [C.c_void_p, C.POINTER(T.kz_counters)]),
    "kz_sim_time_kernel": (C.c_int, [C.c_void_p, C.c_int]),
    "kz_sim_kernel_time": (C.c_int, [C.c_void_p, D, C.POINTER(C.c_uint64)]),
    "kz_sim_set_bl_tables": (C.c_int, [C.c_void_p, D]),
    "kz_calc_rij_visc_total": (C.c_int, [C.c_void_p, C.c_double]),
    "kz_sim_get_rijvisc": (C.c_int, [C.c_void_p, D]),
    "kz_apply_dynamo": (C.c_int, [C.c_void_p, C.c_double]),
    "kz_kdmp_parse_header": (C.c_int, [C.c_void_p, C.c_uint64, C.POINTER(T.kz_dump_header)]),
    "kz_sim_kdmp_size": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint64)]),
    "kz_sim_write_kdmp": (C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint64, C.POINTER(C.c_uint64)]),
    "kz_sim_load_kdmp": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64, C.POINTER(T.kz_dump_header)]),
    "kz_sim_finish_init": (C.c_int, [C.c_void_p]),
    # kz_implicit.h
    "kz_implicit_batch_create": (C.c_void_p, [C.c_int64, C.POINTER(T.kz_physics), I32]),
    "kz_implicit_batch_destroy": (None, [C.c_void_p]),
    "kz_implicit_batch_set": (C.c_int, [C.c_void_p, D, GEOM, D]),
    "kz_implicit_batch_solve": (C.c_int, [C.c_void_p, D]),
    "kz_implicit_batch_get": (C.c_int, [C.c_void_p, D, D, C.c_void_p, C.POINTER(C.c_uint64)]),
    "kz_implicit_solve_noalloc": (C.c_int, [C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                            C.c_void_p, C.c_void_p, C.POINTER(T.kz_physics), C.c_void_p]),
    # kz_cellops.h
    "kz_p2u_batch": (C.c_int, [C.c_int64, D, GEOM, C.c_double, D]),
    "kz_u2p_mhd_batch": (C.c_int, [C.c_int64, D, D, GEOM, C.c_double, C.POINTER(T.kz_floor_params), C.c_void_p]),
    "kz_check_floors_mhd_batch": (C.c_int, [C.c_int64, D, GEOM, C.c_double, C.POINTER(T.kz_floor_params), I32,
                                            C.POINTER(C.c_uint64)]),
    "kz_u2p_rad_batch": (C.c_int, [C.c_int64, D, D, GEOM, C.POINTER(T.kz_rad_params), C.c_void_p]),
    "kz_check_floors_rad_batch": (C.c_int, [C.c_int64, D, GEOM, C.POINTER(T.kz_rad_params), I32]),
    "kz_solve_implicit_lab_batch": (C.c_int, [C.c_int64, D, D, GEOM, D, C.c_double, C.POINTER(T.kz_rad_params),
                                              C.POINTER(T.kz_opac_params), C.POINTER(T.kz_implicit_params),
                                              C.c_void_p]),
    "kz_reconstruct_batch": (C.c_int, [C.c_int64, I32, D, D, D, D, D]),
    "kz_u2p_solver_w_batch": (C.c_int, [C.c_int64, D, D, GEOM, C.c_int32, C.c_double, I32]),
    "kz_flux_prime_batch": (C.c_int, [C.c_int64, D, I32, GEOM, C.c_double, D, I32]),
    "kz_gas_wavespeeds_batch": (C.c_int, [C.c_int64, D, GEOM, C.c_double, I32, D]),
    "kz_rad_wavespeeds_batch": (C.c_int, [C.c_int64, D, GEOM, D, I32, D]),
    "kz_rij_batch": (C.c_int, [C.c_int64, D, GEOM, D]),
    "kz_radstate_batch": (C.c_int, [C.c_int64, D, GEOM, C.c_double, C.POINTER(T.kz_opac_params), C.c_int32, D]),
    "kz_math_batch": (C.c_int, [C.c_int64, C.c_int32, D, D, D]),
}


def lib():
    """Load libkz_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} missing: the HIP library must be built "
                               "(__graft_entry__.build()); there is no CPU fallback")
        L = C.CDLL(str(LIB_PATH))
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def parse_kdmp_header(data):
    h = T.kz_dump_header()
    buf = (C.c_uint8 * len(data)).from_buffer_copy(data)
    _chk(lib().kz_kdmp_parse_header(buf, len(data), C.byref(h)), "kz_kdmp_parse_header")
    return {k: getattr(h, k) for k, _ in h._fields_ if k != "_pad"}


def _chk(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what}: {T.ERRORS.get(rc, rc)}")


def _f64(a):
    return np.ascontiguousarray(a, np.float64)


def _geoms(g):
    g = _f64(g)
    assert g.shape[-1] == T.GEOM_DOUBLES, g.shape
    return g


class Sim:
    """Device-resident run state driven through the kz_step.h C-ABI.

    `tables` is the dict from koral_zig_amd.host.metric_tables(); it is
    copied to the device at construction.
    """

    def __init__(self, grid, phys, tables):
        self.grid = grid
        self.phys = phys
        err = C.c_int32(0)
        t = {k: _f64(v) for k, v in tables.items() if isinstance(v, np.ndarray)}
        self.h = lib().kz_sim_create(C.byref(grid), C.byref(phys), T.ptr(t["g"]), T.ptr(t["gcon"]),
                                     T.ptr(t["kris"]), T.ptr(t["gb0"]), T.ptr(t["gconb0"]), T.ptr(t["gb1"]),
                                     T.ptr(t["gconb1"]), T.ptr(t["gb2"]), T.ptr(t["gconb2"]), T.ptr(t["bcx"]),
                                     C.byref(err))
        if not self.h:
            raise RuntimeError(f"kz_sim_create failed: {T.ERRORS.get(err.value, err.value)}")
        (sz, sy, sx), _ = T.padded_shape(grid)
        self.shape = (sz, sy, sx)
        self.ncell = sz * sy * sx

    def close(self):
        if getattr(self, "h", None):
            lib().kz_sim_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def set_state(self, p=None, u=None):
        p = None if p is None else _f64(p)
        u = None if u is None else _f64(u)
        _chk(lib().kz_sim_set_state(self.h, T.ptr(p), T.ptr(u)), "kz_sim_set_state")

    def get_state(self):
        p = np.zeros(self.shape + (T.NV,))
        u = np.zeros(self.shape + (T.NV,))
        f = np.zeros(self.shape + (T.NFLAGS,), np.int32)
        _chk(lib().kz_sim_get_state(self.h, T.ptr(p), T.ptr(u), T.ptr(f, C.c_int32)), "kz_sim_get_state")
        return p, u, f

    def counters(self):
        c = T.kz_counters()
        _chk(lib().kz_sim_counters(self.h, C.byref(c)), "kz_sim_counters")
        return {n: getattr(c, n) for n, _ in T.kz_counters._fields_}

    def set_time(self, t, tstepdenmax, tstepdenmin):
        _chk(lib().kz_sim_set_time(self.h, t, tstepdenmax, tstepdenmin), "kz_sim_set_time")

    def p2u_domain(self):
        _chk(lib().kz_sim_p2u_domain(self.h), "kz_sim_p2u_domain")

    def calc_divb(self):
        """ct.zig:336-381 calcDivB at every domain cell, [nz][ny][nx]."""
        g = self.grid
        out = np.zeros((g.nz, g.ny, g.nx))
        _chk(lib().kz_sim_calc_divb(self.h, T.ptr(out)), "kz_sim_calc_divb")
        return out

    def set_bc(self, ifinit=False):
        _chk(lib().kz_sim_set_bc(self.h, int(ifinit)), "kz_sim_set_bc")

    def calc_u2p(self):
        _chk(lib().kz_calc_u2p(self.h), "kz_calc_u2p")

    def op_explicit(self, dt):
        _chk(lib().kz_op_explicit(self.h, dt), "kz_op_explicit")

    def op_implicit(self, dt):
        _chk(lib().kz_op_implicit(self.h, dt), "kz_op_implicit")

    def do_correct(self):
        _chk(lib().kz_do_correct(self.h), "kz_do_correct")

    def init_timestep_guess(self):
        _chk(lib().kz_sim_init_timestep_guess(self.h), "kz_sim_init_timestep_guess")

    def postinit(self, maxbeta):
        fac = C.c_double(0.0)
        _chk(lib().kz_sim_postinit(self.h, maxbeta, C.byref(fac)), "kz_sim_postinit")
        return fac.value

    def step(self, dt=None):
        if dt is None:
            _chk(lib().kz_step(self.h, None), "kz_step")
        else:
            d = C.c_double(dt)
            _chk(lib().kz_step(self.h, C.pointer(d)), "kz_step")

    def step_many(self, nsteps):
        """Enqueue nsteps CFL steps; call sync() to wait."""
        _chk(lib().kz_step_many(self.h, int(nsteps)), "kz_step_many")

    def sync(self):
        _chk(lib().kz_sim_sync(self.h), "kz_sim_sync")

    def stream(self):
        return lib().kz_sim_stream(self.h)

    def time_kernel(self, kernel):
        """Start event-timing one kernel class (KERNELS name, or None to stop)."""
        k = -1 if kernel is None else KERNELS[kernel]
        _chk(lib().kz_sim_time_kernel(self.h, k), "kz_sim_time_kernel")

    def set_bl_tables(self, bl):
        """Load the BL sidecar (host.bl_tables) for radiative viscosity / dynamo."""
        bl = _f64(bl)
        _chk(lib().kz_sim_set_bl_tables(self.h, T.ptr(bl)), "kz_sim_set_bl_tables")

    def calc_rij_visc_total(self, global_dt):
        """Viscous R^i_j for step size global_dt; returns it, 16 per padded cell."""
        _chk(lib().kz_calc_rij_visc_total(self.h, global_dt), "kz_calc_rij_visc_total")
        r = np.zeros(self.shape + (16,))
        _chk(lib().kz_sim_get_rijvisc(self.h, T.ptr(r)), "kz_sim_get_rijvisc")
        return r

    def apply_dynamo(self, dt):
        _chk(lib().kz_apply_dynamo(self.h, dt), "kz_apply_dynamo")

    def capture_steps(self, nsteps):
        """Capture nsteps CFL steps into a HIP graph (see replay)."""
        _chk(lib().kz_sim_capture_steps(self.h, int(nsteps)), "kz_sim_capture_steps")

    def replay(self):
        """Enqueue the captured steps without waiting (sync() waits)."""
        _chk(lib().kz_sim_replay(self.h), "kz_sim_replay")

    def set_implicit_launch(self, mode):
        """0 by size, 1 merged first-rung launch, 2 persistent work stealing."""
        _chk(lib().kz_sim_set_implicit_launch(self.h, int(mode)), "kz_sim_set_implicit_launch")

    def busy(self):
        """True while enqueued work is still running (non-blocking)."""
        rc = lib().kz_sim_query(self.h)
        if rc < 0:
            _chk(rc, "kz_sim_query")
        return rc == 1

    def scalars(self, dt, ix_h, ix_l, ix_s, dth_bl, dph_bl):
        """One scalars.dat row of the current state (kz_sim_scalars) as a dict
        in the row's column order; koral_zig_amd.scalars drives it."""
        dth, dph = _f64(dth_bl), _f64(dph_bl)
        r = T.kz_scalar_row()
        _chk(lib().kz_sim_scalars(self.h, float(dt), int(ix_h), int(ix_l), int(ix_s), dth.ctypes.data,
                                  dph.ctypes.data, C.byref(r)), "kz_sim_scalars")
        return {n: getattr(r, n) for n, _ in T.kz_scalar_row._fields_}

    def device_bytes(self):
        n = C.c_uint64(0)
        _chk(lib().kz_sim_device_bytes(self.h, C.byref(n)), "kz_sim_device_bytes")
        return n.value

    def write_kdmp(self, out_idx=0):
        """KDMP v2 checkpoint of the run's domain primitives and clock (bytes)."""
        n = C.c_uint64(0)
        _chk(lib().kz_sim_kdmp_size(self.h, C.byref(n)), "kz_sim_kdmp_size")
        buf = (C.c_uint8 * n.value)()
        w = C.c_uint64(0)
        _chk(lib().kz_sim_write_kdmp(self.h, out_idx, buf, n.value, C.byref(w)), "kz_sim_write_kdmp")
        return bytes(buf[:w.value])

    def load_kdmp(self, data, finish=True):
        """Restart from a KDMP checkpoint (then ghosts + dt guess unless
        finish=False); returns the header as a dict."""
        buf = (C.c_uint8 * len(data)).from_buffer_copy(data)
        h = T.kz_dump_header()
        _chk(lib().kz_sim_load_kdmp(self.h, buf, len(data), C.byref(h)), "kz_sim_load_kdmp")
        if finish:
            _chk(lib().kz_sim_finish_init(self.h), "kz_sim_finish_init")
        return {k: getattr(h, k) for k, _ in h._fields_ if k != "_pad"}

    def kernel_time(self):
        """(summed ms, launches) of the timed class since the last call."""
        ms, n = C.c_double(0.0), C.c_uint64(0)
        _chk(lib().kz_sim_kernel_time(self.h, C.byref(ms), C.byref(n)), "kz_sim_kernel_time")
        return ms.value, n.value


class Group:
    """Resident runs stepped together (include/kz_step.h kz_group_*): both
    implicit passes of every step solve the cells of all member runs in one
    set of launches; each run's other kernels stay on its own stream.  The
    members end every step bit for bit where Sim.step_many leaves them."""

    def __init__(self, sims=()):
        h = C.c_void_p()
        _chk(lib().kz_group_create(C.byref(h)), "kz_group_create")
        self.h = h
        self.sims = []
        if sims:
            self.set_runs(sims)

    def set_runs(self, sims):
        """New membership (waits for the previous members' work)."""
        sims = list(sims)
        arr = (C.c_void_p * max(len(sims), 1))(*[s.h for s in sims])
        _chk(lib().kz_group_set_runs(self.h, arr, len(sims)), "kz_group_set_runs")
        self.sims = sims

    def set_implicit_launch(self, mode):
        _chk(lib().kz_group_set_implicit_launch(self.h, int(mode)), "kz_group_set_implicit_launch")

    def step_many(self, nsteps):
        _chk(lib().kz_group_step_many(self.h, int(nsteps)), "kz_group_step_many")

    def capture_steps(self, nsteps):
        _chk(lib().kz_group_capture_steps(self.h, int(nsteps)), "kz_group_capture_steps")

    def replay(self):
        _chk(lib().kz_group_replay(self.h), "kz_group_replay")

    def busy(self):
        rc = lib().kz_group_query(self.h)
        if rc < 0:
            _chk(rc, "kz_group_query")
        return rc == 1

    def sync(self):
        _chk(lib().kz_group_sync(self.h), "kz_group_sync")

    def close(self):
        if self.h:
            lib().kz_group_destroy(self.h)
            self.h = None
            self.sims = []

    def __del__(self):
        if _lib is not None and getattr(self, "h", None):
            self.close()


class ImplicitBatch:
    """n independent implicit solves resident on the device (kz_implicit.h)."""

    def __init__(self, n, phys):
        self.n = int(n)
        err = C.c_int32(0)
        self.h = lib().kz_implicit_batch_create(self.n, C.byref(phys), C.byref(err))
        if not self.h:
            raise RuntimeError(f"kz_implicit_batch_create failed: {T.ERRORS.get(err.value, err.value)}")

    def close(self):
        if getattr(self, "h", None):
            lib().kz_implicit_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def set(self, pp, geom, dt):
        pp, dt = _f64(pp), _f64(dt)
        assert pp.shape == (self.n, T.NV) and dt.shape == (self.n,)
        _chk(lib().kz_implicit_batch_set(self.h, T.ptr(pp), T.ptr(_geoms(geom)), T.ptr(dt)), "kz_implicit_batch_set")

    def solve(self):
        """Solve all cells from the loaded inputs; returns the kernel time in ms."""
        ms = C.c_double(0.0)
        _chk(lib().kz_implicit_batch_solve(self.h, C.byref(ms)), "kz_implicit_batch_solve")
        return ms.value

    def get(self):
        """(uu, pp, result dict, counts (failures, iterations, solves))."""
        uu = np.zeros((self.n, T.NV))
        pp = np.zeros((self.n, T.NV))
        r = np.zeros(self.n, T.IMPLICIT_RESULT_DTYPE)
        cnt = (C.c_uint64 * 3)()
        _chk(lib().kz_implicit_batch_get(self.h, T.ptr(uu), T.ptr(pp), r.ctypes.data_as(C.c_void_p), cnt),
             "kz_implicit_batch_get")
        return uu, pp, {k: r[k] for k in r.dtype.names}, tuple(int(c) for c in cnt)


KERNELS = {"wavespeeds": 0, "face_flux": 1, "ct": 2, "update": 3, "u2p": 4, "implicit": 5, "bc": 6, "visc": 7,
           "dynamo": 8}


# ---- per-point batches (include/kz_cellops.h) ------------------------------

def p2u(pp, geom, gam):
    pp = _f64(pp)
    uu = np.zeros_like(pp)
    _chk(lib().kz_p2u_batch(len(pp), T.ptr(pp), T.ptr(_geoms(geom)), gam, T.ptr(uu)), "kz_p2u_batch")
    return uu


def u2p_mhd(uu, pp, geom, gam, floors):
    uu, pp = _f64(uu), _f64(pp).copy()
    r = np.zeros(len(uu), T.U2P_RESULT_DTYPE)
    _chk(lib().kz_u2p_mhd_batch(len(uu), T.ptr(uu), T.ptr(pp), T.ptr(_geoms(geom)), gam, C.byref(floors),
                                r.ctypes.data), "kz_u2p_mhd_batch")
    return pp, r


def check_floors_mhd(pp, geom, gam, floors):
    pp = _f64(pp).copy()
    ret = np.zeros(len(pp), np.int32)
    ng = C.c_uint64(0)
    _chk(lib().kz_check_floors_mhd_batch(len(pp), T.ptr(pp), T.ptr(_geoms(geom)), gam, C.byref(floors),
                                         T.ptr(ret, C.c_int32), C.byref(ng)), "kz_check_floors_mhd_batch")
    return pp, ret, ng.value


def u2p_rad(uu, pp, geom, rad):
    uu, pp = _f64(uu), _f64(pp).copy()
    r = np.zeros(len(uu), T.U2PRAD_RESULT_DTYPE)
    _chk(lib().kz_u2p_rad_batch(len(uu), T.ptr(uu), T.ptr(pp), T.ptr(_geoms(geom)), C.byref(rad), r.ctypes.data),
         "kz_u2p_rad_batch")
    return pp, r


def check_floors_rad(pp, geom, rad):
    pp = _f64(pp).copy()
    ret = np.zeros(len(pp), np.int32)
    _chk(lib().kz_check_floors_rad_batch(len(pp), T.ptr(pp), T.ptr(_geoms(geom)), C.byref(rad),
                                         T.ptr(ret, C.c_int32)), "kz_check_floors_rad_batch")
    return pp, ret


def solve_implicit(uu, pp, geom, dt, gam, rad, opac, imp):
    uu, pp = _f64(uu).copy(), _f64(pp).copy()
    n = len(uu)
    dt = _f64(np.broadcast_to(np.asarray(dt, np.float64), (n,)))
    r = np.zeros(n, T.IMPLICIT_RESULT_DTYPE)
    _chk(lib().kz_solve_implicit_lab_batch(n, T.ptr(uu), T.ptr(pp), T.ptr(_geoms(geom)), T.ptr(dt), gam,
                                           C.byref(rad), C.byref(opac), C.byref(imp), r.ctypes.data),
         "kz_solve_implicit_lab_batch")
    return uu, pp, r


def reconstruct(order, theta, u5, dx5):
    u5, dx5 = _f64(u5), _f64(dx5)
    n = len(u5)
    order = np.ascontiguousarray(np.broadcast_to(np.asarray(order, np.int32), (n,)))
    theta = _f64(np.broadcast_to(np.asarray(theta, np.float64), (n,)))
    l, r = np.zeros(n), np.zeros(n)
    _chk(lib().kz_reconstruct_batch(n, T.ptr(order, C.c_int32), T.ptr(theta), T.ptr(u5), T.ptr(dx5), T.ptr(l),
                                    T.ptr(r)), "kz_reconstruct_batch")
    return l, r


def u2p_solver_w(uu, guess, geom, etype, gam):
    uu, pp = _f64(uu), _f64(guess).copy()
    ret = np.zeros(len(uu), np.int32)
    _chk(lib().kz_u2p_solver_w_batch(len(uu), T.ptr(uu), T.ptr(pp), T.ptr(_geoms(geom)), int(etype), gam,
                                     T.ptr(ret, C.c_int32)), "kz_u2p_solver_w_batch")
    return ret, pp


def flux_prime(pp, idim, geom, gam):
    pp = _f64(pp)
    n = len(pp)
    idim = np.ascontiguousarray(np.broadcast_to(np.asarray(idim, np.int32), (n,)))
    ff = np.zeros_like(pp)
    err = np.zeros(n, np.int32)
    _chk(lib().kz_flux_prime_batch(n, T.ptr(pp), T.ptr(idim, C.c_int32), T.ptr(_geoms(geom)), gam, T.ptr(ff),
                                   T.ptr(err, C.c_int32)), "kz_flux_prime_batch")
    return ff, err


def gas_wavespeeds(pp, geom, gam, active=(1, 1, 1)):
    pp = _f64(pp)
    act = np.asarray(active, np.int32)
    out = np.zeros((len(pp), 6))
    _chk(lib().kz_gas_wavespeeds_batch(len(pp), T.ptr(pp), T.ptr(_geoms(geom)), gam, T.ptr(act, C.c_int32),
                                       T.ptr(out)), "kz_gas_wavespeeds_batch")
    return out


def rad_wavespeeds(pp, geom, tau3, active=(1, 1, 1)):
    pp, tau3 = _f64(pp), _f64(tau3)
    act = np.asarray(active, np.int32)
    out = np.zeros((len(pp), 12))
    _chk(lib().kz_rad_wavespeeds_batch(len(pp), T.ptr(pp), T.ptr(_geoms(geom)), T.ptr(tau3),
                                       T.ptr(act, C.c_int32), T.ptr(out)), "kz_rad_wavespeeds_batch")
    return out


def rij(pp, geom):
    pp = _f64(pp)
    out = np.zeros((len(pp), 21))
    _chk(lib().kz_rij_batch(len(pp), T.ptr(pp), T.ptr(_geoms(geom)), T.ptr(out)), "kz_rij_batch")
    return out


def radstate(pp, geom, gam, opac, full=True):
    pp = _f64(pp)
    out = np.zeros((len(pp), 10))
    _chk(lib().kz_radstate_batch(len(pp), T.ptr(pp), T.ptr(_geoms(geom)), gam, C.byref(opac), int(full),
                                 T.ptr(out)), "kz_radstate_batch")
    return out


MATH_FN = {"exp": 0, "log": 1, "pow": 2, "cbrt": 3, "log1p": 4, "sqrt": 5, "div": 6}


def math_fn(name, x, y=None):
    x = _f64(x)
    yy = None if y is None else _f64(np.broadcast_to(np.asarray(y, np.float64), x.shape))
    out = np.zeros_like(x)
    _chk(lib().kz_math_batch(x.size, MATH_FN[name], T.ptr(x), T.ptr(yy), T.ptr(out)), "kz_math_batch")
    return out
