"""kz_implicit_solve_noalloc: the implicit batch solve over caller device
arrays with its working buffers in the library's static device scratch.

Checked against kz_implicit_batch_solve (the allocating entry point) on the
same inputs, bit for bit: two calls in a row, the second with fewer cells,
so the second call runs on scratch the first left behind.  Sizes above the
scratch's limit are rejected before any device work.  Device memory for the
inputs and outputs is the caller's (torch tensors here).
"""
import ctypes as C

import numpy as np
import pytest
import torch

import kzo
from koral_zig_amd import abi_types as T
from koral_zig_amd import hip, workloads

pytestmark = pytest.mark.gpu

MAX_CELLS = 32768  # KZ_NOALLOC_MAX_CELLS in include/kz_implicit.h


def _inputs(n, seed, phys):
    return workloads.implicit_batch_inputs(
        n, seed, phys, lambda p, g: kzo.opac_entry(p, g, phys.gam, phys.opac)[:, 3])


def _reference(pp, geom, dt, phys):
    b = hip.ImplicitBatch(len(pp), phys)
    b.set(pp, geom, dt)
    b.solve()
    out = b.get()
    b.close()
    return out


def _noalloc(pp, geom, dt, phys):
    n = len(pp)
    dev = torch.device("cuda", 0)
    t_pp = torch.from_numpy(np.ascontiguousarray(pp)).to(dev)
    t_g = torch.from_numpy(np.ascontiguousarray(geom, np.float64)).to(dev)
    t_dt = torch.from_numpy(np.ascontiguousarray(dt)).to(dev)
    t_uu = torch.full((n, T.NV), 7.0, dtype=torch.float64, device=dev)
    t_po = torch.full((n, T.NV), 7.0, dtype=torch.float64, device=dev)
    t_r = torch.full((n * T.IMPLICIT_RESULT_DTYPE.itemsize // 4,), 7, dtype=torch.int32, device=dev)
    t_c = torch.full((3,), 7, dtype=torch.int64, device=dev)
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream(dev).cuda_stream
    rc = hip.lib().kz_implicit_solve_noalloc(n, t_pp.data_ptr(), t_g.data_ptr(), t_dt.data_ptr(), t_uu.data_ptr(),
                                             t_po.data_ptr(), t_r.data_ptr(), t_c.data_ptr(), C.byref(phys),
                                             C.c_void_p(stream))
    assert rc == 0, T.ERRORS.get(rc, rc)
    torch.cuda.synchronize()
    r = np.frombuffer(t_r.cpu().numpy().tobytes(), T.IMPLICIT_RESULT_DTYPE)
    return (t_uu.cpu().numpy(), t_po.cpu().numpy(), {k: r[k] for k in r.dtype.names},
            tuple(int(c) for c in t_c.cpu().numpy()))


def test_noalloc_matches_batch_twice():
    phys = T.physics_puffy()
    for n, seed in ((20000, 11), (5000, 12)):
        print(f"no-alloc solve of {n} cells, seed {seed}", flush=True)
        pp, geom, dt = _inputs(n, seed, phys)
        uu_a, pp_a, r_a, c_a = _reference(pp, geom, dt, phys)
        uu_b, pp_b, r_b, c_b = _noalloc(pp, geom, dt, phys)
        for k in ("ok", "rung", "iters", "whichprim", "whicheq", "whichframe"):
            assert np.array_equal(r_a[k], r_b[k]), k
        assert pp_a.tobytes() == pp_b.tobytes()
        assert uu_a.tobytes() == uu_b.tobytes()
        assert c_a == c_b and c_a[2] > 0.9 * n


def test_noalloc_bridge_matches_batch():
    """The synchrotron bridge on: the no-alloc entry point launches the
    bridge variants of the implicit kernels, as the batch does."""
    phys = T.physics_puffy()
    phys.opac.synchrotron_bridge = 1
    print("no-alloc solve of 8000 cells with the synchrotron bridge", flush=True)
    pp, geom, dt = _inputs(8000, 13, phys)
    uu_a, pp_a, r_a, c_a = _reference(pp, geom, dt, phys)
    uu_b, pp_b, r_b, c_b = _noalloc(pp, geom, dt, phys)
    for k in ("ok", "rung", "iters", "whichprim", "whicheq", "whichframe"):
        assert np.array_equal(r_a[k], r_b[k]), k
    assert pp_a.tobytes() == pp_b.tobytes()
    assert uu_a.tobytes() == uu_b.tobytes()
    assert c_a == c_b and c_a[2] > 0.9 * 8000


def test_noalloc_rejects_oversize():
    phys = T.physics_puffy()
    rc = hip.lib().kz_implicit_solve_noalloc(MAX_CELLS + 1, None, None, None, None, None, None, None,
                                             C.byref(phys), None)
    assert T.ERRORS.get(rc) == "InvalidConfig" or rc == -5, rc
