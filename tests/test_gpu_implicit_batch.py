"""The device-resident implicit batch (kz_implicit.h) vs the oracle.

Cells from the independent-cell workload generator (koral_zig_amd.workloads:
rho over 22 decades, T_gas over 6, E_hat/u over 12, optical depth over 8,
MKS2 a=0 geometry), solved by the split, compacted device solve and by the
oracle's solveImplicitLab (implicit.zig:927-1009): outcome, rung, primitive
block, equation, frame and iteration count identical per cell, solved
states identical bit for bit, failed cells left at their input.  Also the
per-record kernel of kz_cellops.h on the same cells, and two solves in a
row.
"""
import numpy as np
import pytest

import kzo
from koral_zig_amd import abi_types as T
from koral_zig_amd import hip, workloads

pytestmark = pytest.mark.gpu

FIELDS = ("ok", "rung", "iters", "whichprim", "whicheq", "whichframe")


def same(a, b):
    a, b = np.ascontiguousarray(a), np.ascontiguousarray(b)
    return a.shape == b.shape and a.tobytes() == b.tobytes()


@pytest.fixture(scope="module")
def cells():
    phys = T.physics_puffy()
    pp, geom, dt = workloads.implicit_batch_inputs(
        20000, 11, phys, lambda p, g: kzo.opac_entry(p, g, phys.gam, phys.opac)[:, 3])
    return phys, pp, geom, dt


def test_batch_matches_oracle(cells):
    phys, pp, geom, dt = cells
    rc, pc, uc = kzo.solve_implicit(pp, geom, dt, phys.gam, phys.rad, phys.opac, phys.implicit)
    b = hip.ImplicitBatch(len(pp), phys)
    b.set(pp, geom, dt)
    b.solve()
    ug, pg, rg, (nfail, iters, solves) = b.get()
    b.close()
    for k in FIELDS:
        bad = np.nonzero(rc[k] != rg[k])[0]
        assert bad.size == 0, (k, bad[:5], rc[k][bad[:5]], rg[k][bad[:5]])
    ok = rc["ok"] == 1
    assert 0.99 < ok.mean() < 1.0  # both outcomes occur
    assert same(pc[ok], pg[ok])
    assert same(uc[ok], ug[ok])
    assert same(pg[~ok], pp[~ok])
    assert (nfail, iters, solves) == (int((~ok).sum()), int(rc["iters"][ok].sum()), int(ok.sum()))
    assert len(set(rc["rung"][ok].tolist())) >= 3  # the slow path runs


def test_batch_matches_per_record_kernel(cells):
    phys, pp, geom, dt = cells
    m = 4000
    uz, pz, rz = hip.solve_implicit(np.zeros((m, T.NV)), pp[:m], geom[:m], dt[:m], phys.gam, phys.rad, phys.opac,
                                    phys.implicit)
    b = hip.ImplicitBatch(m, phys)
    b.set(pp[:m], geom[:m], dt[:m])
    b.solve()
    ug, pg, rg, _ = b.get()
    b.close()
    for k in FIELDS:
        assert same(rz[k], rg[k]), k
    ok = rz["ok"] == 1
    assert same(pz, pg) and same(uz[ok], ug[ok])


def test_batch_repeat_deterministic(cells):
    phys, pp, geom, dt = cells
    m = 6000
    b = hip.ImplicitBatch(m, phys)
    b.set(pp[:m], geom[:m], dt[:m])
    b.solve()
    u1, p1, r1, c1 = b.get()
    b.solve()
    u2, p2, r2, c2 = b.get()
    b.close()
    assert c1 == c2 and same(p1, p2) and same(u1, u2)
    for k in FIELDS:
        assert same(r1[k], r2[k])


def test_fp32_jacobian_matches_oracle(cells):
    """The mixed-precision Newton (jacobian_fp32: FP32 quotients, scaling and
    LU; FP64 residuals, inversions and update) on the device equals the
    oracle's restatement of it bit for bit; it is a different solve from the
    all-FP64 one, which it must still agree with on most outcomes."""
    phys, pp, geom, dt = cells
    mp = T.physics_puffy()
    mp.implicit.jacobian_fp32 = 1
    rc, pc, uc = kzo.solve_implicit(pp, geom, dt, mp.gam, mp.rad, mp.opac, mp.implicit)
    b = hip.ImplicitBatch(len(pp), mp)
    b.set(pp, geom, dt)
    b.solve()
    ug, pg, rg, (nfail, iters, solves) = b.get()
    b.close()
    for k in FIELDS:
        bad = np.nonzero(rc[k] != rg[k])[0]
        assert bad.size == 0, (k, bad[:5], rc[k][bad[:5]], rg[k][bad[:5]])
    ok = rc["ok"] == 1
    assert same(pc[ok], pg[ok]) and same(uc[ok], ug[ok])
    r64, _, _ = kzo.solve_implicit(pp, geom, dt, phys.gam, phys.rad, phys.opac, phys.implicit)
    assert (r64["ok"] == rc["ok"]).mean() > 0.99
    assert (r64["rung"] == rc["rung"]).mean() > 0.95
    assert not np.array_equal(r64["iters"], rc["iters"])  # the FP32 path is really taken
