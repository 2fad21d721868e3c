"""Pin the CPU restatement (oracle/) against the reference's committed
C-oracle golden records, at the reference's own gates.

Each test mirrors one reference golden test (file:line in its docstring).
"""
import numpy as np
import pytest

import kzo
from goldens import DevTracker, geoms, read_kgld
from koral_zig_amd import abi_types as T

GAM = 5.0 / 3.0


@pytest.fixture(scope="module")
def P():
    return dict(floors=T.floor_params_puffy(), rad=T.rad_params_puffy(), opac=T.opac_params_puffy(),
                imp=T.implicit_params_puffy())


def test_recon_golden():
    """flux_golden_tests.zig:18-41 — avg2point (donor/minmod/PPM), 1e-14."""
    inp, out = read_kgld("flux/recon.kgld", 12, 2)
    order = np.maximum(0, 2 - inp[:, 10].astype(np.int32))
    l, r = kzo.reconstruct(order, inp[:, 11], inp[:, 0:5], inp[:, 5:10])
    t = DevTracker().add(out[:, 0], l).add(out[:, 1], r)
    assert t.ok(1e-14), t.max_dev


def test_gas_wavespeeds_golden():
    """flux_golden_tests.zig:43-59 — gas wavespeeds l/r (x,y), 1e-13."""
    inp, out = read_kgld("flux/wavespeed.kgld", 36, 4)
    a = kzo.gas_wavespeeds(inp[:, 23:36], geoms(inp), GAM, (1, 1, 0))
    t = DevTracker().add(out, a[:, :4])
    assert t.ok(1e-13), t.max_dev


def _family_scaled(t, c, z):
    smhd = np.maximum(1e-300, np.max(np.abs(c[:, :9]), axis=1, keepdims=True))
    srad = np.maximum(1e-300, np.max(np.abs(c[:, 9:]), axis=1, keepdims=True))
    t.add(c[:, :9] / smhd, z[:, :9] / smhd)
    t.add(c[:, 9:] / srad, z[:, 9:] / srad)
    return t


def test_flux_prime_golden():
    """flux_golden_tests.zig:61-87 — f_flux_prime all 13 rows, family-normalized 1e-13."""
    inp, out = read_kgld("flux/fluxprime.kgld", 37, 13)
    ff, err = kzo.flux_prime(inp[:, 24:37], inp[:, 23].astype(np.int32), geoms(inp), GAM)
    assert (err == 0).all()
    t = _family_scaled(DevTracker(), out, ff)
    assert t.ok(1e-13), t.max_dev


def test_p2u_golden():
    """state_golden_tests.zig:161-189 — p2u family-normalized 1e-13."""
    inp, out = read_kgld("state/p2u.kgld", 36, 13)
    uu = kzo.p2u(inp[:, 23:36], geoms(inp), GAM)
    t = _family_scaled(DevTracker(), out, uu)
    assert t.ok(1e-13), t.max_dev


def test_tij_golden():
    """state_golden_tests.zig:114-159 — T^munu, u, b, 1e-13."""
    inp, out = read_kgld("state/physics_tij.kgld", 33, 33)
    o = kzo.tij(inp[:, 20:33], geoms(inp, 20), GAM)
    scale = np.maximum(1e-300, np.max(np.abs(out[:, :16]), axis=1, keepdims=True))
    t = DevTracker().add(out[:, :16] / scale, o[:, :16] / scale).add(out[:, 16:], o[:, 16:])
    assert t.ok(1e-13), t.max_dev


def test_u2p_solver_golden():
    """state_golden_tests.zig:191-236 — u2p_solver_W hot + entropy: outcome exact, prims 1e-8."""
    inp, out = read_kgld("state/u2p_solver.kgld", 49, 20)
    g = geoms(inp)
    uu, guess = inp[:, 23:36], inp[:, 36:49]
    rh, ph = kzo.u2p_solver_w(uu, guess, g, 0, GAM)
    re, pe = kzo.u2p_solver_w(uu, guess, g, 1, GAM)
    c_hot, c_ent = out[:, 0].astype(int), out[:, 10].astype(int)
    assert ((c_hot == 0) == (rh == 0)).all()
    assert ((c_ent == 0) == (re == 0)).all()
    th = DevTracker()
    te = DevTracker()
    for i in range(len(inp)):
        if c_hot[i] == 0 and rh[i] == 0:
            th.add(out[i, 1:10], ph[i, :9], i)
        if c_ent[i] == 0 and re[i] == 0:
            te.add(out[i, 11:20], pe[i, :9], i)
    assert th.ok(1e-8), th.max_dev
    assert te.ok(1e-8), te.max_dev


def test_floors_mhd_golden(P):
    """state_golden_tests.zig:238-260 — check_floors_mhd, ret exact, 1e-12."""
    inp, out = read_kgld("state/floors.kgld", 36, 14)
    ret, pp, _ = kzo.check_floors_mhd(inp[:, 23:36], geoms(inp), GAM, P["floors"])
    assert (ret == out[:, 0].astype(int)).all()
    t = DevTracker().add(out[:, 1:], pp)
    assert t.ok(1e-12), t.max_dev


def test_rij_golden():
    """rad_golden_tests.zig:27-59 — R^munu 1e-14, Rtt 1e-13, ucon 1e-14."""
    inp, out = read_kgld("rad/rad_rij.kgld", 33, 21)
    o = kzo.rij(inp[:, 20:33], geoms(inp, 20))
    scale = np.maximum(1e-300, np.max(np.abs(out[:, :16]), axis=1, keepdims=True))
    assert DevTracker().add(out[:, :16] / scale, o[:, :16] / scale).ok(1e-14)
    assert DevTracker().add(out[:, 16:17] / scale, o[:, 16:17] / scale).ok(1e-13)
    assert DevTracker().add(out[:, 17:], o[:, 17:]).ok(1e-14)


def test_u2prad_golden(P):
    """rad_golden_tests.zig:61-100 — u2p_rad: ret and cold branch exact, 1e-12."""
    inp, out = read_kgld("rad/rad_u2prad.kgld", 31, 6)
    n = len(inp)
    uu = np.zeros((n, 13))
    uu[:, 9:13] = inp[:, 23:27]
    pp = np.zeros((n, 13))
    pp[:, 9:13] = inp[:, 27:31]
    res, po = kzo.u2p_rad(uu, pp, geoms(inp), P["rad"])
    assert (res["ret"] == out[:, 0].astype(int)).all()
    assert (res["corrected"] == out[:, 1].astype(int)).all()
    assert res["corrected"].sum() > 0  # the cold branches are exercised
    scale = np.maximum(np.abs(out[:, 2]), 1e-300)
    t = DevTracker().add(out[:, 2] / scale, po[:, 9] / scale).add(out[:, 3:6], po[:, 10:13])
    assert t.ok(1e-12), t.max_dev


def test_rad_wavespeeds_golden():
    """rad_golden_tests.zig:101-127 — calc_rad_wavespeeds (tau limiter), 1e-12."""
    inp, out = read_kgld("rad/rad_wavespeeds.kgld", 36, 8)
    a = kzo.rad_wavespeeds(inp[:, 20:33], geoms(inp, 20), inp[:, 33:36], (1, 1, 0))
    z = np.concatenate([a[:, 0:4], a[:, 6:10]], axis=1)
    t = DevTracker().add(out, z)
    assert t.ok(1e-12), t.max_dev


def test_floors_rad_golden(P):
    """rad_golden_tests.zig:129-149 — check_floors_rad, ret exact, 1e-13."""
    inp, out = read_kgld("rad/rad_floors.kgld", 36, 14)
    ret, pp = kzo.check_floors_rad(inp[:, 23:36], geoms(inp), P["rad"])
    assert (ret == out[:, 0].astype(int)).all()
    scale = np.maximum(np.abs(out[:, 1:]), 1e-300)
    t = DevTracker().add(out[:, 1:] / scale, pp / scale)
    assert t.ok(1e-13), t.max_dev


def test_radstate_golden(P):
    """opac_golden_tests.zig:35-70 — fill_struct_of_state thermo 1e-13, opacities 1e-12."""
    inp, out = read_kgld("rad/rad_thermo.kgld", 36, 13)
    o = kzo.radstate(inp[:, 23:36], geoms(inp), GAM, P["opac"], 1)
    assert DevTracker().add(out[:, 0:6], o[:, 0:6]).ok(1e-13)
    scale = np.maximum(1e-300, np.max(np.abs(out[:, 6:13]), axis=1, keepdims=True))
    t = DevTracker().add(out[:, 6:13] / scale, o[:, 6:13] / scale)
    assert t.ok(1e-12), t.max_dev
    # the slim (residual-level) fill reproduces every consumed field bit-for-bit
    s = kzo.radstate(inp[:, 23:36], geoms(inp), GAM, P["opac"], 0)
    for k in (0, 1, 2, 3, 4, 5, 6, 7, 11, 12):
        assert np.array_equal(s[:, k], o[:, k]), k


def test_opac_entry_golden(P):
    """opac_golden_tests.zig:72-92 — calc_kappa/kappaes/chi 1e-12; calcChiSlim == calcChi bitwise."""
    inp, out = read_kgld("rad/rad_opac.kgld", 36, 3)
    o = kzo.opac_entry(inp[:, 23:36], geoms(inp), GAM, P["opac"])
    t = DevTracker().add(out, o[:, :3])
    assert t.ok(1e-12), t.max_dev
    assert np.array_equal(o[:, 2], o[:, 3])


def test_gi_golden(P):
    """opac_golden_tests.zig:94-130 — calc_Gi lab/ff, force-scale normalized 1e-12."""
    inp, out = read_kgld("rad/rad_gi.kgld", 36, 8)
    pp = inp[:, 23:36]
    g = geoms(inp)
    o, slim = kzo.gi(pp, g, GAM, P["opac"])
    st = kzo.radstate(pp, g, GAM, P["opac"], 1)
    op = P["opac"]
    te, ehat, kes = st[:, 1], st[:, 3], st[:, 5]
    b = op.sigma_rad_over_pi * te * te * te * te
    thetae = op.kb_over_me * te
    coeff = kes * ehat * ((4.0 * op.kb_over_me) * (2.0 * te - te)) * \
        (1.0 + 3.683 * thetae + 4.0 * thetae * thetae) / (1.0 + 4.0 * thetae)
    scale = (st[:, 11] + kes) * ehat + st[:, 6] * op.fourpi * b + np.abs(coeff)
    t = DevTracker().add(out / scale[:, None], o / scale[:, None])
    assert t.ok(1e-12), t.max_dev
    # slim Gi: lab[0..3] and ff[0] bit-identical to the full evaluation
    assert np.array_equal(slim[:, :5], o[:, :5])


def test_invert4_closed_form():
    """implicit.zig:176-224 invert4 vs numpy LU inverse on well-conditioned matrices."""
    rng = np.random.default_rng(7)
    a = rng.normal(size=(500, 4, 4)) + 4 * np.eye(4)
    inv = kzo.invert4(a.reshape(-1, 16)).reshape(-1, 4, 4)
    ref = np.linalg.inv(a)
    assert np.max(np.abs(inv - ref) / np.maximum(1, np.abs(ref))) < 1e-13
    # a pivoting case: zero leading entry
    m = np.array([[0.0, 2, 1, 0], [1, 0, 0, 3], [0, 1, 4, 0], [2, 0, 1, 1]])
    assert np.allclose(kzo.invert4(m.reshape(1, 16)).reshape(4, 4), np.linalg.inv(m), rtol=1e-14, atol=1e-15)


def test_laxf_closed_form():
    """laxf.zig:13-15 F*(U,U) = F(U) and the linear formula (flux_tests.zig:128)."""
    rng = np.random.default_rng(3)
    f, u, a = rng.normal(size=(3, 1000))
    assert np.array_equal(kzo.laxf(f, f, u, u, a), 0.5 * (f + f - a * (u - u)))
    fl, fr, ul, ur, ag = rng.normal(size=(5, 1000))
    assert np.array_equal(kzo.laxf(fl, fr, ul, ur, ag), 0.5 * (fr + fl - ag * (ur - ul)))


def test_ppm_exact_on_linear_and_bounded():
    """flux_tests.zig:46-127 — PPM reproduces a linear profile on a uniform
    grid exactly and never overshoots the stencil's neighbours."""
    rng = np.random.default_rng(11)
    a, b = rng.normal(size=(2, 200))
    x = np.arange(5)[None, :] - 2.0
    u5 = a[:, None] + b[:, None] * x
    dx = np.ones((200, 5))
    l, r = kzo.reconstruct(2, 1.5, u5, dx)
    assert np.allclose(l, a - 0.5 * b, rtol=1e-13, atol=1e-13)
    assert np.allclose(r, a + 0.5 * b, rtol=1e-13, atol=1e-13)
    u5 = rng.normal(size=(2000, 5))
    l, r = kzo.reconstruct(2, 1.5, u5, np.ones((2000, 5)) * 0.3)
    lo = np.minimum(u5[:, 1], np.minimum(u5[:, 2], u5[:, 3]))
    hi = np.maximum(u5[:, 1], np.maximum(u5[:, 2], u5[:, 3]))
    assert (l >= lo - 1e-12).all() and (l <= hi + 1e-12).all()
    assert (r >= lo - 1e-12).all() and (r <= hi + 1e-12).all()


def test_implicit_golden(P):
    """implicit_golden_tests.zig:50-168 — full rung ladder vs C: ok-agreement >= 97%,
    entropy-rung flips <= 30, rung agreement >= 90%, >= 40 residual-verified
    records compared, worst pp deviation <= 1e-6."""
    inp, out = read_kgld("rad/rad_implicit.kgld", 37, 18)
    g = geoms(inp)
    dt = inp[:, 23]
    pp0 = inp[:, 24:37]
    res, pp, uu = kzo.solve_implicit(pp0, g, dt, GAM, P["rad"], P["opac"], P["imp"])
    c_ok = out[:, 0] == 0.0
    n_ok_agree = n_basis = n_flip = n_both = n_rung = 0
    cmp_idx = []
    for i in range(len(inp)):
        z_ok = bool(res["ok"][i])
        z_ent = z_ok and res["whicheq"][i] == 1
        c_ent = c_ok[i] and int(out[i, 2]) == 1
        if c_ok[i] != z_ok and (z_ent or c_ent):
            n_flip += 1
            continue
        n_basis += 1
        n_ok_agree += int(c_ok[i] == z_ok)
        if c_ok[i] and z_ok:
            n_both += 1
            same = (int(out[i, 1]) == res["whichprim"][i] and int(out[i, 2]) == res["whicheq"][i]
                    and int(out[i, 3]) == res["whichframe"][i])
            if same:
                n_rung += 1
                cmp_idx.append(i)
    assert n_ok_agree * 100 >= n_basis * 97, (n_ok_agree, n_basis)
    assert n_flip <= 30
    assert n_rung * 10 >= n_both * 9, (n_rung, n_both)
    # residual-verified comparison (trueResidual, implicit_golden_tests.zig:25-40)
    idx = np.array(cmp_idx)
    c_pp = out[idx, 5:18]
    wp, we, wf = res["whichprim"][idx], res["whicheq"][idx], res["whicheq"][idx] * 0 + res["whichframe"][idx]
    err_z = kzo.true_residual(pp[idx], pp0[idx], dt[idx], g[idx], GAM, P["opac"], wp, we, wf)
    err_c = kzo.true_residual(c_pp, pp0[idx], dt[idx], g[idx], GAM, P["opac"], wp, we, wf)
    keep = (err_z < 1e-8) & (err_c < 1e-8)
    assert keep.sum() >= 40, keep.sum()
    vel = [2, 3, 4, 10, 11, 12]
    worst = 0.0
    for k in np.nonzero(keep)[0]:
        i = idx[k]
        cp = out[i, 5:18]
        vmax = max(1e-14, np.max(np.abs(np.concatenate([cp[vel], pp[i, vel]]))))
        for iv in range(13):
            sc = vmax if iv in vel else max(abs(cp[iv]), abs(pp[i, iv]), 1e-300)
            worst = max(worst, abs(cp[iv] - pp[i, iv]) / sc)
    assert worst <= 1e-6, worst


def test_flux_ct_golden():
    """step_golden_tests.zig:299-336 — flux-CT (Toth EMF averaging) on a 32x32
    grid: the B rows of the face fluxes set from the records, fluxCt, every
    record's rebuilt B rows compared at 1e-15."""
    from koral_zig_amd import host
    inp, out = read_kgld("flux/ct.kgld", 6, 3)
    tp = 2.0 * np.pi
    grid = T.kz_grid(nx=32, ny=32, nz=1, ng=3, minx=0.0, maxx=tp, miny=0.0, maxy=tp, minz=0.0, maxz=1.0)
    sim = kzo.Sim(grid, T.physics_puffy(), host.metric_tables(grid))
    flb = [sim.get_flb(d) for d in range(2)]
    ng = 3
    for r in inp:
        d, ix, iy = int(r[0]), int(r[1]), int(r[2])
        flb[d][0, iy + ng, ix + ng, 6:9] = r[3:6]
    for d in range(2):
        sim.set_flb(d, flb[d])
    sim.flux_ct()
    res = [sim.get_flb(d) for d in range(2)]
    got = np.array([res[int(r[0])][0, int(r[2]) + ng, int(r[1]) + ng, 6:9] for r in inp])
    t = DevTracker().add(out, got)
    assert t.ok(1e-15), t.max_dev
    # the normal B flux of every rebuilt face is zeroed (ct.zig:160,166)
    dims, ix, iy = inp[:, 0].astype(int), inp[:, 1].astype(int), inp[:, 2].astype(int)
    rx = (dims == 0) & (iy >= 0) & (iy < 32)
    ry = (dims == 1) & (ix >= 0) & (ix < 32)
    assert rx.sum() == 33 * 32 and ry.sum() == 33 * 32
    assert (got[rx, 0] == 0).all() and (got[ry, 1] == 0).all()
