/* kzo_abi.c — batched entry points of the oracle for the test harness
 * (ctypes).  Each loops a per-point restatement over n records whose
 * geometry arrives as kz_geom (koral/geometry.zig:29-49).
 * TEST INFRASTRUCTURE (see kzo.h).
 */
#include "kzo.h"

#define GEO(i) kzo_geo gm; kzo_geo_from_kz(&g[i], &gm)

int kzo_p2u_batch(int64_t n, const double *pp, const kz_geom *g, double gam, double *uu) {
  for (int64_t i = 0; i < n; ++i) { GEO(i); kzo_p2u(pp + i * NV, uu + i * NV, &gm, gam); }
  return 0;
}

int kzo_u2p_solver_w_batch(int64_t n, const double *uu, double *pp, const kz_geom *g, int32_t etype, double gam,
                           int32_t *ret) {
  for (int64_t i = 0; i < n; ++i) { GEO(i); ret[i] = kzo_u2p_solver_w(uu + i * NV, pp + i * NV, &gm, etype, gam); }
  return 0;
}

int kzo_u2p_mhd_batch(int64_t n, const double *uu, double *pp, const kz_geom *g, double gam,
                      const kz_floor_params *f, kz_u2p_result *r) {
  for (int64_t i = 0; i < n; ++i) { GEO(i); r[i] = kzo_u2p_mhd(uu + i * NV, pp + i * NV, &gm, gam, f); }
  return 0;
}

int kzo_check_floors_mhd_batch(int64_t n, double *pp, const kz_geom *g, double gam, const kz_floor_params *f,
                               int32_t *ret, uint64_t *nguard) {
  uint64_t ng = 0;
  if (f->b2rhofloorframe < 0 || f->b2rhofloorframe > 1) return KZ_E_INVALID_CONFIG;
  for (int64_t i = 0; i < n; ++i) { GEO(i); ret[i] = kzo_check_floors_mhd(pp + i * NV, &gm, gam, f, &ng); }
  if (nguard) *nguard = ng;
  return 0;
}

int kzo_u2p_rad_batch(int64_t n, const double *uu, double *pp, const kz_geom *g, const kz_rad_params *rp,
                      kz_u2prad_result *r) {
  for (int64_t i = 0; i < n; ++i) { GEO(i); r[i] = kzo_u2p_rad(uu + i * NV, pp + i * NV, &gm, rp); }
  return 0;
}

int kzo_check_floors_rad_batch(int64_t n, double *pp, const kz_geom *g, const kz_rad_params *rp, int32_t *ret) {
  for (int64_t i = 0; i < n; ++i) { GEO(i); ret[i] = kzo_check_floors_rad(pp + i * NV, &gm, rp); }
  return 0;
}

int kzo_solve_implicit_lab_batch(int64_t n, double *uu, double *pp, const kz_geom *g, const double *dt, double gam,
                                 const kz_rad_params *rp, const kz_opac_params *op, const kz_implicit_params *ip,
                                 kz_implicit_result *r) {
  /* records are independent: the oracle's thread count splits them, each
   * solved exactly as serially */
#pragma omp parallel for num_threads(kzo_get_threads()) schedule(dynamic, 256)
  for (int64_t i = 0; i < n; ++i) {
    GEO(i);
    r[i] = kzo_solve_implicit_lab(uu + i * NV, pp + i * NV, &gm, dt[i], gam, rp, op, ip);
  }
  return 0;
}

int kzo_reconstruct_batch(int64_t n, const int32_t *order, const double *theta, const double *u5, const double *dx5,
                          double *l, double *r) {
  for (int64_t i = 0; i < n; ++i) kzo_reconstruct(order[i], theta[i], u5 + i * 5, dx5 + i * 5, l + i, r + i);
  return 0;
}

int kzo_laxf_batch(int64_t n, const double *fl, const double *fr, const double *ul, const double *ur,
                   const double *ag, double *out) {
  for (int64_t i = 0; i < n; ++i) out[i] = kzo_laxf(fl[i], fr[i], ul[i], ur[i], ag[i]);
  return 0;
}

int kzo_flux_prime_batch(int64_t n, const double *pp, const int32_t *idim, const kz_geom *g, double gam, double *ff,
                         int32_t *err) {
  for (int64_t i = 0; i < n; ++i) { GEO(i); err[i] = kzo_f_flux_prime(pp + i * NV, idim[i], &gm, gam, ff + i * NV); }
  return 0;
}

int kzo_gas_wavespeeds_batch(int64_t n, const double *pp, const kz_geom *g, double gam, const int32_t *act,
                             double *a6) {
  int active[3] = {act[0], act[1], act[2]};
  for (int64_t i = 0; i < n; ++i) { GEO(i); kzo_gas_wavespeeds_lr(pp + i * NV, &gm, gam, active, a6 + i * 6); }
  return 0;
}

int kzo_rad_wavespeeds_batch(int64_t n, const double *pp, const kz_geom *g, const double *tau3, const int32_t *act,
                             double *a12) {
  int active[3] = {act[0], act[1], act[2]};
  for (int64_t i = 0; i < n; ++i) { GEO(i); kzo_calc_rad_wavespeeds(pp + i * NV, &gm, tau3 + i * 3, active, a12 + i * 12); }
  return 0;
}

/* out21: Rij16, Rtt, ucon4 */
int kzo_rij_batch(int64_t n, const double *pp, const kz_geom *g, double *out21) {
  for (int64_t i = 0; i < n; ++i) {
    GEO(i);
    double rij[4][4], ucon[4];
    kzo_calc_rij(pp + i * NV, &gm, rij);
    double rtt = kzo_calc_ff_rtt(pp + i * NV, &gm, ucon);
    double *o = out21 + i * 21;
    for (int a = 0; a < 4; ++a)
      for (int b = 0; b < 4; ++b) o[a * 4 + b] = rij[a][b];
    o[16] = rtt;
    for (int a = 0; a < 4; ++a) o[17 + a] = ucon[a];
  }
  return 0;
}

/* out33: Tij16 ucon4 ucov4 bcon4 bcov4 bsq */
int kzo_tij_batch(int64_t n, const double *pp, const kz_geom *g, double gam, double *out33) {
  for (int64_t i = 0; i < n; ++i) {
    GEO(i);
    const double *p = pp + i * NV;
    double t[4][4], ucon[4], ucov[4], bcon[4], bcov[4], bsq;
    kzo_calc_tij(p, &gm, gam, t);
    kzo_ucon_ucov_from_prims(p, &gm, ucon, ucov);
    double B[3] = {p[B1], p[B2], p[B3]};
    kzo_bcon_bcov_bsq(B, ucon, ucov, &gm, bcon, bcov, &bsq);
    double *o = out33 + i * 33;
    for (int a = 0; a < 4; ++a)
      for (int b = 0; b < 4; ++b) o[a * 4 + b] = t[a][b];
    for (int a = 0; a < 4; ++a) { o[16 + a] = ucon[a]; o[20 + a] = ucov[a]; o[24 + a] = bcon[a]; o[28 + a] = bcov[a]; }
    o[32] = bsq;
  }
  return 0;
}

/* out14: Tgas Te ne Ehat TradBB kappaes gas_abs rad_abs gas_num rad_num gas_ross rad_ross kappa sgas */
int kzo_radstate_batch(int64_t n, const double *pp, const kz_geom *g, double gam, const kz_opac_params *op,
                       int32_t full, double *out14) {
  for (int64_t i = 0; i < n; ++i) {
    GEO(i);
    kzo_radstate st;
    kzo_fill_radstate(pp + i * NV, &gm, gam, op, full, &st);
    double *o = out14 + i * 14;
    o[0] = st.tgas; o[1] = st.te; o[2] = st.ne; o[3] = st.ehat; o[4] = st.tradbb; o[5] = st.kappaes;
    o[6] = st.opac.gas_abs; o[7] = st.opac.rad_abs; o[8] = st.opac.gas_num; o[9] = st.opac.rad_num;
    o[10] = st.opac.gas_ross; o[11] = st.opac.rad_ross; o[12] = st.kappa; o[13] = st.sgas;
  }
  return 0;
}

/* out4: kappa (calc_kappa) kappaes (standalone) chi (calc_chi) chi_slim */
int kzo_opac_entry_batch(int64_t n, const double *pp, const kz_geom *g, double gam, const kz_opac_params *op,
                         double *out4) {
  for (int64_t i = 0; i < n; ++i) {
    GEO(i);
    kzo_radstate st;
    kzo_fill_radstate(pp + i * NV, &gm, gam, op, 1, &st);
    double kes = kzo_calc_kappaes_standalone(pp + i * NV, gam, op);
    double *o = out4 + i * 4;
    o[0] = st.kappa;
    o[1] = kes;
    o[2] = st.kappa + kes;
    o[3] = kzo_calc_chi_slim(pp + i * NV, &gm, gam, op);
  }
  return 0;
}

/* out8: Gi_lab4 Gi_ff4 (full boost), plus the slim ff0/lab for identity checks in out8b */
int kzo_gi_batch(int64_t n, const double *pp, const kz_geom *g, double gam, const kz_opac_params *op,
                 double *out8, double *out8_slim) {
  for (int64_t i = 0; i < n; ++i) {
    GEO(i);
    kzo_radstate st;
    kzo_fill_radstate(pp + i * NV, &gm, gam, op, 1, &st);
    double ff[4], lab[4];
    kzo_calc_gi(&st, pp + i * NV, &gm, op, 1, ff, lab);
    for (int a = 0; a < 4; ++a) { out8[i * 8 + a] = lab[a]; out8[i * 8 + 4 + a] = ff[a]; }
    if (out8_slim) {
      kzo_radstate st2;
      kzo_fill_radstate(pp + i * NV, &gm, gam, op, 0, &st2);
      kzo_calc_gi(&st2, pp + i * NV, &gm, op, 0, ff, lab);
      for (int a = 0; a < 4; ++a) { out8_slim[i * 8 + a] = lab[a]; out8_slim[i * 8 + 4 + a] = ff[a]; }
    }
  }
  return 0;
}

int kzo_invert4_batch(int64_t n, const double *a16, double *inv16) {
  for (int64_t i = 0; i < n; ++i) kzo_invert4((const double(*)[4])(a16 + i * 16), (double(*)[4])(inv16 + i * 16));
  return 0;
}

int kzo_invert4f_batch(int64_t n, const float *a16, float *inv16) {
  for (int64_t i = 0; i < n; ++i) kzo_invert4f((const float(*)[4])(a16 + i * 16), (float(*)[4])(inv16 + i * 16));
  return 0;
}

/* fn: 0 exp 1 log 2 pow 3 cbrt 4 log1p 5 sqrt 6 div 7 pow through
 * kzm_pow_from_log (finite x > 0 only) */
int kzo_math_batch(int64_t n, int32_t fn, const double *x, const double *y, double *out) {
  for (int64_t i = 0; i < n; ++i) {
    double a = x[i], b = y ? y[i] : 0.0, r;
    switch (fn) {
      case 0: r = kzm_exp(a); break;
      case 1: r = kzm_log(a); break;
      case 2: r = kzm_pow(a, b); break;
      case 3: r = kzm_cbrt(a); break;
      case 4: r = kzm_log1p(a); break;
      case 5: r = kzm_sqrt(a); break;
      case 7: {
        double lh, ll;
        kzm_log_dd(a, &lh, &ll);
        r = kzm_pow_from_log(lh, ll, b);
        break;
      }
      case 6:
      default: r = a / b; break;
    }
    out[i] = r;
  }
  return 0;
}
