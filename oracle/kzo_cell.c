/* kzo_cell.c — oracle restatement of the per-point algebra: relele, p2u,
 * bfield, hydro, thermo, radiation, wavespeeds, opacities, radforce, recon,
 * laxf, flux.  TEST INFRASTRUCTURE (see kzo.h).
 */
#include "kzo.h"

static const double four_third = 4.0 / 3.0;
static const double one_third = 1.0 / 3.0;
static const double BIG_THERMO = 1.0 / 1.0e-80; /* units.zig:33-34 */

/* ------------------------------------------------------------------------ */
/* geometry (koral/metric/precompute.zig:466-496 geometryFromBlocks)         */

void kzo_geo_from_kz(const kz_geom *k, kzo_geo *o) {
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) {
      o->g[i][j] = k->gg[i][j];
      o->G[i][j] = k->GG[i][j];
    }
  o->gdet = k->gdet;
  o->alpha = k->alpha;
  o->gttpert = k->gttpert;
  o->x1 = k->xx[1];
}

/* metric.zig:20-22 rHorizonBL */
double kzo_r_horizon_bl(double a) { return 1.0 + kzm_sqrt(1.0 - a * a); }

/* metric.zig:26-32 rIscoBL.  The reference's pow takes y == 1/2 as a square
 * root (std.math.pow's special case); the cube roots go through kzm_pow. */
double kzo_r_isco_bl(double ac) {
  const double c3 = 1.0 / 3.0;
  double z1 = 1.0 + kzm_pow(1.0 - ac * ac, c3) * (kzm_pow(1.0 + ac, c3) + kzm_pow(1.0 - ac, c3));
  double z2 = kzm_sqrt(3.0 * ac * ac + z1 * z1);
  return 3.0 + z2 - kzm_sqrt((3.0 - z1) * (3.0 + z1 + 2.0 * z2));
}

/* precompute.zig:466-496: alpha = sqrt(-1/GG00) recomputed from the block */
void kzo_geo_from_blocks(const double *g20, const double *G20, double x1, kzo_geo *o) {
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) {
      o->g[i][j] = g20[i * 5 + j];
      o->G[i][j] = G20[i * 5 + j];
    }
  o->gdet = g20[3 * 5 + 4];
  o->gttpert = G20[3 * 5 + 4];
  o->alpha = kzm_sqrt(-1.0 / o->G[0][0]);
  o->x1 = x1;
}

/* ------------------------------------------------------------------------ */
/* relele.zig                                                                */

/* relele.zig:89-97 raiseVecG */
void kzo_raise(const double a[4], const kzo_geo *gm, double out[4]) {
  double r[4] = {0, 0, 0, 0};
  for (int i = 0; i < 4; ++i)
    for (int k = 0; k < 4; ++k) r[i] += a[k] * gm->G[i][k];
  for (int i = 0; i < 4; ++i) out[i] = r[i];
}

/* relele.zig:104-112 lowerVecG */
void kzo_lower(const double a[4], const kzo_geo *gm, double out[4]) {
  double r[4] = {0, 0, 0, 0};
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) r[i] += a[j] * gm->g[i][j];
  for (int i = 0; i < 4; ++i) out[i] = r[i];
}

/* relele.zig:72-74 dotG */
double kzo_dot(const double a[4], const double b[4]) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2] + a[3] * b[3];
}

/* relele.zig:147-158 lowerSecond */
void kzo_lower_second(const double t[4][4], const kzo_geo *gm, double out[4][4]) {
  double r[4][4] = {{0}};
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j)
      for (int k = 0; k < 4; ++k) r[i][j] += t[i][k] * gm->g[k][j];
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) out[i][j] = r[i][j];
}

/* relele.zig:165-174 lowerSecondRow */
void kzo_lower_second_row(const double t[4][4], const kzo_geo *gm, int row, double out[4]) {
  double r[4] = {0, 0, 0, 0};
  for (int j = 0; j < 4; ++j)
    for (int k = 0; k < 4; ++k) r[j] += t[row][k] * gm->g[k][j];
  for (int j = 0; j < 4; ++j) out[j] = r[j];
}

/* relele.zig:200-212 utFromVel3 */
static double ut_from_vel3(const double v[4], const kzo_geo *gm) {
  double b = 0, c = 0;
  for (int i = 1; i < 4; ++i) {
    b += v[i] * gm->g[0][i];
    for (int j = 1; j < 4; ++j) c += v[i] * v[j] * gm->g[i][j];
  }
  return kzm_sqrt(-1.0 / (gm->g[0][0] + 2.0 * b + c));
}

/* relele.zig:228-242 utFromSpatialUcon */
static int ut_from_spatial_ucon(const double u[4], const kzo_geo *gm, double *ut) {
  double a = gm->g[0][0];
  double b = 0, c = 1;
  for (int i = 1; i < 4; ++i) {
    b += u[i] * gm->g[0][i];
    for (int j = 1; j < 4; ++j) c += u[i] * u[j] * gm->g[i][j];
  }
  double delta = b * b - a * c;
  if (delta < 0) return KZO_ESPACELIKE;
  *ut = (-b - kzm_sqrt(delta)) / a;
  return KZO_OK;
}

/* relele.zig:250-260 alphaGammaG */
double kzo_alpha_gamma(const double u[4], const kzo_geo *gm) {
  double qsq = 0;
  for (int i = 1; i < 4; ++i)
    for (int j = 1; j < 4; ++j) qsq += u[i] * u[j] * gm->g[i][j];
  double alpgam2 = (-1.0 / gm->G[0][0]) * (1.0 + qsq);
  return alpgam2 < 0 ? 1.0 : kzm_sqrt(alpgam2);
}

/* relele.zig:265-273 velrToVel4G */
void kzo_velr_to_vel4(const double uin[4], const kzo_geo *gm, double out[4]) {
  double alpgam = kzo_alpha_gamma(uin, gm);
  double o0 = -alpgam * gm->G[0][0];
  double o1 = uin[1] - alpgam * gm->G[0][1];
  double o2 = uin[2] - alpgam * gm->G[0][2];
  double o3 = uin[3] - alpgam * gm->G[0][3];
  out[0] = o0; out[1] = o1; out[2] = o2; out[3] = o3;
}

/* relele.zig:282-348 convert */
int kzo_convert(const double uin[4], int from, int to, const kzo_geo *gm, int trust_ut, double out[4]) {
  double uo[4];
  if (from == to) {
    for (int i = 0; i < 4; ++i) uo[i] = uin[i];
    if (from == VEL4 && !trust_ut) {
      int e = ut_from_spatial_ucon(uin, gm, &uo[0]);
      if (e) return e;
    }
    for (int i = 0; i < 4; ++i) out[i] = uo[i];
    return KZO_OK;
  }
  if (from == VEL4) {
    double ut = uin[0];
    if (!trust_ut) {
      int e = ut_from_spatial_ucon(uin, gm, &ut);
      if (e) return e;
    }
    if (to == VEL3) {
      uo[0] = 1.0; uo[1] = uin[1] / ut; uo[2] = uin[2] / ut; uo[3] = uin[3] / ut;
    } else { /* VELR */
      uo[0] = ut;
      for (int i = 1; i < 4; ++i) uo[i] = uin[i] - ut * gm->G[0][i] / gm->G[0][0];
    }
  } else if (from == VEL3) {
    double ut = ut_from_vel3(uin, gm);
    if (ut < 1.0 || ut != ut) return KZO_EVELCONV;
    uo[0] = ut; uo[1] = uin[1] * ut; uo[2] = uin[2] * ut; uo[3] = uin[3] * ut;
    if (to == VELR) {
      for (int i = 1; i < 4; ++i) uo[i] = uo[i] - uo[0] * gm->G[0][i] / gm->G[0][0];
    }
  } else { /* VELR */
    if (to == VEL4) {
      kzo_velr_to_vel4(uin, gm, uo);
    } else { /* VEL3 */
      double alpgam = kzo_alpha_gamma(uin, gm);
      uo[0] = -alpgam * gm->G[0][0];
      uo[1] = uin[1] / uo[0] + gm->G[0][1] / gm->G[0][0];
      uo[2] = uin[2] / uo[0] + gm->G[0][2] / gm->G[0][0];
      uo[3] = uin[3] / uo[0] + gm->G[0][3] / gm->G[0][0];
    }
  }
  for (int i = 0; i < 4; ++i) out[i] = uo[i];
  return KZO_OK;
}

/* relele.zig:375-377 uconUcovFromPrimsG (VELR, infallible) */
void kzo_ucon_ucov_from_prims(const double *pp, const kzo_geo *gm, double ucon[4], double ucov[4]) {
  double v[4] = {0, pp[VX], pp[VY], pp[VZ]};
  kzo_velr_to_vel4(v, gm, ucon);
  kzo_lower(ucon, gm, ucov);
}

/* ------------------------------------------------------------------------ */
/* bfield.zig / hydro.zig                                                    */

/* bfield.zig:23-31 + :44-48 */
void kzo_bcon_bcov_bsq(const double B[3], const double ucon[4], const double ucov[4],
                       const kzo_geo *gm, double bcon[4], double bcov[4], double *bsq) {
  double b[4];
  b[0] = B[0] * ucov[1] + B[1] * ucov[2] + B[2] * ucov[3];
  double inv_ut = 1.0 / ucon[0];
  for (int j = 1; j < 4; ++j) b[j] = (B[j - 1] + b[0] * ucon[j]) * inv_ut;
  for (int j = 0; j < 4; ++j) bcon[j] = b[j];
  kzo_lower(b, gm, bcov);
  *bsq = kzo_dot(bcon, bcov);
}

/* hydro.zig:32-37 sFromUG */
double kzo_s_from_u(double rho, double u, double gam) {
  double gammam1 = gam - 1.0;
  double indexn = 1.0 / gammam1;
  double pre = gammam1 * u;
  return rho * kzm_log(kzm_pow(pre, indexn) / kzm_pow(rho, indexn + 1.0));
}

/* ------------------------------------------------------------------------ */
/* p2u.zig                                                                   */

/* p2u.zig:17-36 calcUtp1 */
double kzo_calc_utp1(const double vcon[4], const kzo_geo *gm) {
  double qsq = 0;
  for (int i = 1; i < 4; ++i)
    for (int j = 1; j < 4; ++j) qsq += vcon[i] * vcon[j] * gm->g[i][j];
  double alpha = gm->alpha;
  double alphasq = alpha * alpha;
  double alpgam = kzm_sqrt(alphasq * (1.0 + qsq));
  double betasqoalphasq = gm->g[0][1] * gm->G[0][1] + gm->g[0][2] * gm->G[0][2] + gm->g[0][3] * gm->G[0][3];
  double ud0tilde = 0;
  for (int j = 1; j < 4; ++j) ud0tilde += vcon[j] * gm->g[0][j];
  return ud0tilde + (gm->gttpert - alphasq * (betasqoalphasq + qsq)) / (1.0 + alpgam);
}

/* p2u.zig:39-100 p2uMhd */
void kzo_p2u_mhd(const double *pp, double *uu, const kzo_geo *gm, double gam) {
  double gdetu = gm->gdet;
  double rho = pp[RHO], ugas = pp[UU], s = pp[ENTR];
  double vcon[4] = {0, pp[VX], pp[VY], pp[VZ]};
  double ucon[4], ucov[4];
  kzo_velr_to_vel4(vcon, gm, ucon);
  kzo_lower(ucon, gm, ucov);
  double B[3] = {pp[B1], pp[B2], pp[B3]};
  double bcon[4], bcov[4], bsq;
  kzo_bcon_bcov_bsq(B, ucon, ucov, gm, bcon, bcov, &bsq);

  double ut = ucon[0];
  double rhout = rho * ut;
  double sut = s * ut;
  double pre = (gam - 1.0) * ugas;
  double w = rho + ugas + pre;
  double eta = w + bsq;
  double etap = ugas + pre + bsq;
  double ptot = pre + 0.5 * bsq;
  double utp1 = kzo_calc_utp1(vcon, gm);

  double tttt = etap * ucon[0] * ucov[0] + rho * ucon[0] * utp1 + ptot - bcon[0] * bcov[0];
  double ttr = eta * ucon[0] * ucov[1] - bcon[0] * bcov[1];
  double ttth = eta * ucon[0] * ucov[2] - bcon[0] * bcov[2];
  double ttph = eta * ucon[0] * ucov[3] - bcon[0] * bcov[3];

  uu[RHO] = gdetu * rhout;
  uu[UU] = gdetu * tttt;
  uu[VX] = gdetu * ttr;
  uu[VY] = gdetu * ttth;
  uu[VZ] = gdetu * ttph;
  uu[ENTR] = gdetu * sut;
  uu[B1] = gdetu * pp[B1];
  uu[B2] = gdetu * pp[B2];
  uu[B3] = gdetu * pp[B3];
}

/* p2u.zig:103-132 p2uRad */
void kzo_p2u_rad(const double *pp, double *uu, const kzo_geo *gm) {
  double gdetu = gm->gdet;
  double erf = pp[EE];
  double v[4] = {0, pp[FX], pp[FY], pp[FZ]};
  double urf[4];
  kzo_velr_to_vel4(v, gm, urf);
  double rtopp[4];
  for (int i = 0; i < 4; ++i) rtopp[i] = four_third * erf * urf[0] * urf[i] + (one_third) * erf * gm->G[0][i];
  kzo_lower(rtopp, gm, rtopp);
  for (int i = 0; i < 4; ++i) uu[EE + i] = gdetu * rtopp[i];
}

/* p2u.zig:135-148 */
void kzo_p2u(const double *pp, double *uu, const kzo_geo *gm, double gam) {
  for (int i = 0; i < NV; ++i) uu[i] = 0;
  kzo_p2u_mhd(pp, uu, gm, gam);
  kzo_p2u_rad(pp, uu, gm);
}

/* hydro.zig:64-112 calcTij (VELR: infallible) */
int kzo_calc_tij(const double *pp, const kzo_geo *gm, double gam, double t[4][4]) {
  double ucon[4], ucov[4];
  kzo_ucon_ucov_from_prims(pp, gm, ucon, ucov);
  double B[3] = {pp[B1], pp[B2], pp[B3]};
  double bcon[4], bcov[4], bsq;
  kzo_bcon_bcov_bsq(B, ucon, ucov, gm, bcon, bcov, &bsq);
  double rho = pp[RHO], u = pp[UU];
  double p = (gam - 1.0) * u;
  double w = rho + u + p;
  double eta = w + bsq;
  double ptot = p + 0.5 * bsq;
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) t[i][j] = eta * ucon[i] * ucon[j] + ptot * gm->G[i][j] - bcon[i] * bcon[j];
  return KZO_OK;
}

/* ------------------------------------------------------------------------ */
/* thermo.zig                                                                */

/* thermo.zig:166-169 tFromUrhoG */
double kzo_t_from_urho(const kz_opac_params *c, double uint, double rho, double gam) {
  double p = uint * (gam - 1.0);
  return c->mugas_mp_over_kb * p / rho;
}

/* thermo.zig:155-157 lteTfromEG */
double kzo_lte_t_from_e(const kz_opac_params *c, double e) {
  return kzm_sqrt(kzm_sqrt(c->one_over_four_sigmarad * e));
}

/* thermo.zig:191-201 tempsFromUrhoG (single temperature) */
static void temps_from_urho(const kz_opac_params *c, double uint, double rho, double gam,
                            double *tgas, double *te, double *ti) {
  double tg = kzo_t_from_urho(c, uint, rho, gam);
  double e = tg, i = tg;
  e = kzm_isfinite(e) ? e : BIG_THERMO;
  i = kzm_isfinite(i) ? i : BIG_THERMO;
  e = e < 1.0e2 ? 1.0e2 : e;
  i = i < 1.0e2 ? 1.0e2 : i;
  *tgas = tg; *te = e; *ti = i;
}

/* ------------------------------------------------------------------------ */
/* radiation.zig                                                             */

/* radiation.zig:79-98 calcRijG */
void kzo_calc_rij(const double *pp, const kzo_geo *gm, double rij[4][4]) {
  double erf = pp[EE];
  double v[4] = {0, pp[FX], pp[FY], pp[FZ]};
  double urf[4];
  kzo_velr_to_vel4(v, gm, urf);
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j)
      rij[i][j] = four_third * erf * urf[i] * urf[j] + one_third * erf * gm->G[i][j];
}

/* radiation.zig:104-124 calcFfRtt; returns rtt (Ehat = -rtt) */
double kzo_calc_ff_rtt(const double *pp, const kzo_geo *gm, double ucon_out[4]) {
  double ucon[4], ucov[4];
  kzo_ucon_ucov_from_prims(pp, gm, ucon, ucov);
  double rij[4][4], rud[4][4];
  kzo_calc_rij(pp, gm, rij);
  kzo_lower_second(rij, gm, rud);
  double rtt = 0;
  for (int ia = 0; ia < 4; ++ia)
    for (int ib = 0; ib < 4; ++ib) rtt += -rud[ia][ib] * ucon[ib] * ucov[ia];
  if (ucon_out)
    for (int i = 0; i < 4; ++i) ucon_out[i] = ucon[i];
  return rtt;
}

/* ------------------------------------------------------------------------ */
/* wavespeeds.zig                                                            */

/* wavespeeds.zig:32-81 lrCore */
void kzo_lr_core(const double ucon[4], const kzo_geo *gm, const double ws2[3], const int active[3],
                 double aret[6]) {
  for (int i = 0; i < 6; ++i) aret[i] = 0;
  double Bcov[4] = {1, 0, 0, 0};
  double Bcon[4];
  kzo_raise(Bcov, gm, Bcon);
  double Bsq = kzo_dot(Bcon, Bcov);
  double Bu = kzo_dot(Bcov, ucon);
  double Bu2 = Bu * Bu;
  for (int dim = 0; dim < 3; ++dim) {
    if (!active[dim]) continue;
    double wspeed2 = ws2[dim];
    double Acov[4] = {0, 0, 0, 0};
    Acov[dim + 1] = 1;
    double Acon[4];
    kzo_raise(Acov, gm, Acon);
    double Asq = kzo_dot(Acon, Acov);
    double Au = kzo_dot(Acov, ucon);
    double AB = kzo_dot(Acon, Bcov);
    double Au2 = Au * Au;
    double AuBu = Au * Bu;
    double b = -2.0 * (AuBu * (1.0 - wspeed2) - AB * wspeed2);
    double a = Bu2 * (1.0 - wspeed2) - Bsq * wspeed2;
    double discr = 4.0 * wspeed2 * ((AB * AB - Asq * Bsq) * wspeed2 +
                                    (2.0 * AB * Au * Bu - Asq * Bu2 - Bsq * Au2) * (wspeed2 - 1.0));
    discr = kzm_sqrt(discr);
    double cst1 = (-b + discr) / (2.0 * a);
    double cst2 = (-b - discr) / (2.0 * a);
    if (cst2 > cst1) {
      aret[2 * dim] = cst1;
      aret[2 * dim + 1] = cst2;
    } else {
      aret[2 * dim] = cst2;
      aret[2 * dim + 1] = cst1;
    }
  }
}

/* wavespeeds.zig:85-94 gasWspeed2 */
static double gas_wspeed2(double rho, double uu, double bsq, double gam) {
  double pre = (gam - 1.0) * uu;
  double cs2 = gam * pre / (rho + uu + pre);
  if (cs2 > 0.95) cs2 = 0.95;
  double va2 = bsq / (bsq + (rho + uu + pre));
  if (va2 < 0.0) va2 = 0.0;
  return cs2 + va2 - cs2 * va2;
}

/* wavespeeds.zig:98-132 gasWavespeedsLr */
void kzo_gas_wavespeeds_lr(const double *pp, const kzo_geo *gm, double gam, const int active[3], double a[6]) {
  double ucon[4], ucov[4];
  kzo_ucon_ucov_from_prims(pp, gm, ucon, ucov);
  double B[3] = {pp[B1], pp[B2], pp[B3]};
  double bcon[4], bcov[4], bsq;
  kzo_bcon_bcov_bsq(B, ucon, ucov, gm, bcon, bcov, &bsq);
  double vtot2 = gas_wspeed2(pp[RHO], pp[UU], bsq, gam);
  double ws[3] = {vtot2, vtot2, vtot2};
  kzo_lr_core(ucon, gm, ws, active, a);
  for (int d = 0; d < 3; ++d) {
    if (a[2 * d] > 0.0) a[2 * d] = 0.0;
    if (a[2 * d + 1] < 0.0) a[2 * d + 1] = 0.0;
  }
}

/* radiation.zig:177-211 calcRadWavespeeds */
void kzo_calc_rad_wavespeeds(const double *pp, const kzo_geo *gm, const double tautot[3],
                             const int active[3], double aval[12]) {
  double v[4] = {0, pp[FX], pp[FY], pp[FZ]};
  double urf[4];
  kzo_velr_to_vel4(v, gm, urf);
  double rv2rad = one_third;
  double ws0[3] = {rv2rad, rv2rad, rv2rad};
  kzo_lr_core(urf, gm, ws0, active, aval);
  double rv2dim[3];
  for (int d = 0; d < 3; ++d) {
    if (tautot[d] <= 0.0) {
      rv2dim[d] = rv2rad;
    } else {
      double rv2tau = (four_third * four_third) / (tautot[d] * tautot[d]);
      rv2dim[d] = kzm_fmin(rv2rad, rv2tau);
    }
  }
  double ws1[3] = {rv2dim[0], rv2dim[1], rv2dim[1]}; /* radiation.zig:205 z uses y */
  kzo_lr_core(urf, gm, ws1, active, aval + 6);
}

/* ------------------------------------------------------------------------ */
/* opacities.zig                                                             */

/* opacities.zig:131-315 calcOpacitiesFromStateG (no MESA), with the
   synchrotron bridge of opacities.zig:208-297 when c->synchrotron_bridge */
void kzo_calc_opacities(const kz_opac_params *c, double rho, double tgas, double te, double trad,
                        double tradbb, double ne, double bsq, int level, kzo_opac *o) {
  (void)tgas;
  int want_rad = level != OPAC_CHI;
  int want_full = level == OPAC_FULL;
  double rt_te = kzm_sqrt(te);
  double rhocgs = c->rhogu2cgs * rho;
  double nethcgs = ne * c->numdensgu2cgs;
  (void)rhocgs;
  double bsqcgs = c->fourpi * (bsq * c->enden_gu2cgs_factor);
  double bmagcgs = kzm_sqrt(bsqcgs);
  double bbenergy = (4.0 * c->sigma_rad_cgs) * te * te * te * te;

  double kappagasff = 0, kapparadff = 0, kappagasffross = 0, kapparadffross = 0;
  double kappagassyn = 0, kapparadsyn = 0, kapparadnumsyn = 0, kapparadsynross = 0, kappagassynross = 0;

  if (c->bremsstrahlung) {
    double mfrac = 1.0 - c->hfrac - c->hefrac;
    double emis_ff = 1.4e-27 * nethcgs * ((c->hfrac + c->hefrac + c->z2diva_mean * mfrac) / c->m_proton_cgs) *
                     rt_te * 1.2 * (1.0 + 4.4e-10 * te);
    kappagasff = c->kappacgs2gu * (emis_ff / bbenergy) * rho;
    if (want_rad) {
      double zeta = trad / te;
      double zeta_inv = 1.0 / zeta;
      double zeta_inv_3 = zeta_inv * zeta_inv * zeta_inv;
      kapparadff = kappagasff * kzm_log1p(1.6 * zeta) * c->one_over_log_2p6 * zeta_inv_3;
      double zeta_root5 = kzm_pow(zeta, 0.2);
      double zeta_root5_inv_4 = 1.0 / (zeta_root5 * zeta_root5 * zeta_root5 * zeta_root5);
      double zeta_root5_inv_3 = zeta_root5 * zeta_root5_inv_4;
      double scale_pla_ros_ff = 14.12 / (432.7 - 106.8 * zeta_root5_inv_3 + 43.17 * zeta_root5_inv_4 + 57.88 * zeta_inv);
      kapparadffross = kappagasff * scale_pla_ros_ff * zeta_inv_3;
    }
    if (want_full) kappagasffross = kappagasff * 0.0330;
  }

  if (c->synchrotron) {
    /* the bridge clamps the synchrotron Trad up to TradBB^(4/3) / Te^(1/3),
       adds the non-relativistic components to the gas and radiation
       absorption, applies the number-opacity crossover at tc_n and drops the
       Terelfactor suppression */
    const int bridge = c->synchrotron_bridge != 0;
    const double tc_n = 5.07783e9;
    double trad_syn = trad;
    if (bridge) {
      double trad_lim = kzm_pow(tradbb, 4.0 / 3.0) / kzm_pow(te, 1.0 / 3.0);
      trad_syn = trad < trad_lim ? trad_lim : trad;
    }
    double emis_synchro = 3.61e-34 * (nethcgs / rhocgs) * te * te * bmagcgs * bmagcgs;
    kappagassyn = c->kappacgs2gu * (emis_synchro / bbenergy) * rho;
    if (want_rad) {
      double nu_mbsyn = 1.19e-13 * te * te;
      double zbr = c->k_boltz_cgs * trad_syn / c->h_cgs / nu_mbsyn;
      double zeta_a_denom_rad =
          1.79 * kzm_cbrt(zbr * zbr * zbr * zbr * zbr * bmagcgs * bmagcgs * bmagcgs * bmagcgs) +
          1.35 * kzm_cbrt(zbr * zbr * zbr * zbr * zbr * zbr * zbr * bmagcgs * bmagcgs) +
          0.248 * zbr * zbr * zbr;
      double ia_by_b_rad = bmagcgs * bmagcgs / zeta_a_denom_rad;
      double te5 = te * te * te * te * te;
      kapparadsyn = c->kappacgs2gu * (2.13e39 * (nethcgs / rhocgs) / te5 * ia_by_b_rad) * rho;
      if (bridge)
        kapparadsyn += c->kappacgs2gu *
                       (2.35869e-21 * (nethcgs / rhocgs) * (bmagcgs * bmagcgs) / (trad_syn * trad_syn * trad_syn)) * rho;
      if (want_full) {
        double zeta_a_denom_num = (0.025 * kzm_cbrt(zbr * zbr * zbr * zbr * bmagcgs * bmagcgs) +
                                   0.169 * kzm_cbrt(zbr * zbr * zbr * zbr * zbr * bmagcgs) + 0.287 * zbr * zbr);
        double ia_by_b_num = bmagcgs / zeta_a_denom_num;
        kapparadnumsyn = c->kappacgs2gu * (2.13e39 * (nethcgs / rhocgs) / te5 * ia_by_b_num) * rho;
        if (bridge) kapparadnumsyn *= (te / tc_n) / (1.0 + te / tc_n);
      }
      double bmb = bmagcgs > 0.0 ? kzm_pow(bmagcgs, -0.463) : 0.0;
      double ia_by_b_ross_rad = 0.13 * kzm_pow(bmagcgs, 0.69) /
                                (kzm_pow(zbr, 1.69) * kzm_exp(1.6 * kzm_pow(zbr, 0.463) * bmb));
      kapparadsynross = c->kappacgs2gu * (2.13e39 * (nethcgs / rhocgs) / te5 * ia_by_b_ross_rad) * rho;
      if (want_full) {
        double zbg = c->k_boltz_cgs * te / c->h_cgs / nu_mbsyn;
        double ia_by_b_ross_gas = 0.13 * kzm_pow(bmagcgs, 0.69) /
                                  (kzm_pow(zbg, 1.69) * kzm_exp(1.6 * kzm_pow(zbg, 0.463) * bmb));
        kappagassynross = c->kappacgs2gu * (2.13e39 * (nethcgs / rhocgs) / te5 * ia_by_b_ross_gas) * rho;
      }
    }
    if (bridge)
      kappagassyn += c->kappacgs2gu * (2.35869e-21 * (nethcgs / rhocgs) * (bmagcgs * bmagcgs) / (te * te * te)) * rho;
    if (!bridge) {
      double terel = te * c->k_over_mecsq;
      double terelfactor = (terel * terel) / (1.0 + terel * terel);
      kappagassyn *= terelfactor;
      if (want_rad) {
        kapparadsyn *= terelfactor;
        kapparadsynross *= terelfactor;
        if (want_full) {
          kapparadnumsyn *= terelfactor;
          kappagassynross *= terelfactor;
        }
      }
    }
  }

  o->gas_abs = kappagasff + kappagassyn;
  o->rad_abs = kapparadff + kapparadsyn;
  o->gas_num = kappagasff;
  o->rad_num = kapparadff + kapparadnumsyn;
  o->gas_ross = kappagasffross + kappagassynross;
  o->rad_ross = kapparadffross + kapparadsynross;
  o->tot_emissivity = 0;
}

/* opacities.zig:352-360 kappaEsPuffyG */
double kzo_kappa_es_puffy(const kz_opac_params *c, double rho, double trad) {
  double corr = 1.0;
  if (c->kleinnishina) corr = 1.0 / (1.0 + kzm_pow(trad / 4.5e8, 0.86));
  return ((0.2 * (1.0 + c->hfrac)) * corr) * c->kappa_cgs2gu_factor * rho;
}

/* radforce.zig:140-151 kappaesAtG */
static double kappaes_at(const kz_opac_params *c, double rho, double trad) {
  if (c->kappaes_mode == 1) return kzo_kappa_es_puffy(c, rho, trad);
  if (c->kappaes_mode == 2) return c->kappaes_grey * rho;
  return 0.0;
}

/* ------------------------------------------------------------------------ */
/* radforce.zig                                                              */

/* radforce.zig:241-351 fillRadStateCoreG */
void kzo_fill_radstate(const double *pp, const kzo_geo *gm, double gam, const kz_opac_params *c,
                       int full, kzo_radstate *st) {
  double rho = pp[RHO], uint = pp[UU];
  double tgas, te, ti;
  temps_from_urho(c, uint, rho, gam, &tgas, &te, &ti);
  double sgas = full ? c->kb_over_mugas_mp * kzo_s_from_u(rho, uint, gam) : 0.0;
  double ne = c->one_over_mue_mp * rho;
  double ucon[4], ucov[4];
  kzo_ucon_ucov_from_prims(pp, gm, ucon, ucov);
  double B[3] = {pp[B1], pp[B2], pp[B3]};
  double bcon[4], bcov[4], bsq;
  kzo_bcon_bcov_bsq(B, ucon, ucov, gm, bcon, bcov, &bsq);
  kzo_calc_rij(pp, gm, st->rij);
  double ehat = 0;
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) ehat += st->rij[i][j] * ucov[i] * ucov[j];
  double tradbb = kzo_lte_t_from_e(c, ehat);
  double trad = tradbb;
  double kappaes = kappaes_at(c, rho, trad);
  if (c->kappa_mode == 0) {
    kzo_calc_opacities(c, rho, tgas, te, trad, tradbb, ne, bsq, full ? OPAC_FULL : OPAC_RESIDUAL, &st->opac);
    if (full) {
      double b = c->sigma_rad_over_pi * te * te * te * te;
      st->opac.tot_emissivity = st->opac.gas_abs * c->fourpi * b;
    }
    st->kappa = st->opac.gas_abs;
  } else {
    double kap = c->kappa_grey * rho;
    double b = c->sigma_rad_over_pi * te * te * te * te;
    st->opac.gas_abs = kap; st->opac.rad_abs = kap; st->opac.gas_num = kap; st->opac.rad_num = kap;
    st->opac.gas_ross = kap; st->opac.rad_ross = kap;
    st->opac.tot_emissivity = kap * c->fourpi * b;
    st->kappa = kap;
  }
  st->rho = rho; st->uint = uint; st->tgas = tgas; st->te = te; st->ti = ti; st->sgas = sgas; st->ne = ne;
  for (int i = 0; i < 4; ++i) { st->ucon[i] = ucon[i]; st->ucov[i] = ucov[i]; }
  st->bsq = bsq; st->ehat = ehat; st->tradbb = tradbb; st->trad = trad; st->kappaes = kappaes;
}

/* radforce.zig:370-375 comptonGiCoeffG */
double kzo_compton_coeff(const kz_opac_params *c, const kzo_radstate *st) {
  double thetae = c->kb_over_me * st->te;
  return st->kappaes * st->ehat * ((4.0 * c->kb_over_me) * (st->trad - st->te)) *
         (1.0 + 3.683 * thetae + 4.0 * thetae * thetae) / (1.0 + 4.0 * thetae);
}

/* frames.zig lorentzFromPairG / normalObsConCovG / boost2Lab2FfG (full Gi only) */
static void boost2_lab2ff(const double a1[4], const double *pp, const kzo_geo *gm, double a2[4]) {
  double ucon[4], ucov[4];
  kzo_ucon_ucov_from_prims(pp, gm, ucon, ucov);
  double alpha = kzm_sqrt(-1.0 / gm->G[0][0]);
  double wcov[4] = {-alpha, 0, 0, 0};
  double wcon[4];
  kzo_raise(wcov, gm, wcon);
  double om[4][4];
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) om[i][j] = ucon[i] * wcov[j] - wcon[i] * ucov[j];
  double gam = -kzo_dot(wcon, ucov);
  double l[4][4];
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) {
      double omsum = 0;
      for (int k = 0; k < 4; ++k) omsum += om[i][k] * om[k][j];
      l[i][j] = (i == j ? 1.0 : 0.0) + 1.0 / (1.0 + gam) * omsum + om[i][j];
    }
  double r[4] = {0, 0, 0, 0};
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) r[i] += l[i][j] * a1[j];
  double alpha2 = kzm_sqrt(-1.0 / gm->G[0][0]);
  for (int i = 0; i < 4; ++i) a2[i] = r[i] * alpha2;
}

/* radforce.zig:432-494 calcGiFromStateCoreG */
void kzo_calc_gi(const kzo_radstate *st, const double *pp, const kzo_geo *gm,
                 const kz_opac_params *c, int full, double gi_ff[4], double gi_lab[4]) {
  double b = c->sigma_rad_over_pi * st->te * st->te * st->te * st->te;
  double kra = st->opac.rad_abs, kga = st->opac.gas_abs, krr = st->opac.rad_ross;
  double kes = st->kappaes, ruu = st->ehat;
  for (int i = 0; i < 4; ++i) {
    double ru = 0;
    for (int j = 0; j < 4; ++j) ru += st->rij[i][j] * st->ucov[j];
    gi_lab[i] = -(krr + kes) * ru - ((krr + kes - kra) * ruu + kga * c->fourpi * b) * st->ucon[i];
  }
  if (full) {
    boost2_lab2ff(gi_lab, pp, gm, gi_ff);
  } else {
    gi_ff[0] = gi_ff[1] = gi_ff[2] = gi_ff[3] = 0;
  }
  gi_ff[0] = -kga * c->fourpi * b + kra * st->ehat;
  if (c->comptonization) {
    double coeff = kzo_compton_coeff(c, st);
    for (int i = 0; i < 4; ++i) gi_lab[i] += coeff * st->ucon[i];
    gi_ff[0] += coeff;
  }
  if (c->opdamp != 1.0) {
    double od = c->opdamp;
    for (int i = 0; i < 4; ++i) {
      gi_ff[i] *= od;
      gi_lab[i] *= od;
    }
  }
}

/* radforce.zig:515-527 calcKappaes (standalone, Trad = Te) */
double kzo_calc_kappaes_standalone(const double *pp, double gam, const kz_opac_params *c) {
  double tg, te, ti;
  temps_from_urho(c, pp[UU], pp[RHO], gam, &tg, &te, &ti);
  return kappaes_at(c, pp[RHO], te);
}

/* radforce.zig:562-609 calcChiSlim */
double kzo_calc_chi_slim(const double *pp, const kzo_geo *gm, double gam, const kz_opac_params *c) {
  double rho = pp[RHO], uint = pp[UU];
  double tgas, te, ti;
  temps_from_urho(c, uint, rho, gam, &tgas, &te, &ti);
  double ne = c->one_over_mue_mp * rho;
  double ucon[4], ucov[4];
  kzo_ucon_ucov_from_prims(pp, gm, ucon, ucov);
  double B[3] = {pp[B1], pp[B2], pp[B3]};
  double bcon[4], bcov[4], bsq;
  kzo_bcon_bcov_bsq(B, ucon, ucov, gm, bcon, bcov, &bsq);
  double kappa;
  if (c->kappa_mode == 0) {
    kzo_opac o;
    kzo_calc_opacities(c, rho, tgas, te, te, te, ne, bsq, OPAC_CHI, &o);
    kappa = o.gas_abs;
  } else {
    kappa = c->kappa_grey * rho;
  }
  double kappaes = kappaes_at(c, rho, te);
  return kappa + kappaes;
}

/* ------------------------------------------------------------------------ */
/* fv/recon.zig, fv/laxf.zig                                                 */

static double sgn(double x) { return x >= 0.0 ? 1.0 : -1.0; }

/* recon.zig:77-91 linearMinmod */
static void linear_minmod(const double u[5], double theta, double *l, double *r) {
  double deltam = u[2] - u[1];
  double deltap = u[3] - u[2];
  double smin = kzm_fmin(kzm_fmin(theta * deltam, 0.5 * (deltam + deltap)), theta * deltap);
  double smax = kzm_fmax(kzm_fmax(theta * deltam, 0.5 * (deltam + deltap)), theta * deltap);
  double slope = deltam > 0.0 ? smin : smax;
  slope = (deltam * deltap <= 0.0) ? 0.0 : slope;
  *l = u[2] - 0.5 * slope;
  *r = u[2] + 0.5 * slope;
}

/* recon.zig:99-195 ppm */
static void ppm(const double u[5], const double dx[5], double *lo, double *ro) {
  double dxm2 = dx[0], dxm1 = dx[1], dx0 = dx[2], dxp1 = dx[3], dxp2 = dx[4];
  double dxp2_plus_dxp1 = dxp2 + dxp1;
  double dxp2_plus_dxp1_inv = 1.0 / dxp2_plus_dxp1;
  double dxp1_plus_dx0 = dxp1 + dx0;
  double dxp1_plus_dx0_inv = 1.0 / dxp1_plus_dx0;
  double dx0_plus_dxm1 = dx0 + dxm1;
  double dx0_plus_dxm1_inv = 1.0 / dx0_plus_dxm1;
  double dxm1_plus_dxm2 = dxm1 + dxm2;
  double dxm1_plus_dxm2_inv = 1.0 / dxm1_plus_dxm2;
  double dxm1_plus_dx0_plus_dxp1_inv = 1.0 / (dxm1 + dx0 + dxp1);
  double dx0_plus_dxp1_plus_dxp2_inv = 1.0 / (dx0 + dxp1 + dxp2);
  double dxm2_plus_dxm1_plus_dx0_inv = 1.0 / (dxm2 + dxm1 + dx0);
  double dx0_plus_twodxm1 = dx0 + 2.0 * dxm1;
  double dxp1_plus_twodx0 = dxp1 + 2.0 * dx0;
  double dxm1_plus_twodxm2 = dxm1 + 2.0 * dxm2;
  double twodxp1_plus_dx0 = 2.0 * dxp1 + dx0;
  double twodxp2_plus_dxp1 = 2.0 * dxp2 + dxp1;
  double twodx0_plus_dxm1 = 2.0 * dx0 + dxm1;

  double dri = dx0 * dxm1_plus_dx0_plus_dxp1_inv *
               (dx0_plus_twodxm1 * dxp1_plus_dx0_inv * (u[3] - u[2]) +
                twodxp1_plus_dx0 * dx0_plus_dxm1_inv * (u[2] - u[1]));
  double drip1 = dxp1 * dx0_plus_dxp1_plus_dxp2_inv *
                 (dxp1_plus_twodx0 * dxp2_plus_dxp1_inv * (u[4] - u[3]) +
                  twodxp2_plus_dxp1 * dxp1_plus_dx0_inv * (u[3] - u[2]));
  double drim1 = dxm1 * dxm2_plus_dxm1_plus_dx0_inv *
                 (dxm1_plus_twodxm2 * dx0_plus_dxm1_inv * (u[2] - u[1]) +
                  twodx0_plus_dxm1 * dxm1_plus_dxm2_inv * (u[1] - u[0]));

  dri = ((u[3] - u[2]) * (u[2] - u[1]) > 0.0)
            ? kzm_fmin(kzm_fabs(dri), kzm_fmin(2.0 * kzm_fabs(u[2] - u[1]), 2.0 * kzm_fabs(u[2] - u[3]))) * sgn(dri)
            : 0.0;
  drip1 = ((u[4] - u[3]) * (u[3] - u[2]) > 0.0)
              ? kzm_fmin(kzm_fabs(drip1), kzm_fmin(2.0 * kzm_fabs(u[3] - u[2]), 2.0 * kzm_fabs(u[3] - u[4]))) * sgn(drip1)
              : 0.0;
  drim1 = ((u[2] - u[1]) * (u[1] - u[0]) > 0.0)
              ? kzm_fmin(kzm_fabs(drim1), kzm_fmin(2.0 * kzm_fabs(u[1] - u[0]), 2.0 * kzm_fabs(u[1] - u[2]))) * sgn(drim1)
              : 0.0;

  double z1 = dx0_plus_dxm1 / dxp1_plus_twodx0;
  double z2 = dxp2_plus_dxp1 / twodxp1_plus_dx0;
  double dx_inv = 1.0 / (dxm1 + dx0 + dxp1 + dxp2);
  double right = u[2] + dx0 * dxp1_plus_dx0_inv * (u[3] - u[2]) +
                 dx_inv * ((2.0 * dxp1 * dx0) * dxp1_plus_dx0_inv * (z1 - z2) * (u[3] - u[2]) -
                           dx0 * z1 * drip1 + dxp1 * z2 * dri);
  z1 = dxm1_plus_dxm2 / dx0_plus_twodxm1;
  z2 = dxp1_plus_dx0 / twodx0_plus_dxm1;
  dx_inv = 1.0 / (dxm2 + dxm1 + dx0 + dxp1);
  double left = u[1] + dxm1 * dx0_plus_dxm1_inv * (u[2] - u[1]) +
                dx_inv * ((2.0 * dx0 * dxm1) * dx0_plus_dxm1_inv * (z1 - z2) * (u[2] - u[1]) -
                          dxm1 * z1 * dri + dx0 * z2 * drim1);

  int flat = (right - u[2]) * (u[2] - left) <= 0.0;
  if (flat) {
    left = u[2];
    right = u[2];
  }
  if ((right - left) * (left - (3.0 * u[2] - 2.0 * right)) < 0.0) left = 3.0 * u[2] - 2.0 * right;
  if ((right - left) * ((3.0 * u[2] - 2.0 * left) - right) < 0.0) right = 3.0 * u[2] - 2.0 * left;
  *lo = left;
  *ro = right;
}

/* recon.zig:47-53 reconstruct; order 0 donor, 1 linear, 2 PPM */
void kzo_reconstruct(int order, double theta, const double u[5], const double dx[5], double *left, double *right) {
  if (order <= 0) {
    *left = u[2];
    *right = u[2];
  } else if (order == 1) {
    linear_minmod(u, theta, left, right);
  } else {
    ppm(u, dx, left, right);
  }
}

/* laxf.zig:13-15 */
double kzo_laxf(double fl, double fr, double ul, double ur, double ag) { return 0.5 * (fr + fl - ag * (ur - ul)); }

/* ------------------------------------------------------------------------ */
/* physics/flux.zig:40-134 fFluxPrime                                        */
int kzo_f_flux_prime(const double *pp, int idim, const kzo_geo *gm, double gam, double *ff) {
  for (int i = 0; i < NV; ++i) ff[i] = 0;
  double gdetu = gm->gdet;
  double rho = pp[RHO], ugas = pp[UU], s = pp[ENTR];
  double vcon[4] = {0, pp[VX], pp[VY], pp[VZ]};
  double ucon[4], ucov[4];
  kzo_velr_to_vel4(vcon, gm, ucon);
  kzo_lower(ucon, gm, ucov);
  double B[3] = {pp[B1], pp[B2], pp[B3]};
  double bcon[4], bcov[4], bsq;
  kzo_bcon_bcov_bsq(B, ucon, ucov, gm, bcon, bcov, &bsq);

  double pre0 = (gam - 1.0) * ugas;
  double w = rho + ugas + pre0;
  double eta = w + bsq;
  double ptot = pre0 + 0.5 * bsq;
  double tij[4][4];
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) tij[i][j] = eta * ucon[i] * ucon[j] + ptot * gm->G[i][j] - bcon[i] * bcon[j];
  double trow[4];
  kzo_lower_second_row(tij, gm, idim + 1, trow);

  double pre = (gam - 1.0) * ugas;
  double etap = ugas + pre + bsq;
  double utp1 = kzo_calc_utp1(vcon, gm);
  int d = idim + 1;
  ff[RHO] = gdetu * rho * ucon[d];
  ff[ENTR] = gdetu * s * ucon[d];
  ff[UU] = gdetu * (etap * ucon[d] * ucov[0] + rho * ucon[d] * utp1);
  ff[UU] += -gdetu * bcon[d] * bcov[0];
  ff[VX] = gdetu * trow[1];
  ff[VY] = gdetu * trow[2];
  ff[VZ] = gdetu * trow[3];
  ff[B1] = gdetu * (bcon[1] * ucon[d] - bcon[d] * ucon[1]);
  ff[B2] = gdetu * (bcon[2] * ucon[d] - bcon[d] * ucon[2]);
  ff[B3] = gdetu * (bcon[3] * ucon[d] - bcon[d] * ucon[3]);

  double rij[4][4], rrow[4];
  kzo_calc_rij(pp, gm, rij);
  kzo_lower_second_row(rij, gm, d, rrow);
  for (int nu = 0; nu < 4; ++nu) ff[EE + nu] = gdetu * rrow[nu];

  for (int i = 0; i < NV; ++i)
    if (!kzm_isfinite(ff[i])) return KZO_ENANFLUX;
  return KZO_OK;
}
