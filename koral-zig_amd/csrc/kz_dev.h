// kz_dev.h — per-point physics of the koral step as gfx950 device code.
//
// One GPU thread evaluates one cell or face; every function below restates
// one reference function with the same operand order and association (the
// file:line is cited on each), so that with -ffp-contract=off and the shared
// kz_math.h transcendentals the kernels reproduce the CPU path bit-for-bit.
// Geometry lives in registers as a Geo (full 4x4 blocks: the reference's
// contractions include the zero entries, and skipping them would change the
// result when an operand is inf/NaN).
#pragma once

#include <stdint.h>
#include <hip/hip_runtime.h>

#include "../../include/kz_types.h"

#define KZM_QUAL __host__ __device__ static inline
#include "kz_math.h"

#define KZD __device__ static inline

namespace kzd {

constexpr int NV = KZ_NV;
enum { RHO = 0, UU, VX, VY, VZ, ENTR, B1, B2, B3, EE, FX, FY, FZ };
enum { WP_RAD = 1, WP_MHD = 2, WE_ENERGY = 0, WE_ENTROPY = 1, WF_LAB = 0, WF_FF = 1 };
enum { OPAC_FULL = 0, OPAC_RESIDUAL = 1, OPAC_CHI = 2 };
enum { E_OK = 0, E_SPACELIKE = KZ_E_SPACELIKE, E_VELCONV = KZ_E_VELCONV, E_NANFLUX = KZ_E_NAN_IN_FLUX };

constexpr double four_third = 4.0 / 3.0;
constexpr double one_third = 1.0 / 3.0;

struct Geo {
  double g[4][4];
  double G[4][4];
  double gdet, alpha, gttpert, x1;
};

// ---- relele.zig ---------------------------------------------------------
KZD void raise4(const double a[4], const Geo &gm, double o[4]) {  // :89-97
  double r[4] = {0, 0, 0, 0};
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int k = 0; k < 4; ++k) r[i] += a[k] * gm.G[i][k];
#pragma unroll
  for (int i = 0; i < 4; ++i) o[i] = r[i];
}
KZD void lower4(const double a[4], const Geo &gm, double o[4]) {  // :104-112
  double r[4] = {0, 0, 0, 0};
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) r[i] += a[j] * gm.g[i][j];
#pragma unroll
  for (int i = 0; i < 4; ++i) o[i] = r[i];
}
KZD double dot4(const double a[4], const double b[4]) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2] + a[3] * b[3]; }
KZD void lower_second(const double t[4][4], const Geo &gm, double o[4][4]) {  // :147-158
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      double s = 0;
#pragma unroll
      for (int k = 0; k < 4; ++k) s += t[i][k] * gm.g[k][j];
      o[i][j] = s;
    }
}
KZD void lower_second_row(const double t[4][4], const Geo &gm, int row, double o[4]) {  // :165-174
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    double s = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) s += t[row][k] * gm.g[k][j];
    o[j] = s;
  }
}
KZD double ut_from_vel3(const double v[4], const Geo &gm) {  // :200-212
  double b = 0, c = 0;
#pragma unroll
  for (int i = 1; i < 4; ++i) {
    b += v[i] * gm.g[0][i];
#pragma unroll
    for (int j = 1; j < 4; ++j) c += v[i] * v[j] * gm.g[i][j];
  }
  return kzm_sqrt(-1.0 / (gm.g[0][0] + 2.0 * b + c));
}
KZD int ut_from_spatial_ucon(const double u[4], const Geo &gm, double *ut) {  // :228-242
  double a = gm.g[0][0], b = 0, c = 1;
#pragma unroll
  for (int i = 1; i < 4; ++i) {
    b += u[i] * gm.g[0][i];
#pragma unroll
    for (int j = 1; j < 4; ++j) c += u[i] * u[j] * gm.g[i][j];
  }
  double delta = b * b - a * c;
  if (delta < 0) return E_SPACELIKE;
  *ut = (-b - kzm_sqrt(delta)) / a;
  return E_OK;
}
KZD double alpha_gamma(const double u[4], const Geo &gm) {  // :250-260
  double qsq = 0;
#pragma unroll
  for (int i = 1; i < 4; ++i)
#pragma unroll
    for (int j = 1; j < 4; ++j) qsq += u[i] * u[j] * gm.g[i][j];
  double ag2 = (-1.0 / gm.G[0][0]) * (1.0 + qsq);
  return ag2 < 0 ? 1.0 : kzm_sqrt(ag2);
}
KZD void velr_to_vel4(const double u[4], const Geo &gm, double o[4]) {  // :265-273
  double ag = alpha_gamma(u, gm);
  double o0 = -ag * gm.G[0][0], o1 = u[1] - ag * gm.G[0][1], o2 = u[2] - ag * gm.G[0][2], o3 = u[3] - ag * gm.G[0][3];
  o[0] = o0; o[1] = o1; o[2] = o2; o[3] = o3;
}
// relele.zig:282-348 convert, the pairs on the path
KZD int vel4_to_velr(const double u[4], const Geo &gm, bool trust_ut, double o[4]) {
  double ut = u[0];
  if (!trust_ut) {
    int e = ut_from_spatial_ucon(u, gm, &ut);
    if (e) return e;
  }
  double r1 = u[1] - ut * gm.G[0][1] / gm.G[0][0];
  double r2 = u[2] - ut * gm.G[0][2] / gm.G[0][0];
  double r3 = u[3] - ut * gm.G[0][3] / gm.G[0][0];
  o[0] = ut; o[1] = r1; o[2] = r2; o[3] = r3;
  return E_OK;
}
KZD int vel3_to_velr(const double u[4], const Geo &gm, double o[4]) {
  double ut = ut_from_vel3(u, gm);
  if (ut < 1.0 || ut != ut) return E_VELCONV;
  double w[4] = {ut, u[1] * ut, u[2] * ut, u[3] * ut};
#pragma unroll
  for (int i = 1; i < 4; ++i) w[i] = w[i] - w[0] * gm.G[0][i] / gm.G[0][0];
#pragma unroll
  for (int i = 0; i < 4; ++i) o[i] = w[i];
  return E_OK;
}
KZD void ucon_ucov(const double *pp, const Geo &gm, double ucon[4], double ucov[4]) {  // :375-377
  double v[4] = {0, pp[VX], pp[VY], pp[VZ]};
  velr_to_vel4(v, gm, ucon);
  lower4(ucon, gm, ucov);
}

// ---- bfield.zig / hydro.zig ---------------------------------------------
KZD void bcon_bcov_bsq(double B0, double B1_, double B2_, const double ucon[4], const double ucov[4], const Geo &gm,
                       double bcon[4], double bcov[4], double *bsq) {  // bfield.zig:23-48
  double b[4];
  b[0] = B0 * ucov[1] + B1_ * ucov[2] + B2_ * ucov[3];
  double inv_ut = 1.0 / ucon[0];
  b[1] = (B0 + b[0] * ucon[1]) * inv_ut;
  b[2] = (B1_ + b[0] * ucon[2]) * inv_ut;
  b[3] = (B2_ + b[0] * ucon[3]) * inv_ut;
#pragma unroll
  for (int j = 0; j < 4; ++j) bcon[j] = b[j];
  lower4(b, gm, bcov);
  *bsq = dot4(bcon, bcov);
}
KZD double s_from_u(double rho, double u, double gam) {  // hydro.zig:32-37
  double gammam1 = gam - 1.0;
  double indexn = 1.0 / gammam1;
  double pre = gammam1 * u;
  return rho * kzm_log(kzm_pow(pre, indexn) / kzm_pow(rho, indexn + 1.0));
}

// ---- p2u.zig ---------------------------------------------------------------
KZD double calc_utp1(const double v[4], const Geo &gm) {  // :17-36
  double qsq = 0;
#pragma unroll
  for (int i = 1; i < 4; ++i)
#pragma unroll
    for (int j = 1; j < 4; ++j) qsq += v[i] * v[j] * gm.g[i][j];
  double alphasq = gm.alpha * gm.alpha;
  double alpgam = kzm_sqrt(alphasq * (1.0 + qsq));
  double bsa = gm.g[0][1] * gm.G[0][1] + gm.g[0][2] * gm.G[0][2] + gm.g[0][3] * gm.G[0][3];
  double ud0 = 0;
#pragma unroll
  for (int j = 1; j < 4; ++j) ud0 += v[j] * gm.g[0][j];
  return ud0 + (gm.gttpert - alphasq * (bsa + qsq)) / (1.0 + alpgam);
}
KZD void p2u_mhd(const double *pp, double *uu, const Geo &gm, double gam) {  // :39-100
  double gdetu = gm.gdet;
  double rho = pp[RHO], ugas = pp[UU], s = pp[ENTR];
  double vcon[4] = {0, pp[VX], pp[VY], pp[VZ]};
  double ucon[4], ucov[4];
  velr_to_vel4(vcon, gm, ucon);
  lower4(ucon, gm, ucov);
  double bcon[4], bcov[4], bsq;
  bcon_bcov_bsq(pp[B1], pp[B2], pp[B3], ucon, ucov, gm, bcon, bcov, &bsq);
  double ut = ucon[0];
  double rhout = rho * ut, sut = s * ut;
  double pre = (gam - 1.0) * ugas;
  double w = rho + ugas + pre;
  double eta = w + bsq;
  double etap = ugas + pre + bsq;
  double ptot = pre + 0.5 * bsq;
  double utp1 = calc_utp1(vcon, gm);
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
KZD void p2u_rad(const double *pp, double *uu, const Geo &gm) {  // :103-132
  double erf = pp[EE];
  double v[4] = {0, pp[FX], pp[FY], pp[FZ]};
  double urf[4], r[4];
  velr_to_vel4(v, gm, urf);
#pragma unroll
  for (int i = 0; i < 4; ++i) r[i] = four_third * erf * urf[0] * urf[i] + (one_third)*erf * gm.G[0][i];
  lower4(r, gm, r);
#pragma unroll
  for (int i = 0; i < 4; ++i) uu[EE + i] = gm.gdet * r[i];
}
KZD void p2u(const double *pp, double *uu, const Geo &gm, double gam) {  // :135-148
  p2u_mhd(pp, uu, gm, gam);
  p2u_rad(pp, uu, gm);
}

// ---- thermo.zig / radiation.zig -----------------------------------------
KZD double t_from_urho(const kz_opac_params &c, double uint, double rho, double gam) {  // thermo.zig:166-169
  double p = uint * (gam - 1.0);
  return c.mugas_mp_over_kb * p / rho;
}
KZD double lte_t_from_e(const kz_opac_params &c, double e) {  // thermo.zig:155-157
  return kzm_sqrt(kzm_sqrt(c.one_over_four_sigmarad * e));
}
KZD void temps_from_urho(const kz_opac_params &c, double uint, double rho, double gam, double *tg, double *te) {
  double t = t_from_urho(c, uint, rho, gam);  // thermo.zig:191-201
  double e = kzm_isfinite(t) ? t : (1.0 / 1.0e-80);
  e = e < 1.0e2 ? 1.0e2 : e;
  *tg = t;
  *te = e;
}
KZD void calc_rij(const double *pp, const Geo &gm, double rij[4][4]) {  // radiation.zig:79-98
  double erf = pp[EE];
  double v[4] = {0, pp[FX], pp[FY], pp[FZ]};
  double urf[4];
  velr_to_vel4(v, gm, urf);
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) rij[i][j] = four_third * erf * urf[i] * urf[j] + one_third * erf * gm.G[i][j];
}
KZD double calc_ff_rtt(const double *pp, const Geo &gm) {  // radiation.zig:104-124
  double ucon[4], ucov[4];
  ucon_ucov(pp, gm, ucon, ucov);
  double rij[4][4], rud[4][4];
  calc_rij(pp, gm, rij);
  lower_second(rij, gm, rud);
  double rtt = 0;
#pragma unroll
  for (int ia = 0; ia < 4; ++ia)
#pragma unroll
    for (int ib = 0; ib < 4; ++ib) rtt += -rud[ia][ib] * ucon[ib] * ucov[ia];
  return rtt;
}

// ---- wavespeeds.zig -------------------------------------------------------
KZD void lr_core(const double ucon[4], const Geo &gm, const double ws2[3], const int active[3], double aret[6]) {
#pragma unroll
  for (int i = 0; i < 6; ++i) aret[i] = 0;  // :32-81
  double Bcov[4] = {1, 0, 0, 0}, Bcon[4];
  raise4(Bcov, gm, Bcon);
  double Bsq = dot4(Bcon, Bcov);
  double Bu = dot4(Bcov, ucon);
  double Bu2 = Bu * Bu;
#pragma unroll
  for (int dim = 0; dim < 3; ++dim) {
    if (!active[dim]) continue;
    double w2 = ws2[dim];
    double Acov[4] = {0, 0, 0, 0}, Acon[4];
    Acov[dim + 1] = 1;
    raise4(Acov, gm, Acon);
    double Asq = dot4(Acon, Acov), Au = dot4(Acov, ucon), AB = dot4(Acon, Bcov);
    double Au2 = Au * Au, AuBu = Au * Bu;
    double b = -2.0 * (AuBu * (1.0 - w2) - AB * w2);
    double a = Bu2 * (1.0 - w2) - Bsq * w2;
    double discr = 4.0 * w2 * ((AB * AB - Asq * Bsq) * w2 + (2.0 * AB * Au * Bu - Asq * Bu2 - Bsq * Au2) * (w2 - 1.0));
    discr = kzm_sqrt(discr);
    double c1 = (-b + discr) / (2.0 * a), c2 = (-b - discr) / (2.0 * a);
    if (c2 > c1) {
      aret[2 * dim] = c1;
      aret[2 * dim + 1] = c2;
    } else {
      aret[2 * dim] = c2;
      aret[2 * dim + 1] = c1;
    }
  }
}
KZD void gas_wavespeeds(const double *pp, const Geo &gm, double gam, const int active[3], double a[6]) {
  double ucon[4], ucov[4];  // :85-132
  ucon_ucov(pp, gm, ucon, ucov);
  double bcon[4], bcov[4], bsq;
  bcon_bcov_bsq(pp[B1], pp[B2], pp[B3], ucon, ucov, gm, bcon, bcov, &bsq);
  double rho = pp[RHO], uu = pp[UU];
  double pre = (gam - 1.0) * uu;
  double cs2 = gam * pre / (rho + uu + pre);
  if (cs2 > 0.95) cs2 = 0.95;
  double va2 = bsq / (bsq + (rho + uu + pre));
  if (va2 < 0.0) va2 = 0.0;
  double vt2 = cs2 + va2 - cs2 * va2;
  double ws[3] = {vt2, vt2, vt2};
  lr_core(ucon, gm, ws, active, a);
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    if (a[2 * d] > 0.0) a[2 * d] = 0.0;
    if (a[2 * d + 1] < 0.0) a[2 * d + 1] = 0.0;
  }
}
KZD void rad_wavespeeds(const double *pp, const Geo &gm, const double tau[3], const int active[3], double av[12]) {
  double v[4] = {0, pp[FX], pp[FY], pp[FZ]}, urf[4];  // radiation.zig:177-211
  velr_to_vel4(v, gm, urf);
  double ws0[3] = {one_third, one_third, one_third};
  lr_core(urf, gm, ws0, active, av);
  double r2[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    if (tau[d] <= 0.0) r2[d] = one_third;
    else r2[d] = kzm_fmin(one_third, (four_third * four_third) / (tau[d] * tau[d]));
  }
  double ws1[3] = {r2[0], r2[1], r2[1]};
  lr_core(urf, gm, ws1, active, av + 6);
}

// ---- opacities.zig ---------------------------------------------------------
struct Opac {
  double gas_abs, rad_abs, rad_ross;
};
// opacities.zig:131-315 at the .residual / .chi levels (the channels the
// step consumes); `want_rad` false is the .chi level.
KZD Opac calc_opacities(const kz_opac_params &c, double rho, double te, double trad, double ne, double bsq,
                        bool want_rad) {
  double rt_te = kzm_sqrt(te);
  double rhocgs = c.rhogu2cgs * rho;
  double nethcgs = ne * c.numdensgu2cgs;
  double bsqcgs = c.fourpi * (bsq * c.enden_gu2cgs_factor);
  double bmagcgs = kzm_sqrt(bsqcgs);
  double bbenergy = (4.0 * c.sigma_rad_cgs) * te * te * te * te;
  double kgff = 0, krff = 0, krffross = 0, kgsyn = 0, krsyn = 0, krsynross = 0;
  if (c.bremsstrahlung) {
    double mfrac = 1.0 - c.hfrac - c.hefrac;
    double emis_ff = 1.4e-27 * nethcgs * ((c.hfrac + c.hefrac + c.z2diva_mean * mfrac) / c.m_proton_cgs) * rt_te *
                     1.2 * (1.0 + 4.4e-10 * te);
    kgff = c.kappacgs2gu * (emis_ff / bbenergy) * rho;
    if (want_rad) {
      double zeta = trad / te;
      double zi = 1.0 / zeta;
      double zi3 = zi * zi * zi;
      krff = kgff * kzm_log1p(1.6 * zeta) * c.one_over_log_2p6 * zi3;
      double zr5 = kzm_pow(zeta, 0.2);
      double zr5i4 = 1.0 / (zr5 * zr5 * zr5 * zr5);
      double zr5i3 = zr5 * zr5i4;
      double sc = 14.12 / (432.7 - 106.8 * zr5i3 + 43.17 * zr5i4 + 57.88 * zi);
      krffross = kgff * sc * zi3;
    }
  }
  if (c.synchrotron) {
    // opacities.zig:208-297: with the bridge the synchrotron Trad is clamped
    // up to TradBB^(4/3) / Te^(1/3) (TradBB = trad here: no photon-number
    // evolution), non-relativistic components are added to the gas and
    // radiation absorption, and the Terelfactor suppression is dropped.  The
    // reference adds the gas term after the Trad block; nothing there reads
    // or writes kgsyn, so adding it here gives the same bits and ends the
    // live range of its inputs before that block.
    const bool bridge = c.synchrotron_bridge != 0;
    double emis_syn = 3.61e-34 * (nethcgs / rhocgs) * te * te * bmagcgs * bmagcgs;
    kgsyn = c.kappacgs2gu * (emis_syn / bbenergy) * rho;
    if (bridge)
      kgsyn += c.kappacgs2gu * (2.35869e-21 * (nethcgs / rhocgs) * (bmagcgs * bmagcgs) / (te * te * te)) * rho;
    if (want_rad) {
      double trad_syn = trad;
      if (bridge) {
        const double trad_lim = kzm_pow(trad, 4.0 / 3.0) / kzm_pow(te, 1.0 / 3.0);
        trad_syn = trad < trad_lim ? trad_lim : trad;
      }
      double nu_mb = 1.19e-13 * te * te;
      double zbr = c.k_boltz_cgs * trad_syn / c.h_cgs / nu_mb;
      double den = 1.79 * kzm_cbrt(zbr * zbr * zbr * zbr * zbr * bmagcgs * bmagcgs * bmagcgs * bmagcgs) +
                   1.35 * kzm_cbrt(zbr * zbr * zbr * zbr * zbr * zbr * zbr * bmagcgs * bmagcgs) +
                   0.248 * zbr * zbr * zbr;
      double iab = bmagcgs * bmagcgs / den;
      double te5 = te * te * te * te * te;
      krsyn = c.kappacgs2gu * (2.13e39 * (nethcgs / rhocgs) / te5 * iab) * rho;
      if (bridge)
        krsyn += c.kappacgs2gu *
                 (2.35869e-21 * (nethcgs / rhocgs) * (bmagcgs * bmagcgs) / (trad_syn * trad_syn * trad_syn)) * rho;
      // two powers of each base: one logarithm per base (kzm_pow_from_log
      // returns kzm_pow's bits for a finite positive base)
      double bmb, pb69, pz169, pz463;
      if (bmagcgs > 0.0 && kzm_isfinite(bmagcgs)) {
        double lh, ll;
        kzm_log_dd(bmagcgs, &lh, &ll);
        bmb = kzm_pow_from_log(lh, ll, -0.463);
        pb69 = kzm_pow_from_log(lh, ll, 0.69);
      } else {
        bmb = bmagcgs > 0.0 ? kzm_pow(bmagcgs, -0.463) : 0.0;
        pb69 = kzm_pow(bmagcgs, 0.69);
      }
      if (zbr > 0.0 && kzm_isfinite(zbr)) {
        double lh, ll;
        kzm_log_dd(zbr, &lh, &ll);
        pz169 = kzm_pow_from_log(lh, ll, 1.69);
        pz463 = kzm_pow_from_log(lh, ll, 0.463);
      } else {
        pz169 = kzm_pow(zbr, 1.69);
        pz463 = kzm_pow(zbr, 0.463);
      }
      double iabr = 0.13 * pb69 / (pz169 * kzm_exp(1.6 * pz463 * bmb));
      krsynross = c.kappacgs2gu * (2.13e39 * (nethcgs / rhocgs) / te5 * iabr) * rho;
    }
    if (!bridge) {
      double terel = te * c.k_over_mecsq;
      double tf = (terel * terel) / (1.0 + terel * terel);
      kgsyn *= tf;
      if (want_rad) {
        krsyn *= tf;
        krsynross *= tf;
      }
    }
  }
  Opac o;
  o.gas_abs = kgff + kgsyn;
  o.rad_abs = krff + krsyn;
  o.rad_ross = krffross + krsynross;
  return o;
}
KZD double kappaes_at(const kz_opac_params &c, double rho, double trad) {  // radforce.zig:140-151
  if (c.kappaes_mode == 1) {  // opacities.zig:352-360
    double corr = 1.0;
    if (c.kleinnishina) corr = 1.0 / (1.0 + kzm_pow(trad / 4.5e8, 0.86));
    return ((0.2 * (1.0 + c.hfrac)) * corr) * c.kappa_cgs2gu_factor * rho;
  }
  if (c.kappaes_mode == 2) return c.kappaes_grey * rho;
  return 0.0;
}

// ---- radforce.zig -----------------------------------------------------------
// the fields of RadState the implicit residual reads (radforce.zig:156-178)
struct RadState {
  double tgas, te, sgas, uint;
  double ucon[4], ucov[4];
  double rij[4][4];
  double ehat, trad, kappaes;
  Opac op;
};
// radforce.zig:241-351 fillRadStateCoreG; `full` computes sgas
KZD void fill_radstate(const double *pp, const Geo &gm, double gam, const kz_opac_params &c, bool full,
                       RadState &st) {
  double rho = pp[RHO], uint = pp[UU];
  double tgas, te;
  temps_from_urho(c, uint, rho, gam, &tgas, &te);
  st.sgas = full ? c.kb_over_mugas_mp * s_from_u(rho, uint, gam) : 0.0;
  double ne = c.one_over_mue_mp * rho;
  ucon_ucov(pp, gm, st.ucon, st.ucov);
  double bcon[4], bcov[4], bsq;
  bcon_bcov_bsq(pp[B1], pp[B2], pp[B3], st.ucon, st.ucov, gm, bcon, bcov, &bsq);
  calc_rij(pp, gm, st.rij);
  double ehat = 0;
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) ehat += st.rij[i][j] * st.ucov[i] * st.ucov[j];
  double trad = lte_t_from_e(c, ehat);
  st.kappaes = kappaes_at(c, rho, trad);
  if (c.kappa_mode == 0) {
    st.op = calc_opacities(c, rho, te, trad, ne, bsq, true);
  } else {
    double kap = c.kappa_grey * rho;
    st.op.gas_abs = kap; st.op.rad_abs = kap; st.op.rad_ross = kap;
  }
  st.tgas = tgas; st.te = te; st.uint = uint; st.ehat = ehat; st.trad = trad;
}
// radforce.zig:432-494 calcGiFromStateCoreG (slim: ff[0] and lab only)
KZD void calc_gi(const RadState &st, const kz_opac_params &c, double opdamp, double *giff0, double gilab[4]) {
  double b = c.sigma_rad_over_pi * st.te * st.te * st.te * st.te;
  double kra = st.op.rad_abs, kga = st.op.gas_abs, krr = st.op.rad_ross, kes = st.kappaes, ruu = st.ehat;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    double ru = 0;
#pragma unroll
    for (int j = 0; j < 4; ++j) ru += st.rij[i][j] * st.ucov[j];
    gilab[i] = -(krr + kes) * ru - ((krr + kes - kra) * ruu + kga * c.fourpi * b) * st.ucon[i];
  }
  double ff0 = -kga * c.fourpi * b + kra * st.ehat;
  if (c.comptonization) {
    double th = c.kb_over_me * st.te;  // radforce.zig:370-375
    double coeff = st.kappaes * st.ehat * ((4.0 * c.kb_over_me) * (st.trad - st.te)) *
                   (1.0 + 3.683 * th + 4.0 * th * th) / (1.0 + 4.0 * th);
#pragma unroll
    for (int i = 0; i < 4; ++i) gilab[i] += coeff * st.ucon[i];
    ff0 += coeff;
  }
  if (opdamp != 1.0) {
    ff0 *= opdamp;
#pragma unroll
    for (int i = 0; i < 4; ++i) gilab[i] *= opdamp;
  }
  *giff0 = ff0;
}
// radforce.zig:562-609 calcChiSlim
KZD double calc_chi_slim(const double *pp, const Geo &gm, double gam, const kz_opac_params &c) {
  double rho = pp[RHO], uint = pp[UU];
  double tgas, te;
  temps_from_urho(c, uint, rho, gam, &tgas, &te);
  double ne = c.one_over_mue_mp * rho;
  double ucon[4], ucov[4];
  ucon_ucov(pp, gm, ucon, ucov);
  double bcon[4], bcov[4], bsq;
  bcon_bcov_bsq(pp[B1], pp[B2], pp[B3], ucon, ucov, gm, bcon, bcov, &bsq);
  double kappa = c.kappa_mode == 0 ? calc_opacities(c, rho, te, te, ne, bsq, false).gas_abs : c.kappa_grey * rho;
  return kappa + kappaes_at(c, rho, te);
}

// ---- fv/recon.zig (PPM, :99-195) and laxf.zig ------------------------------
KZD double sgn(double x) { return x >= 0.0 ? 1.0 : -1.0; }
KZD void ppm(const double u[5], const double dx[5], double *lo, double *ro) {
  double dxm2 = dx[0], dxm1 = dx[1], dx0 = dx[2], dxp1 = dx[3], dxp2 = dx[4];
  double a1 = dxp2 + dxp1, a1i = 1.0 / a1;
  double a2 = dxp1 + dx0, a2i = 1.0 / a2;
  double a3 = dx0 + dxm1, a3i = 1.0 / a3;
  double a4 = dxm1 + dxm2, a4i = 1.0 / a4;
  double s1i = 1.0 / (dxm1 + dx0 + dxp1), s2i = 1.0 / (dx0 + dxp1 + dxp2), s3i = 1.0 / (dxm2 + dxm1 + dx0);
  double b1 = dx0 + 2.0 * dxm1, b2 = dxp1 + 2.0 * dx0, b3 = dxm1 + 2.0 * dxm2;
  double b4 = 2.0 * dxp1 + dx0, b5 = 2.0 * dxp2 + dxp1, b6 = 2.0 * dx0 + dxm1;
  double dri = dx0 * s1i * (b1 * a2i * (u[3] - u[2]) + b4 * a3i * (u[2] - u[1]));
  double drip1 = dxp1 * s2i * (b2 * a1i * (u[4] - u[3]) + b5 * a2i * (u[3] - u[2]));
  double drim1 = dxm1 * s3i * (b3 * a3i * (u[2] - u[1]) + b6 * a4i * (u[1] - u[0]));
  dri = ((u[3] - u[2]) * (u[2] - u[1]) > 0.0)
            ? kzm_fmin(kzm_fabs(dri), kzm_fmin(2.0 * kzm_fabs(u[2] - u[1]), 2.0 * kzm_fabs(u[2] - u[3]))) * sgn(dri)
            : 0.0;
  drip1 = ((u[4] - u[3]) * (u[3] - u[2]) > 0.0)
              ? kzm_fmin(kzm_fabs(drip1), kzm_fmin(2.0 * kzm_fabs(u[3] - u[2]), 2.0 * kzm_fabs(u[3] - u[4]))) * sgn(drip1)
              : 0.0;
  drim1 = ((u[2] - u[1]) * (u[1] - u[0]) > 0.0)
              ? kzm_fmin(kzm_fabs(drim1), kzm_fmin(2.0 * kzm_fabs(u[1] - u[0]), 2.0 * kzm_fabs(u[1] - u[2]))) * sgn(drim1)
              : 0.0;
  double z1 = a3 / b2, z2 = a1 / b4;
  double dxi = 1.0 / (dxm1 + dx0 + dxp1 + dxp2);
  double right = u[2] + dx0 * a2i * (u[3] - u[2]) +
                 dxi * ((2.0 * dxp1 * dx0) * a2i * (z1 - z2) * (u[3] - u[2]) - dx0 * z1 * drip1 + dxp1 * z2 * dri);
  z1 = a4 / b1;
  z2 = a2 / b6;
  dxi = 1.0 / (dxm2 + dxm1 + dx0 + dxp1);
  double left = u[1] + dxm1 * a3i * (u[2] - u[1]) +
                dxi * ((2.0 * dx0 * dxm1) * a3i * (z1 - z2) * (u[2] - u[1]) - dxm1 * z1 * dri + dx0 * z2 * drim1);
  if ((right - u[2]) * (u[2] - left) <= 0.0) {
    left = u[2];
    right = u[2];
  }
  if ((right - left) * (left - (3.0 * u[2] - 2.0 * right)) < 0.0) left = 3.0 * u[2] - 2.0 * right;
  if ((right - left) * ((3.0 * u[2] - 2.0 * left) - right) < 0.0) right = 3.0 * u[2] - 2.0 * left;
  *lo = left;
  *ro = right;
}
// recon.zig:77-91 linear reconstruction with the generalized minmod limiter
KZD void minmod_linear(const double u[5], double theta, double *lo, double *ro) {
  double dm = u[2] - u[1], dp = u[3] - u[2];
  double avg = 0.5 * (dm + dp);
  double lo_s = kzm_fmin(kzm_fmin(theta * dm, avg), theta * dp);
  double hi_s = kzm_fmax(kzm_fmax(theta * dm, avg), theta * dp);
  double slope = (dm * dp <= 0.0) ? 0.0 : (dm > 0.0 ? lo_s : hi_s);
  *lo = u[2] - 0.5 * slope;
  *ro = u[2] + 0.5 * slope;
}
// recon.zig:47-53 reconstruct: order 0 donor cell, 1 linear minmod, 2 PPM
KZD void reconstruct(int order, double theta, const double u[5], const double dx[5], double *lo, double *ro) {
  if (order <= 0) {
    *lo = u[2];
    *ro = u[2];
  } else if (order == 1) {
    minmod_linear(u, theta, lo, ro);
  } else {
    ppm(u, dx, lo, ro);
  }
}
KZD double laxf(double fl, double fr, double ul, double ur, double ag) { return 0.5 * (fr + fl - ag * (ur - ul)); }

// ---- physics/flux.zig:40-134 fFluxPrime -------------------------------------
// IDIM is a template parameter so that the row index d = IDIM + 1 into the
// 4-vectors and T^mu_nu is a compile-time constant (registers, not scratch).
template <int IDIM>
KZD int flux_prime_d(const double *pp, const Geo &gm, double gam, double *ff) {
  constexpr int idim = IDIM;
  double gdetu = gm.gdet;
  double rho = pp[RHO], ugas = pp[UU], s = pp[ENTR];
  double vcon[4] = {0, pp[VX], pp[VY], pp[VZ]};
  double ucon[4], ucov[4];
  velr_to_vel4(vcon, gm, ucon);
  lower4(ucon, gm, ucov);
  double bcon[4], bcov[4], bsq;
  bcon_bcov_bsq(pp[B1], pp[B2], pp[B3], ucon, ucov, gm, bcon, bcov, &bsq);
  double p0 = (gam - 1.0) * ugas;
  double eta = rho + ugas + p0 + bsq;
  double ptot = p0 + 0.5 * bsq;
  int d = idim + 1;
  double tij[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) tij[i][j] = eta * ucon[i] * ucon[j] + ptot * gm.G[i][j] - bcon[i] * bcon[j];
  double trow[4];
  lower_second_row(tij, gm, d, trow);
  double etap = ugas + p0 + bsq;
  double utp1 = calc_utp1(vcon, gm);
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
  calc_rij(pp, gm, rij);
  lower_second_row(rij, gm, d, rrow);
#pragma unroll
  for (int nu = 0; nu < 4; ++nu) ff[EE + nu] = gdetu * rrow[nu];
  int bad = 0;
#pragma unroll
  for (int i = 0; i < NV; ++i) bad |= !kzm_isfinite(ff[i]);
  return bad ? E_NANFLUX : E_OK;
}
KZD int flux_prime(const double *pp, int idim, const Geo &gm, double gam, double *ff) {
  if (idim == 0) return flux_prime_d<0>(pp, gm, gam, ff);
  if (idim == 1) return flux_prime_d<1>(pp, gm, gam, ff);
  return flux_prime_d<2>(pp, gm, gam, ff);
}

}  // namespace kzd
