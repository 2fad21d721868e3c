/* kzo_solve.c — oracle restatement of koral/solve/{invert,invert_rad,
 * implicit}.zig.  TEST INFRASTRUCTURE (see kzo.h).
 */
#include <string.h>
#include "kzo.h"

static const double U2PCONV = 1.0e-10; /* invert.zig:29 */
static const double BIGW = 1.0 / 1.0e-80; /* invert.zig:31 */

typedef struct { double f, df, err; } resid_t;
typedef struct { double v2, gamma, gamma2, rho0, wmrho0; } kin_t;

/* invert.zig:94-113 wKinematics */
static kin_t w_kinematics(double W, double Wp, double D, double Qtsq, double QdotBsq, double Bsq) {
  kin_t k;
  double x = Bsq + W;
  double wsq = W * W;
  k.v2 = (wsq * Qtsq + QdotBsq * (Bsq + 2.0 * W)) / (wsq * (x * x));
  k.gamma2 = 1.0 / (1.0 - k.v2);
  k.gamma = kzm_sqrt(k.gamma2);
  k.rho0 = D / k.gamma;
  k.wmrho0 = Wp / k.gamma2 - D * k.v2 / (1.0 + k.gamma);
  return k;
}

/* invert.zig:116-158 fU2pHot */
static resid_t f_u2p_hot(double Wp, const double cons[7], double pgamma) {
  double Qn = cons[0], Qtsq = cons[1], D = cons[2], QdotBsq = cons[3], Bsq = cons[4];
  double W = Wp + D;
  double x = Bsq + W;
  double wsq = W * W;
  double w3 = wsq * W;
  double x3 = x * x * x;
  kin_t k = w_kinematics(W, Wp, D, Qtsq, QdotBsq, Bsq);
  double u = k.wmrho0 / pgamma;
  double p = (pgamma - 1.0) * u;
  resid_t r;
  r.f = Qn + W - p + 0.5 * Bsq * (1.0 + k.v2) - QdotBsq / 2.0 / wsq;
  r.err = kzm_fabs(r.f) / (kzm_fabs(Qn) + kzm_fabs(W) + kzm_fabs(p) + kzm_fabs(0.5 * Bsq * (1.0 + k.v2)) +
                           kzm_fabs(QdotBsq / 2.0 / wsq));
  double dvsq = -2.0 / x3 * (Qtsq + QdotBsq * (3.0 * W * x + Bsq * Bsq) / w3);
  double dp1 = (pgamma - 1.0) * (1.0 - k.v2) / pgamma;
  double idwmrho0dp = (pgamma - 1.0) / pgamma;
  double dwmrho0dvsq = D * (k.gamma * 0.5 - 1.0) - Wp;
  double dp2 = dwmrho0dvsq * idwmrho0dp;
  double dpdW = dp1 + dp2 * dvsq;
  r.df = 1.0 - dpdW + QdotBsq / (wsq * W) + 0.5 * Bsq * dvsq;
  return r;
}

/* invert.zig:161-210 fU2pEntropy */
static resid_t f_u2p_entropy(double Wp, const double cons[7], double pgamma) {
  double Qtsq = cons[1], D = cons[2], QdotBsq = cons[3], Bsq = cons[4], Sc = cons[5];
  double W = Wp + D;
  double x = Bsq + W;
  double wsq = W * W;
  double w3 = wsq * W;
  double x3 = x * x * x;
  kin_t k = w_kinematics(W, Wp, D, Qtsq, QdotBsq, Bsq);
  double pressure = (pgamma - 1.0) / pgamma * k.wmrho0;
  double indexn = 1.0 / (pgamma - 1.0);
  double ssofchi = kzm_log(kzm_pow(pressure, indexn) / kzm_pow(k.rho0, indexn + 1.0));
  resid_t r;
  r.f = -Sc + D * ssofchi;
  r.err = kzm_fabs(r.f) / (kzm_fabs(Sc) + kzm_fabs(D * ssofchi));
  double dssdrho = pgamma / ((1.0 - pgamma) * k.rho0);
  double dssdchi = 1.0 / ((pgamma - 1.0) * k.wmrho0);
  double dwmrho0dw = 1.0 / (k.gamma * k.gamma);
  double dwmrho0dvsq = D * (k.gamma * 0.5 - 1.0) - Wp;
  double drho0dvsq = -D * k.gamma * 0.5;
  double dvsq = -2.0 / x3 * (Qtsq + QdotBsq * (3.0 * W * x + Bsq * Bsq) / w3);
  double dssdw = dwmrho0dw * dssdchi;
  double dssdvsq = drho0dvsq * dssdrho + dwmrho0dvsq * dssdchi;
  double dssdwp = dssdw + dssdvsq * dvsq;
  r.df = D * dssdwp;
  return r;
}

static resid_t u2p_residual(int etype, double Wp, const double cons[7], double pgamma) {
  return etype == 0 ? f_u2p_hot(Wp, cons, pgamma) : f_u2p_entropy(Wp, cons, pgamma);
}

/* invert.zig:221-224 wOutOfBounds */
static int w_out_of_bounds(double W, double Qtsq, double QdotBsq, double Bsq, resid_t res) {
  return (W * W * W * (W + 2.0 * Bsq) - QdotBsq * (2.0 * W + Bsq)) <= W * W * (Qtsq - Bsq * Bsq) ||
         !kzm_isfinite(res.f) || !kzm_isfinite(res.df);
}

/* invert.zig:235-404 u2pSolverW; etype 0 = hot, 1 = entropy */
int kzo_u2p_solver_w(const double *uu, double *pp, const kzo_geo *gm, int etype, double pgamma) {
  double gdetu_inv = 1.0 / gm->gdet;
  double alpha = gm->alpha;
  double D = uu[RHO] * gdetu_inv * alpha;
  double Sc = uu[ENTR] * gdetu_inv * alpha;
  double Qcov[4] = {(uu[UU] * gdetu_inv - uu[RHO] * gdetu_inv) * alpha, uu[VX] * gdetu_inv * alpha,
                    uu[VY] * gdetu_inv * alpha, uu[VZ] * gdetu_inv * alpha};
  double Qcon[4];
  kzo_raise(Qcov, gm, Qcon);
  double Bcon[4] = {0, uu[B1] * gdetu_inv * alpha, uu[B2] * gdetu_inv * alpha, uu[B3] * gdetu_inv * alpha};
  double Bcov[4];
  kzo_lower(Bcon, gm, Bcov);
  double Bsq = kzo_dot(Bcon, Bcov);
  double QdotB = kzo_dot(Qcov, Bcon);
  double QdotBsq = QdotB * QdotB;
  double ncov[4] = {-alpha, 0, 0, 0};
  double ncon[4];
  kzo_raise(ncov, gm, ncon);
  double Qn = Qcon[0] * ncov[0];
  double Qtcon[4];
  for (int i = 0; i < 4; ++i) {
    Qtcon[i] = 0;
    for (int j = 0; j < 4; ++j) Qtcon[i] += ((i == j ? 1.0 : 0.0) + ncon[i] * ncov[j]) * Qcon[j];
  }
  double Qtcov[4];
  kzo_lower(Qtcon, gm, Qtcov);
  double Qt2 = kzo_dot(Qtcon, Qtcov);
  double Qtsq = Qt2;

  double rho = pp[RHO];
  double uint = pp[UU];
  double utcon0[4] = {0, pp[VX], pp[VY], pp[VZ]};
  double qsq = 0;
  for (int i = 1; i < 4; ++i)
    for (int j = 1; j < 4; ++j) qsq += utcon0[i] * utcon0[j] * gm->g[i][j];
  double gamma2 = 1.0 + qsq;
  double W = (rho + pgamma * uint) * gamma2;

  double cons[7] = {Qn, Qt2, D, QdotBsq, Bsq, Sc, 0};
  double conv = U2PCONV;

  int i_increase = 0;
  for (;;) {
    resid_t res = u2p_residual(etype, W - D, cons, pgamma);
    if (w_out_of_bounds(W, Qtsq, QdotBsq, Bsq, res) && i_increase < 50) {
      W *= 10.0;
      i_increase += 1;
      continue;
    }
    break;
  }
  if (i_increase >= 50) return -150;

  int iter = 0;
  double f0, dfdW, errr;
  while (iter < 50) {
    double Wprev = W;
    iter += 1;
    resid_t res = u2p_residual(etype, W - D, cons, pgamma);
    f0 = res.f;
    dfdW = res.df;
    errr = res.err;
    if (errr < conv) break;
    if (dfdW == 0.0) {
      W *= 1.1;
      continue;
    }
    double Wnew = W - f0 / dfdW;
    int idump = 0;
    double dumpfac = 1.0;
    for (;;) {
      resid_t rtmp = u2p_residual(etype, Wnew - D, cons, pgamma);
      if (w_out_of_bounds(Wnew, Qtsq, QdotBsq, Bsq, rtmp) && idump < 100) {
        idump += 1;
        dumpfac /= 2.0;
        Wnew = W - dumpfac * f0 / dfdW;
        continue;
      }
      break;
    }
    if (idump >= 100) return -101;
    W = Wnew;
    if (kzm_fabs(W) > BIGW) return -103;
    if (kzm_fabs((W - Wprev) / Wprev) < conv && errr < 1.0e-1) break;
  }
  if (iter >= 50) return -102;
  if (!kzm_isfinite(W)) return -103;

  double wsq = W * W;
  double xsq = (Bsq + W) * (Bsq + W);
  double v2 = (wsq * Qtsq + QdotBsq * (Bsq + 2.0 * W)) / (wsq * xsq);
  gamma2 = 1.0 / (1.0 - v2);
  double gamma = kzm_sqrt(gamma2);
  rho = D / gamma;
  double entr = Sc / gamma;
  uint = 1.0 / pgamma * (W / gamma2 - rho);
  double utcon[4];
  utcon[0] = 0;
  for (int i = 1; i < 4; ++i) utcon[i] = gamma / (W + Bsq) * (Qtcon[i] + QdotB * Bcon[i] / W);

  if (!kzm_isfinite(utcon[1])) return -120;
  if (uint < 0.0 || gamma2 < 0.0 || W != W || !kzm_isfinite(W)) return -104;

  pp[RHO] = rho;
  pp[UU] = uint;
  pp[VX] = utcon[1];
  pp[VY] = utcon[2];
  pp[VZ] = utcon[3];
  if (rho < 0.0) return -105;
  pp[ENTR] = entr;
  pp[B1] = uu[B1] * gdetu_inv;
  pp[B2] = uu[B2] * gdetu_inv;
  pp[B3] = uu[B3] * gdetu_inv;
  return 0;
}

/* invert.zig:422-467 u2pMhd */
kz_u2p_result kzo_u2p_mhd(const double *uu_in, double *pp, const kzo_geo *gm, double gam,
                          const kz_floor_params *f) {
  double uu[NV], ppbak[NV], ppmhd[NV];
  for (int i = 0; i < NV; ++i) {
    uu[i] = uu_in[i];
    ppbak[i] = pp[i];
    ppmhd[i] = pp[i];
  }
  double gdetu = gm->gdet;
  int ret = 0, entropy_used = 0;
  if (uu[RHO] * gm->alpha / gdetu < 0.0) {
    ppmhd[RHO] = f->rhofloor;
    uu[RHO] = f->rhofloor * gdetu / gm->alpha;
    ret = -2;
  }
  int u2pret = kzo_u2p_solver_w(uu, ppmhd, gm, 0, gam);
  if (u2pret < 0) {
    ret = -1;
    entropy_used = 1;
    u2pret = kzo_u2p_solver_w(uu, ppmhd, gm, 1, gam);
    if (u2pret < 0) ret = -2;
  }
  if (u2pret >= 0) {
    for (int i = 0; i < NV; ++i) pp[i] = ppmhd[i];
  } else {
    for (int i = 0; i < NV; ++i) pp[i] = ppbak[i];
    ret = -3;
  }
  kz_u2p_result r;
  r.ret = ret;
  r.entropy_used = entropy_used;
  r.fixup = ret < -1;
  r._pad = 0;
  return r;
}

/* invert.zig:474-715 checkFloorsMhd: the b^2/rho floor in the drift frame
   (b2rhofloorframe 0) or the ZAMO frame (1) */
int kzo_check_floors_mhd(double *pp, const kzo_geo *gm, double pgamma, const kz_floor_params *fp,
                         uint64_t *n_guard) {
  int ret = 0;
  /* invert.zig:485-492: the ZAMO floor adds its delta to the conserveds of
     the primitives as they enter, before any floor */
  double uu_entry[NV] = {0};
  if (fp->b2rhofloorframe == 1) kzo_p2u_mhd(pp, uu_entry, gm, pgamma);
  if (pp[RHO] < fp->rhofloor) {
    pp[RHO] = fp->rhofloor;
    ret = -1;
  }
  if (pp[UU] < fp->uurhoratiomin * pp[RHO]) {
    pp[UU] = fp->uurhoratiomin * pp[RHO];
    ret = -1;
  }
  if (pp[UU] > fp->uurhoratiomax * pp[RHO]) {
    pp[UU] = fp->uurhoratiomax * pp[RHO];
    ret = -1;
  }

  {
    double ucon[4], ucov[4];
    kzo_ucon_ucov_from_prims(pp, gm, ucon, ucov);
    double B[3] = {pp[B1], pp[B2], pp[B3]};
    double bcon[4], bcov[4], bsq;
    kzo_bcon_bcov_bsq(B, ucon, ucov, gm, bcon, bcov, &bsq);
    double f = 1.0, fuu = 1.0;
    int floored = 0;
    if (bsq > fp->b2rhoratiomax * pp[RHO]) {
      f = bsq / (fp->b2rhoratiomax * pp[RHO]);
      floored = 1;
    }
    if (fp->b2rhofloorframe == 0 && fp->b2uufloor && bsq > fp->b2uuratiomax * pp[UU]) {
      fuu = bsq / (fp->b2uuratiomax * pp[UU]);
      floored = 1;
    }
    if (floored && fp->b2rhofloorframe == 1) {
      /* invert.zig:649-684 (ZAMOFRAME): the new mass drho = rho (f - 1) moves
         with the normal observer n^mu = -alpha g^{mu 0}; its conserved vector
         is added to the entry conserveds through B3 and the primitives are
         recovered by the hot, then the entropy solver; if both fail the
         fluid-frame backup scales rho and u by f */
      double pporg[NV];
      for (int iv = 0; iv < NV; ++iv) pporg[iv] = pp[iv];
      double drho = pp[RHO] * (f - 1.0);
      double etacon[4], etarel[4];
      for (int i = 0; i < 4; ++i) etacon[i] = -gm->alpha * gm->G[i][0];
      kzo_convert(etacon, VEL4, VELR, gm, 1, etarel);
      double dpp[NV], duu[NV], uu[NV];
      for (int iv = 0; iv < NV; ++iv) dpp[iv] = 0.0;
      dpp[RHO] = drho;
      dpp[VX] = etarel[1];
      dpp[VY] = etarel[2];
      dpp[VZ] = etarel[3];
      kzo_p2u_mhd(dpp, duu, gm, pgamma);
      for (int iv = 0; iv < NV; ++iv) uu[iv] = uu_entry[iv];
      for (int iv = 0; iv <= B3; ++iv) uu[iv] += duu[iv];
      int rt = kzo_u2p_solver_w(uu, pp, gm, 0, pgamma);
      if (rt < 0) rt = kzo_u2p_solver_w(uu, pp, gm, 1, pgamma);
      if (rt < 0) {
        for (int iv = 0; iv < NV; ++iv) pp[iv] = pporg[iv];
        pp[RHO] *= f;
        pp[UU] *= f;
      }
      ret = -1;
    } else if (floored) {
      double wold = pp[RHO] + pp[UU] * pgamma;
      pp[RHO] *= f;
      pp[UU] *= fp->isentropic_b2rhofloors ? f : fuu;
      double wnew = pp[RHO] + pp[UU] * pgamma;
      int guarded = 0;
      if (gm->x1 >= fp->horizon_x1) {
        do {
          double Bcon[4] = {0, pp[B1], pp[B2], pp[B3]};
          double Bcov[4];
          kzo_lower(Bcon, gm, Bcov);
          double Bsq = kzo_dot(Bcon, Bcov);
          double Bmag = kzm_sqrt(Bsq);
          if (!(bsq > 1.0e-10 * wnew && Bsq > 1.0e-10 * wnew)) {
            guarded = 1;
            break;
          }
          double betapar = -bcon[0] / (bsq * ucon[0]);
          double betasq = betapar * betapar * bsq;
          double betasqmax = 1.0 - 1.0 / (fp->gammamaxhd * fp->gammamaxhd);
          if (betasq > betasqmax) betasq = betasqmax;
          double gammapar = 1.0 / kzm_sqrt(1.0 - betasq);
          double ucondr[4];
          for (int iv = 0; iv < 4; ++iv) ucondr[iv] = gammapar * (ucon[iv] + betapar * bcon[iv]);
          if (!kzm_isfinite(ucondr[0]) || kzm_fabs(ucondr[0]) < 1.0e-10) {
            guarded = 1;
            break;
          }
          double udotB = kzo_dot(ucon, Bcov);
          double QdotB = udotB * wold * ucon[0];
          double xx = 2.0 * QdotB / (Bmag * wnew * ucondr[0]);
          if (!(xx <= 1.0e10)) {
            xx = 1.0e10;
            guarded = 1;
          }
          if (!(xx >= -1.0e10)) {
            xx = -1.0e10;
            guarded = 1;
          }
          double vpar = xx / (ucondr[0] * (1.0 + kzm_sqrt(1.0 + xx * xx)));
          double vcon[4];
          vcon[0] = 1.0;
          for (int iv = 1; iv < 4; ++iv) vcon[iv] = vpar * Bcon[iv] / Bmag + ucondr[iv] / ucondr[0];
          for (int idamp = 0; idamp < 50; ++idamp) {
            double bv = 0.0, cv = 0.0;
            for (int iv = 1; iv < 4; ++iv) {
              bv += vcon[iv] * gm->g[0][iv];
              for (int jv = 1; jv < 4; ++jv) cv += vcon[iv] * vcon[jv] * gm->g[iv][jv];
            }
            if (gm->g[0][0] + 2.0 * bv + cv < -1.0 / (fp->gammamaxhd * fp->gammamaxhd)) break;
            guarded = 1;
            vpar *= 0.5;
            for (int iv = 1; iv < 4; ++iv) vcon[iv] = vpar * Bcon[iv] / Bmag + ucondr[iv] / ucondr[0];
          }
          double ucont[4];
          if (kzo_convert(vcon, VEL3, VELR, gm, 0, ucont) != KZO_OK) {
            guarded = 1;
            break;
          }
          pp[VX] = ucont[1];
          pp[VY] = ucont[2];
          pp[VZ] = ucont[3];
        } while (0);
      }
      if (guarded && n_guard) *n_guard += 1;
      ret = -1;
    }
  }

  {
    double qsq = 0;
    for (int i = 1; i < 4; ++i)
      for (int j = 1; j < 4; ++j) qsq += pp[VX - 1 + i] * pp[VX - 1 + j] * gm->g[i][j];
    double gamma2 = 1.0 + qsq;
    if (gamma2 > fp->gammamaxhd * fp->gammamaxhd) {
      double qsqmax = fp->gammamaxhd * fp->gammamaxhd - 1.0;
      double a = kzm_sqrt(qsqmax / qsq);
      for (int j = 1; j < 4; ++j) pp[VX - 1 + j] *= a;
      ret = -1;
    }
  }
  if (ret < 0) pp[ENTR] = kzo_s_from_u(pp[RHO], pp[UU], pgamma);
  return ret;
}

/* ------------------------------------------------------------------------ */
/* invert_rad.zig                                                            */

static const double GAMMASMALLLIMIT = 1.0 - 1e-10;          /* invert_rad.zig:32 */
static const double NUMEPSILON = 2.220446049250313080847e-16; /* DBL_EPSILON */

/* invert_rad.zig:66-124 m1GammaRel2 */
static double m1_gammarel2(const kzo_geo *gm, const double avcov[4]) {
  double gn11 = gm->G[0][0], gn12 = gm->G[0][1], gn13 = gm->G[0][2], gn14 = gm->G[0][3];
  double gn22 = gm->G[1][1], gn23 = gm->G[1][2], gn24 = gm->G[1][3];
  double gn33 = gm->G[2][2], gn34 = gm->G[2][3], gn44 = gm->G[3][3];
  double rdtt = avcov[0], rdtx = avcov[1], rdty = avcov[2], rdtz = avcov[3];
  double mix = gn12 * rdtx + gn13 * rdty + gn14 * rdtz;
  double sq = kzm_sqrt(4.0 * (gn11 * gn11) * (rdtt * rdtt) + mix * mix +
                       gn11 * (8.0 * gn12 * rdtt * rdtx + 3.0 * gn22 * (rdtx * rdtx) + 8.0 * gn13 * rdtt * rdty +
                               6.0 * gn23 * rdtx * rdty + 3.0 * gn33 * (rdty * rdty) + 8.0 * gn14 * rdtt * rdtz +
                               6.0 * gn24 * rdtx * rdtz + 6.0 * gn34 * rdty * rdtz + 3.0 * gn44 * (rdtz * rdtz)));
  double den = gn11 * (rdtt * rdtt) + 2.0 * gn12 * rdtt * rdtx + gn22 * (rdtx * rdtx) + 2.0 * gn13 * rdtt * rdty +
               2.0 * gn23 * rdtx * rdty + gn33 * (rdty * rdty) + 2.0 * (gn14 * rdtt + gn24 * rdtx + gn34 * rdty) * rdtz +
               gn44 * (rdtz * rdtz);
  double gamma2a = (-0.25 * (2.0 * (gn11 * gn11) * (rdtt * rdtt) + mix * (mix + sq) +
                             gn11 * (4.0 * gn12 * rdtt * rdtx + gn22 * (rdtx * rdtx) + 2.0 * gn23 * rdtx * rdty +
                                     gn33 * (rdty * rdty) + 2.0 * gn24 * rdtx * rdtz + 2.0 * gn34 * rdty * rdtz +
                                     gn44 * (rdtz * rdtz) + rdtt * (4.0 * gn13 * rdty + 4.0 * gn14 * rdtz + sq)))) /
                   den;
  double gamma2;
  if (gamma2a < GAMMASMALLLIMIT || !kzm_isfinite(gamma2a)) {
    double gamma2b = (0.25 * (-2.0 * (gn11 * gn11) * (rdtt * rdtt) -
                              1.0 * gn11 * (4.0 * gn12 * rdtt * rdtx + gn22 * (rdtx * rdtx) +
                                            rdty * (4.0 * gn13 * rdtt + 2.0 * gn23 * rdtx + gn33 * rdty) +
                                            2.0 * (2.0 * gn14 * rdtt + gn24 * rdtx + gn34 * rdty) * rdtz +
                                            gn44 * (rdtz * rdtz)) +
                              gn11 * rdtt * sq +
                              mix * (-1.0 * gn12 * rdtx - 1.0 * gn13 * rdty - 1.0 * gn14 * rdtz + sq))) /
                     den;
    gamma2 = gamma2b;
  } else {
    gamma2 = gamma2a;
  }
  double alpha = gm->alpha;
  double gammarel2 = gamma2 * alpha * alpha;
  if (gammarel2 > GAMMASMALLLIMIT && gammarel2 < 1.0) gammarel2 = 1.0;
  return gammarel2;
}

/* invert_rad.zig:127-130 m1Erf */
static double m1_erf(const kzo_geo *gm, double avcon0, double gammarel2) {
  double alpha = gm->alpha;
  return 3.0 * avcon0 * alpha * alpha / (4.0 * gammarel2 - 1.0);
}

/* invert_rad.zig:140-185 m1Cold */
static void m1_cold(const kzo_geo *gm, const double avcov_in[4], double gammarel2, double avcov[4],
                    double *erf_out, double urfconrel[4]) {
  double gn11 = gm->G[0][0], gn12 = gm->G[0][1], gn13 = gm->G[0][2], gn14 = gm->G[0][3];
  double gn22 = gm->G[1][1], gn23 = gm->G[1][2], gn24 = gm->G[1][3];
  double gn33 = gm->G[2][2], gn34 = gm->G[2][3], gn44 = gm->G[3][3];
  double alpha = gm->alpha;
  double rdtx = avcov_in[1], rdty = avcov_in[2], rdtz = avcov_in[3];
  double utsq = gammarel2 / (alpha * alpha);
  for (int i = 0; i < 4; ++i) avcov[i] = avcov_in[i];
  double mix = gn13 * rdty + gn14 * rdtz;
  double disc = kzm_sqrt(((gn12 * gn12) * (rdtx * rdtx) + 2.0 * gn12 * rdtx * mix + mix * mix -
                          1.0 * gn11 * (gn22 * (rdtx * rdtx) + 2.0 * gn23 * rdtx * rdty + gn33 * (rdty * rdty) +
                                        2.0 * gn24 * rdtx * rdtz + 2.0 * gn34 * rdty * rdtz + gn44 * (rdtz * rdtz))) *
                         utsq * (gn11 + utsq) * ((gn11 + 4.0 * utsq) * (gn11 + 4.0 * utsq)));
  avcov[0] = (0.25 * (-4.0 * (gn12 * rdtx + gn13 * rdty + gn14 * rdtz) * utsq * (gn11 + utsq) + disc)) /
             (gn11 * utsq * (gn11 + utsq));
  double erf = (0.75 * disc) / (utsq * (gn11 + utsq) * (gn11 + 4.0 * utsq));
  double avcon[4];
  kzo_raise(avcov, gm, avcon);
  double gammarel = kzm_sqrt(gammarel2);
  for (int i = 0; i < 4; ++i) urfconrel[i] = 0;
  if (erf > 0.0) {
    for (int jj = 1; jj < 4; ++jj)
      urfconrel[jj] = alpha * (avcon[jj] + 1.0 / 3.0 * erf * gm->G[0][jj] * (4.0 * gammarel2 - 1.0)) /
                      (4.0 / 3.0 * erf * gammarel);
  }
  *erf_out = erf;
}

/* invert_rad.zig:189-261 u2pRad */
kz_u2prad_result kzo_u2p_rad(const double *uu, double *pp, const kzo_geo *gm, const kz_rad_params *rp) {
  double gdetu = gm->gdet;
  double alpha = gm->alpha;
  double avcov[4] = {uu[EE] / gdetu, uu[FX] / gdetu, uu[FY] / gdetu, uu[FZ] / gdetu};
  double avcon[4];
  kzo_raise(avcov, gm, avcon);
  double gammarel2 = m1_gammarel2(gm, avcov);
  double erf = m1_erf(gm, avcon[0], gammarel2);
  double gammamax = rp->gammamaxrad;
  int nonfailure = gammarel2 >= 1.0 && erf > rp->eradfloor &&
                   gammarel2 <= gammamax * gammamax / GAMMASMALLLIMIT / GAMMASMALLLIMIT;
  double urfcon[4];
  int corrected = 0;
  kz_u2prad_result res;
  if (nonfailure) {
    double gammarel = kzm_sqrt(gammarel2);
    for (int ii = 0; ii < 4; ++ii)
      urfcon[ii] = alpha * (avcon[ii] + 1.0 / 3.0 * erf * gm->G[0][ii] * (4.0 * gammarel2 - 1.0)) /
                   (4.0 / 3.0 * erf * gammarel);
  } else {
    double one_eps = 1.0 + 10.0 * NUMEPSILON;
    double s_av[4], s_erf, s_u[4], f_av[4], f_erf, f_u[4];
    m1_cold(gm, avcov, one_eps * one_eps, s_av, &s_erf, s_u);
    m1_cold(gm, avcov, gammamax * gammamax, f_av, &f_erf, f_u);
    if (kzm_fabs(s_av[0] - avcov[0]) > kzm_fabs(f_av[0] - avcov[0])) {
      kzo_raise(f_av, gm, avcon);
      for (int i = 0; i < 4; ++i) urfcon[i] = f_u[i];
      gammarel2 = gammamax * gammamax;
      erf = f_erf;
    } else {
      kzo_raise(s_av, gm, avcon);
      for (int i = 0; i < 4; ++i) urfcon[i] = s_u[i];
      gammarel2 = one_eps * one_eps;
      erf = s_erf;
    }
    corrected = 1;
    if (!kzm_isfinite(erf) || !kzm_isfinite(gammarel2) || !kzm_isfinite(urfcon[1]) ||
        !kzm_isfinite(urfcon[2]) || !kzm_isfinite(urfcon[3])) {
      res.ret = -1;
      res.corrected = corrected;
      return res;
    }
  }
  pp[EE] = erf;
  pp[FX] = urfcon[1];
  pp[FY] = urfcon[2];
  pp[FZ] = urfcon[3];
  res.ret = 0;
  res.corrected = corrected;
  return res;
}

/* invert_rad.zig:266-303 checkFloorsRad */
int kzo_check_floors_rad(double *pp, const kzo_geo *gm, const kz_rad_params *rp) {
  int ret = 0;
  double rtt = kzo_calc_ff_rtt(pp, gm, 0);
  double ehat = -rtt;
  double eratio = pp[EE] / ehat;
  if (pp[EE] < rp->eradfloor) {
    pp[EE] = rp->eradfloor;
    ret = -1;
  }
  if (ehat < rp->eerhoratiomin * pp[RHO]) {
    pp[EE] = eratio * rp->eerhoratiomin * pp[RHO];
    ret = -1;
  }
  if (ehat > rp->eerhoratiomax * pp[RHO]) {
    pp[RHO] = 1.0 / rp->eerhoratiomax * ehat;
    ret = -1;
  }
  if (ehat < rp->eeuuratiomin * pp[UU]) {
    pp[EE] = eratio * rp->eeuuratiomin * pp[UU];
    ret = -1;
  }
  if (ehat > rp->eeuuratiomax * pp[UU]) {
    pp[UU] = 1.0 / rp->eeuuratiomax * ehat;
    ret = -1;
  }
  return ret;
}

/* ------------------------------------------------------------------------ */
/* implicit.zig                                                              */

static const double SMALL_IMP = 1.0e-80; /* implicit.zig:61 */

/* implicit.zig:167-172 mySign */
static double my_sign(double x) {
  if (x > 0.0) return 1.0;
  if (x < 0.0) return -1.0;
  return 1.0;
}

/* implicit.zig:176-224 invert4 */
void kzo_invert4(const double a[4][4], double inv[4][4]) {
  double lu[4][4];
  int perm[4] = {0, 1, 2, 3};
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) lu[i][j] = a[i][j];
  for (int k = 0; k < 4; ++k) {
    int piv = k;
    double pmax = kzm_fabs(lu[k][k]);
    for (int r = k + 1; r < 4; ++r) {
      if (kzm_fabs(lu[r][k]) > pmax) {
        pmax = kzm_fabs(lu[r][k]);
        piv = r;
      }
    }
    if (piv != k) {
      for (int j = 0; j < 4; ++j) {
        double t = lu[k][j];
        lu[k][j] = lu[piv][j];
        lu[piv][j] = t;
      }
      int tp = perm[k];
      perm[k] = perm[piv];
      perm[piv] = tp;
    }
    for (int r = k + 1; r < 4; ++r) {
      lu[r][k] /= lu[k][k];
      for (int cc = k + 1; cc < 4; ++cc) lu[r][cc] -= lu[r][k] * lu[k][cc];
    }
  }
  for (int col = 0; col < 4; ++col) {
    double y[4], x[4];
    for (int r = 0; r < 4; ++r) {
      y[r] = (perm[r] == col) ? 1.0 : 0.0;
      for (int cc = 0; cc < r; ++cc) y[r] -= lu[r][cc] * y[cc];
    }
    for (int r = 3; r >= 0; --r) {
      x[r] = y[r];
      for (int cc = r + 1; cc < 4; ++cc) x[r] -= lu[r][cc] * x[cc];
      x[r] /= lu[r][r];
    }
    for (int rr = 0; rr < 4; ++rr) inv[rr][col] = x[rr];
  }
}

/* kzo_invert4 in FP32 (the mixed-precision Newton's LU, not in the
 * reference): the same pivoting and elimination order on floats. */
void kzo_invert4f(const float a[4][4], float inv[4][4]) {
  float lu[4][4];
  int perm[4] = {0, 1, 2, 3};
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) lu[i][j] = a[i][j];
  for (int k = 0; k < 4; ++k) {
    int piv = k;
    float pmax = __builtin_fabsf(lu[k][k]);
    for (int r = k + 1; r < 4; ++r) {
      if (__builtin_fabsf(lu[r][k]) > pmax) {
        pmax = __builtin_fabsf(lu[r][k]);
        piv = r;
      }
    }
    if (piv != k) {
      for (int j = 0; j < 4; ++j) {
        float t = lu[k][j];
        lu[k][j] = lu[piv][j];
        lu[piv][j] = t;
      }
      int tp = perm[k];
      perm[k] = perm[piv];
      perm[piv] = tp;
    }
    for (int r = k + 1; r < 4; ++r) {
      lu[r][k] /= lu[k][k];
      for (int cc = k + 1; cc < 4; ++cc) lu[r][cc] -= lu[r][k] * lu[k][cc];
    }
  }
  for (int col = 0; col < 4; ++col) {
    float y[4], x[4];
    for (int r = 0; r < 4; ++r) {
      y[r] = (perm[r] == col) ? 1.0f : 0.0f;
      for (int cc = 0; cc < r; ++cc) y[r] -= lu[r][cc] * y[cc];
    }
    for (int r = 3; r >= 0; --r) {
      x[r] = y[r];
      for (int cc = r + 1; cc < 4; ++cc) x[r] -= lu[r][cc] * x[cc];
      x[r] /= lu[r][r];
    }
    for (int rr = 0; rr < 4; ++rr) inv[rr][col] = x[rr];
  }
}

typedef struct {
  const kzo_geo *gm;
  double gam;
  const kz_rad_params *rp;
  const kz_opac_params *op;
  const kz_implicit_params *ip;
} impl_ctx;

/* implicit.zig:240-306 applyConstraints */
static int apply_constraints(double *pp, double *uu, const double *uu0, const impl_ctx *cx, int whichprim) {
  const kzo_geo *gm = cx->gm;
  double gdetu = gm->gdet;
  double gdetu_inv = 1.0 / gdetu;
  if (whichprim == WP_MHD) {
    double ucon[4], ucov[4];
    kzo_ucon_ucov_from_prims(pp, gm, ucon, ucov);
    pp[RHO] = uu0[RHO] * gdetu_inv / ucon[0];
    pp[ENTR] = kzo_s_from_u(pp[RHO], pp[UU], cx->gam);
    kzo_p2u_mhd(pp, uu, gm, cx->gam);
    for (int k = 0; k < 4; ++k) uu[EE + k] = uu0[EE + k] - (uu[UU + k] - uu0[UU + k]);
    kz_u2prad_result rr = kzo_u2p_rad(uu, pp, gm, cx->rp);
    int u2pret = rr.ret;
    if (rr.corrected) u2pret = cx->ip->allow_rad_ceiling ? -1 : -2;
    return u2pret;
  } else {
    kzo_p2u_rad(pp, uu, gm);
    uu[RHO] = uu0[RHO];
    for (int k = 0; k < 4; ++k) uu[UU + k] = uu0[UU + k] - (uu[EE + k] - uu0[EE + k]);
    for (int k = 0; k < 3; ++k) uu[B1 + k] = gdetu * pp[B1 + k];
    int rettemp = kzo_u2p_solver_w(uu, pp, gm, 0, cx->gam);
    if (rettemp < 0 && cx->ip->allow_entr_in_4dprim) rettemp = kzo_u2p_solver_w(uu, pp, gm, 1, cx->gam);
    int u2pret = rettemp < 0 ? -2 : 0;
    double ucon[4], ucov[4];
    kzo_ucon_ucov_from_prims(pp, gm, ucon, ucov);
    pp[ENTR] = kzo_s_from_u(pp[RHO], pp[UU], cx->gam);
    uu[ENTR] = pp[ENTR] * gdetu * ucon[0];
    return u2pret;
  }
}

/* implicit.zig:337-427 residualG (scalar instantiation) */
static double residual(const double *uu, const double *pp, const kzo_radstate *st, const double *uu0,
                       const kzo_radstate *st0, double dt, const impl_ctx *cx, const kz_opac_params *op,
                       int whichprim, int whicheq, int whichframe, double f[4]) {
  const kzo_geo *gm = cx->gm;
  double gdetu = gm->gdet;
  double tgas = st->tgas;
  double ehat0 = st0->ehat;
  double ehat = st->ehat;
  double dtau = dt / st->ucon[0];
  double giff[4], gilab[4], gi[4];
  kzo_calc_gi(st, pp, gm, op, 0, giff, gilab);
  kzo_lower(gilab, gm, gi);
  double err[4] = {0, 0, 0, 0};
  if (whichprim == WP_MHD) {
    for (int k = 1; k < 4; ++k) {
      f[k] = uu[UU + k] - uu0[UU + k] - dt * gdetu * gi[k];
      err[k] = (kzm_fabs(f[k]) > SMALL_IMP)
                   ? kzm_fabs(f[k]) / (1.0e-20 * uu[UU] + kzm_fabs(uu[UU + k]) + kzm_fabs(uu0[UU + k]) +
                                       kzm_fabs(dt * gdetu * gi[k]))
                   : 0.0;
    }
  } else {
    for (int k = 1; k < 4; ++k) {
      f[k] = uu[EE + k] - uu0[EE + k] + dt * gdetu * gi[k];
      err[k] = (kzm_fabs(f[k]) > SMALL_IMP)
                   ? kzm_fabs(f[k]) / (1.0e-20 * uu[EE] + kzm_fabs(uu[EE + k]) + kzm_fabs(uu0[EE + k]) +
                                       kzm_fabs(dt * gdetu * gi[k]))
                   : 0.0;
    }
  }
  if (whichframe == WF_LAB) {
    if (whichprim == WP_RAD) {
      f[0] = uu[EE] - uu0[EE] + dt * gdetu * gi[0];
      err[0] = (kzm_fabs(f[0]) > SMALL_IMP)
                   ? kzm_fabs(f[0]) / (kzm_fabs(uu[EE]) + kzm_fabs(uu0[EE]) + kzm_fabs(dt * gdetu * gi[0]))
                   : 0.0;
    } else {
      f[0] = uu[UU] - uu0[UU] - dt * gdetu * gi[0];
      err[0] = (kzm_fabs(f[0]) > SMALL_IMP)
                   ? kzm_fabs(f[0]) / (kzm_fabs(uu[UU]) + kzm_fabs(uu0[UU]) + kzm_fabs(dt * gdetu * gi[0]))
                   : 0.0;
    }
  } else {
    if (whicheq == WE_ENERGY) {
      if (whichprim == WP_RAD) {
        f[0] = ehat - ehat0 + dtau * giff[0];
        err[0] = kzm_fabs(f[0]) / (kzm_fabs(ehat) + kzm_fabs(ehat0) + kzm_fabs(dtau * giff[0]));
      } else {
        f[0] = pp[UU] - st0->uint - dtau * giff[0];
        err[0] = kzm_fabs(f[0]) / (kzm_fabs(pp[UU]) + kzm_fabs(st0->uint) + kzm_fabs(dtau * giff[0]));
      }
    } else {
      double s0 = st0->sgas;
      double s = st->sgas;
      f[0] = tgas * (s - s0) - dtau * giff[0];
      err[0] = kzm_fabs(f[0]) / (kzm_fabs(tgas * s) + kzm_fabs(tgas * s0) + kzm_fabs(dtau * giff[0]));
    }
  }
  double err0 = 0;
  for (int k = 0; k < 4; ++k) err0 = err[k] > err0 ? err[k] : err0;
  return err0;
}

/* implicit.zig:431-467 f1dErr */
static double f1d_err(double xxx, const double *pp0, double dtau, const impl_ctx *cx, int whichprim,
                      double totenergy, double ratio) {
  double ugas, ehat;
  if (whichprim == WP_RAD) {
    ehat = xxx * ratio;
    ugas = totenergy - ehat;
  } else {
    ugas = xxx;
    ehat = totenergy - ugas;
  }
  double pp[NV];
  for (int i = 0; i < NV; ++i) pp[i] = pp0[i];
  pp[UU] = ugas;
  pp[EE] = ehat / ratio;
  kzo_radstate st;
  kzo_fill_radstate(pp, cx->gm, cx->gam, cx->op, 0, &st);
  double giff[4], gilab[4];
  kzo_calc_gi(&st, pp, cx->gm, cx->op, 0, giff, gilab);
  if (whichprim == WP_RAD) return ehat - pp0[EE] / ratio + dtau * giff[0];
  return ugas - pp0[UU] - dtau * giff[0];
}

/* implicit.zig:473-574 solve1dPrim */
static int solve_1d_prim(const double *pp00, const kzo_radstate *st00, double dt, const impl_ctx *cx,
                         double *pp0out) {
  double ehat = st00->ehat;
  double dtau = dt / st00->ucon[0];
  double totenergy = ehat + pp00[UU];
  int whichprim = (ehat < 1.0e-3 * pp00[UU]) ? WP_RAD : WP_MHD;
  double tgas = st00->tgas;
  double ratio = ehat / pp00[EE];
  double fac = 3.0;
  double fac1 = fac + 1.0;
  double pp_uu = pp00[UU];
  double pp_ee = pp00[EE];
  if (tgas > 1.0e14) {
    pp_uu = pp00[UU] * 1.0e14 / tgas;
    pp_ee = (totenergy - pp_uu) / ratio;
  }
  double xlo, xhi;
  if (whichprim == WP_MHD) {
    xlo = pp_uu / 1.05;
    xhi = kzm_fmin(pp_uu * 1.05, (pp_uu + fac * totenergy) / fac1);
  } else {
    xlo = pp_ee / 1.05;
    xhi = kzm_fmin(pp_ee * 1.05, (pp_ee + fac * totenergy / ratio) / fac1);
  }
  double errlo = f1d_err(xlo, pp00, dtau, cx, whichprim, totenergy, ratio);
  double errhi = f1d_err(xhi, pp00, dtau, cx, whichprim, totenergy, ratio);
  const int maxiter = 50;
  const double conv = 1.0e-3;
  int iter = 0;
  while (errlo * errhi > 0.0) {
    iter += 1;
    if (errlo > 0.0) {
      xhi = xlo;
      errhi = errlo;
      xlo /= fac;
      errlo = f1d_err(xlo, pp00, dtau, cx, whichprim, totenergy, ratio);
    } else {
      xlo = xhi;
      errlo = errhi;
      if (whichprim == WP_RAD)
        xhi = kzm_fmin(xhi * fac, (xhi + fac * totenergy / ratio) / fac1);
      else
        xhi = kzm_fmin(xhi * fac, (xhi + fac * totenergy) / fac1);
      errhi = f1d_err(xhi, pp00, dtau, cx, whichprim, totenergy, ratio);
    }
    if (errlo != errlo || errhi != errhi) iter = maxiter;
    if (iter >= maxiter) break;
  }
  if (iter >= maxiter) {
    for (int i = 0; i < NV; ++i) pp0out[i] = pp00[i];
    return -1;
  }
  double xmid;
  for (;;) {
    iter += 1;
    xmid = 0.5 * (xlo + xhi);
    double errmid = f1d_err(xmid, pp00, dtau, cx, whichprim, totenergy, ratio);
    if (errmid * errlo > 0.0) {
      xlo = xmid;
      errlo = errmid;
    } else {
      xhi = xmid;
      errhi = errmid;
    }
    if (xmid != xmid || errmid != errmid) {
      for (int i = 0; i < NV; ++i) pp0out[i] = pp00[i];
      return -1;
    }
    if (!(kzm_fabs((xhi - xlo) / xmid) > conv)) break;
  }
  for (int i = 0; i < NV; ++i) pp0out[i] = pp00[i];
  if (whichprim == WP_RAD) {
    pp0out[EE] = xmid;
    pp0out[UU] = totenergy - xmid * ratio;
  } else {
    pp0out[UU] = xmid;
    pp0out[EE] = (totenergy - xmid) / ratio;
  }
  return 0;
}

/* implicit.zig:586-591 fdDel */
static double fd_del(int j, double sign, const double *ppp, int sh, double eps, const kzo_geo *gm) {
  if (j == 0) return sign * eps * ppp[sh];
  double veleps = eps * my_sign(ppp[j + sh]) * kzm_fmax(1.0e-8 / kzm_sqrt(gm->g[j][j]), kzm_fabs(ppp[j + sh]));
  return ppp[j + sh] >= 0.0 ? sign * veleps : -sign * veleps;
}

/* implicit.zig:631-640 lagOpac */
static void lag_opac(kzo_radstate *st, const kzo_radstate *st00) {
  st->kappaes = st00->kappaes;
  st->opac.gas_abs = st00->opac.gas_abs;
  st->opac.rad_abs = st00->opac.rad_abs;
  st->opac.gas_num = st00->opac.gas_num;
  st->opac.rad_num = st00->opac.rad_num;
  st->opac.gas_ross = st00->opac.gas_ross;
  st->opac.rad_ross = st00->opac.rad_ross;
}

static void fill_state(const double *pp, const impl_ctx *cx, const kz_opac_params *op, int slim,
                       const kzo_radstate *st00, kzo_radstate *st) {
  kzo_fill_radstate(pp, cx->gm, cx->gam, op, slim ? 0 : 1, st);
  if (cx->ip->lag_opac) lag_opac(st, st00);
}

/* implicit.zig:642-922 solve4dPrimCore (scalar Jacobian path; the reference's
 * 4-lane batch is bit-identical to it, simd_tests.zig). Returns ok; iters. */
static int solve_4d_prim_core(const double *uu00, const kzo_radstate *state00, const double *pp0in, double dt,
                              const impl_ctx *cx, const kz_opac_params *op, int whichprim, int whicheq,
                              int whichframe, double *pp_out, int *iters_out, double *uu_out) {
  const kz_implicit_params *ip = cx->ip;
  const kzo_geo *gm = cx->gm;
  double uu[NV], pp[NV], ppp[NV];
  for (int i = 0; i < NV; ++i) {
    uu[i] = uu00[i];
    pp[i] = pp0in[i];
  }
  int sh = (whichprim == WP_MHD) ? UU : EE;
  double conv = (whicheq == WE_ENTROPY) ? ip->entrconv : ip->conv;
  int fret = apply_constraints(pp, uu, uu00, cx, whichprim);
  int slim = (whicheq != WE_ENTROPY);
  kzo_radstate state;
  fill_state(pp, cx, op, slim, state00, &state);

  double f1[4], f2[4], jac[4][4];
  int iter = 0;
  int failed = 0;
  for (;;) {
    iter += 1;
    for (int i = 0; i < NV; ++i) ppp[i] = pp[i];
    double xxx[4];
    for (int k = 0; k < 4; ++k) xxx[k] = ppp[k + sh];
    if (fret < -1) return 0;
    double errbase = residual(uu, pp, &state, uu00, state00, dt, cx, op, whichprim, whicheq, whichframe, f1);
    if (errbase < conv) break;

    for (int j = 0; j < 4; ++j) {
      double sign = -1.0;
      for (;;) {
        double del = fd_del(j, sign, ppp, sh, ip->eps, gm);
        pp[j + sh] = ppp[j + sh] + del;
        fret = apply_constraints(pp, uu, uu00, cx, whichprim);
        fill_state(pp, cx, op, slim, state00, &state);
        if (fret < -1) {
          if (sign > 0.0) {
            failed = 1;
            break;
          } else {
            pp[j + sh] = ppp[j + sh];
            sign *= -1.0;
            continue;
          }
        }
        residual(uu, pp, &state, uu00, state00, dt, cx, op, whichprim, whicheq, whichframe, f2);
        for (int i = 0; i < 4; ++i) jac[i][j] = (f2[i] - f1[i]) / del;
        pp[j + sh] = ppp[j + sh];
        break;
      }
      if (failed) break;
    }
    if (failed) break;

    double scalevec[4] = {pp[sh], 1.0, 1.0, 1.0};
    double invscalevec[4] = {1.0, 1.0, 1.0, 1.0};
    if (ip->scale_jacobian) {
      for (int j = 0; j < 4; ++j) {
        if (scalevec[j] == 0.0) scalevec[j] = 1.0;
        invscalevec[j] = 1.0 / scalevec[j];
      }
    }
    double ijac[4][4];
    if (ip->jacobian_fp32) {
      /* mixed precision: FP64 quotients rounded to FP32, scaled and
       * factored in FP32, the inverse applied in FP64 */
      float jf[4][4], ijf[4][4];
      for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) jf[i][j] = (float)jac[i][j];
      if (ip->scale_jacobian)
        for (int i = 0; i < 4; ++i)
          for (int j = 0; j < 4; ++j)
            if (i != j) jf[i][j] = jf[i][j] * ((float)scalevec[j] * (float)invscalevec[i]);
      kzo_invert4f(jf, ijf);
      for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) ijac[i][j] = (double)ijf[i][j];
    } else {
      if (ip->scale_jacobian)
        for (int i = 0; i < 4; ++i)
          for (int j = 0; j < 4; ++j)
            if (i != j) jac[i][j] = jac[i][j] * (scalevec[j] * invscalevec[i]);
      kzo_invert4(jac, ijac);
    }

    double xiapp = 1.0;
    for (;;) {
      for (int k = 0; k < 4; ++k) xxx[k] = ppp[k + sh];
      for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) {
          if (ip->scale_jacobian)
            xxx[i] -= xiapp * (ijac[i][j] * scalevec[i] * invscalevec[j]) * f1[j];
          else
            xxx[i] -= xiapp * ijac[i][j] * f1[j];
        }
      for (int k = 0; k < 4; ++k) pp[k + sh] = xxx[k];
      fret = apply_constraints(pp, uu, uu00, cx, whichprim);
      fill_state(pp, cx, op, slim, state00, &state);

      int okcheck = 0;
      double edenmin = ppp[sh] / ip->max_en_change_down;
      double edenmax = ppp[sh] * ip->max_en_change_up;
      if (xxx[0] > edenmin && xxx[0] < edenmax && fret >= -1) okcheck = 1;
      const kz_opac_params *c = cx->op;
      double tgasold = kzo_t_from_urho(c, ppp[UU], ppp[RHO], cx->gam);
      double tgasnew = kzo_t_from_urho(c, pp[UU], pp[RHO], cx->gam);
      if (okcheck && (tgasnew > tgasold * ip->max_tgas_change || tgasnew < tgasold / ip->max_tgas_change))
        okcheck = 0;
      double tradnew = kzo_lte_t_from_e(c, pp[EE]);
      double tradold = kzo_lte_t_from_e(c, ppp[EE]);
      if (okcheck && (tradnew > tradold * ip->max_trad_change || tradnew < tradold / ip->max_trad_change))
        okcheck = 0;
      if (okcheck) break;
      if (xxx[0] <= 0.0) {
        xiapp *= ppp[sh] / (ppp[sh] + kzm_fabs(xxx[0]));
        xiapp *= 1.0e-1;
      } else {
        xiapp /= ip->damping_factor;
      }
      if (xiapp < ip->max_damping) {
        failed = 1;
        break;
      }
    }

    if (!failed) {
      double f3[4];
      f3[0] = kzm_fabs((pp[sh] - ppp[sh]) / ppp[sh]);
      for (int i = 1; i < 4; ++i) {
        f3[i] = pp[i + sh] - ppp[i + sh];
        f3[i] = kzm_fabs(f3[i] / kzm_fmax(ip->eps, kzm_fabs(ppp[i + sh])));
      }
      double convrel = (whicheq == WE_ENTROPY) ? ip->convrelentr : ip->convrel;
      double convrelerr = (whicheq == WE_ENTROPY) ? ip->convrelentrerr : ip->convrelerr;
      int convrelcheck = 1;
      for (int ib = 0; ib < 4; ++ib) {
        if (f3[ib] > convrel) {
          convrelcheck = 0;
          break;
        }
      }
      if (convrelcheck || errbase <= convrelerr) break;
    }
    if (iter > ip->maxiter || failed) break;
  }
  if (iter > ip->maxiter || failed) return 0;
  for (int i = 0; i < NV; ++i) pp_out[i] = pp[i];
  if (uu_out)
    for (int i = 0; i < NV; ++i) uu_out[i] = uu[i];
  *iters_out = iter;
  return 1;
}

/* implicit.zig:927-1004 solveImplicitLab */
kz_implicit_result kzo_solve_implicit_lab(double *uu, double *pp, const kzo_geo *gm, double dt, double gam,
                                          const kz_rad_params *rp, const kz_opac_params *op,
                                          const kz_implicit_params *ip) {
  kz_implicit_result fail = {0, -1, 0, WP_RAD, WE_ENERGY, WF_LAB};
  impl_ctx cx = {gm, gam, rp, op, ip};
  double pp00[NV], uu00[NV];
  for (int i = 0; i < NV; ++i) pp00[i] = pp[i];
  kzo_p2u(pp00, uu00, gm, gam);
  double rtt = kzo_calc_ff_rtt(pp00, gm, 0);
  double ehat = -rtt;
  kzo_radstate state00;
  kzo_fill_radstate(pp00, gm, gam, op, 1, &state00);
  double pp0[NV];
  for (int i = 0; i < NV; ++i) pp0[i] = pp00[i];
  if (ip->start_with_bisect) solve_1d_prim(pp00, &state00, dt, &cx, pp0);

  int startwith = (ehat < ip->threshold * pp00[UU]) ? WP_RAD : WP_MHD;
  int swapped = (startwith == WP_RAD) ? WP_MHD : WP_RAD;
  const int rp_[6] = {startwith, startwith, swapped, swapped, startwith, swapped};
  const int re_[6] = {WE_ENERGY, WE_ENERGY, WE_ENERGY, WE_ENERGY, WE_ENTROPY, WE_ENTROPY};
  const int rf_[6] = {WF_LAB, WF_FF, WF_LAB, WF_FF, WF_FF, WF_FF};

  for (int level = 0; level <= ip->opdamp_maxlevels; ++level) {
    kz_opac_params opl = *op;
    double od = 1.0;
    for (int i = 0; i < level; ++i) od /= ip->opdamp_factor;
    opl.opdamp = od;
    for (int ir = 0; ir < 6; ++ir) {
      double ppw[NV];
      for (int i = 0; i < NV; ++i) ppw[i] = pp[i];
      int iters = 0;
      int ok = solve_4d_prim_core(uu00, &state00, pp0, dt, &cx, &opl, rp_[ir], re_[ir], rf_[ir], ppw, &iters, 0);
      if (ok) {
        for (int i = 0; i < NV; ++i) pp[i] = ppw[i];
        kzo_p2u(ppw, uu, gm, gam);
        kz_implicit_result r = {1, ir, iters, rp_[ir], re_[ir], rf_[ir]};
        return r;
      }
    }
  }
  return fail;
}

/* implicit.zig:924-926 solve4dPrim called on its own (no bisection warm
 * start, no ladder), as the reference's implicit theory tests call it
 * (koral/tests/implicit_tests.zig:89-130, 305-340): one rung from pp00 with
 * the given primitive block, equation and frame.  uu_out gets the solver's
 * internal conserved vector (the constraint construction), pp_out the
 * primitives; ok[i] 0 leaves both untouched. */
int kzo_solve_4d_prim_batch(int64_t n, const double *pp00, const kz_geom *g, const double *dt, double gam,
                            const kz_rad_params *rp, const kz_opac_params *op, const kz_implicit_params *ip,
                            int32_t whichprim, int32_t whicheq, int32_t whichframe, double *pp_out, double *uu_out,
                            int32_t *ok, int32_t *iters) {
  for (int64_t i = 0; i < n; ++i) {
    kzo_geo gm;
    kzo_geo_from_kz(&g[i], &gm);
    impl_ctx cx = {&gm, gam, rp, op, ip};
    double uu00[NV];
    kzo_p2u(pp00 + i * NV, uu00, &gm, gam);
    kzo_radstate st00;
    kzo_fill_radstate(pp00 + i * NV, &gm, gam, op, 1, &st00);
    int it = 0;
    ok[i] = solve_4d_prim_core(uu00, &st00, pp00 + i * NV, dt[i], &cx, op, whichprim, whicheq, whichframe,
                               pp_out + i * NV, &it, uu_out + i * NV);
    iters[i] = it;
  }
  return 0;
}

/* implicit_golden_tests.zig:25-40 trueResidual: the max normalized error of
 * the converged primitives pp against the reference state pp00. */
int kzo_true_residual_batch(int64_t n, const double *pp, const double *pp00, const double *dt, const kz_geom *g,
                            double gam, const kz_opac_params *op, const int32_t *wp, const int32_t *we,
                            const int32_t *wf, double *err) {
  kz_implicit_params ip;
  memset(&ip, 0, sizeof ip);
  for (int64_t i = 0; i < n; ++i) {
    kzo_geo gm;
    kzo_geo_from_kz(&g[i], &gm);
    impl_ctx cx = {&gm, gam, 0, op, &ip};
    double uu00[NV], uux[NV], f[4];
    kzo_radstate st00, stx;
    kzo_p2u(pp00 + i * NV, uu00, &gm, gam);
    kzo_fill_radstate(pp00 + i * NV, &gm, gam, op, 1, &st00);
    kzo_p2u(pp + i * NV, uux, &gm, gam);
    kzo_fill_radstate(pp + i * NV, &gm, gam, op, 1, &stx);
    err[i] = residual(uux, pp + i * NV, &stx, uu00, &st00, dt[i], &cx, op, wp[i], we[i], wf[i], f);
  }
  return 0;
}
