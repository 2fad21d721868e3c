// kz_dev_solve.h — the per-cell inversions and the implicit four-force
// Newton of the koral step as gfx950 device code: koral/solve/invert.zig
// (u2p W-Newton, MHD floors), koral/solve/invert_rad.zig (closed-form M1
// u2p, radiation floors) and koral/solve/implicit.zig (6-rung ladder).
// One thread runs one cell's whole solve; branch decisions follow the
// reference exactly, so iteration counts and flags match the CPU path.
#pragma once

#include "kz_dev.h"

namespace kzd {

// ---- invert.zig ------------------------------------------------------------
struct Resid {
  double f, df, err;
};
struct Cons {
  double Qn, Qtsq, D, QdotBsq, Bsq, Sc;
};

KZD Resid f_u2p(int etype, double Wp, const Cons &c, double pg) {
  double W = Wp + c.D;  // invert.zig:116-210 (both residuals share wKinematics)
  double x = c.Bsq + W;
  double wsq = W * W;
  double w3 = wsq * W;
  double x3 = x * x * x;
  double v2 = (wsq * c.Qtsq + c.QdotBsq * (c.Bsq + 2.0 * W)) / (wsq * (x * x));
  double gamma2 = 1.0 / (1.0 - v2);
  double gamma = kzm_sqrt(gamma2);
  double rho0 = c.D / gamma;
  double wmrho0 = Wp / gamma2 - c.D * v2 / (1.0 + gamma);
  double dvsq = -2.0 / x3 * (c.Qtsq + c.QdotBsq * (3.0 * W * x + c.Bsq * c.Bsq) / w3);
  double dwmrho0dvsq = c.D * (gamma * 0.5 - 1.0) - Wp;
  Resid r;
  if (etype == 0) {
    double u = wmrho0 / pg;
    double p = (pg - 1.0) * u;
    r.f = c.Qn + W - p + 0.5 * c.Bsq * (1.0 + v2) - c.QdotBsq / 2.0 / wsq;
    r.err = kzm_fabs(r.f) / (kzm_fabs(c.Qn) + kzm_fabs(W) + kzm_fabs(p) + kzm_fabs(0.5 * c.Bsq * (1.0 + v2)) +
                             kzm_fabs(c.QdotBsq / 2.0 / wsq));
    double dp1 = (pg - 1.0) * (1.0 - v2) / pg;
    double idw = (pg - 1.0) / pg;
    double dp2 = dwmrho0dvsq * idw;
    double dpdW = dp1 + dp2 * dvsq;
    r.df = 1.0 - dpdW + c.QdotBsq / (wsq * W) + 0.5 * c.Bsq * dvsq;
  } else {
    double pressure = (pg - 1.0) / pg * wmrho0;
    double indexn = 1.0 / (pg - 1.0);
    double ss = kzm_log(kzm_pow(pressure, indexn) / kzm_pow(rho0, indexn + 1.0));
    r.f = -c.Sc + c.D * ss;
    r.err = kzm_fabs(r.f) / (kzm_fabs(c.Sc) + kzm_fabs(c.D * ss));
    double dssdrho = pg / ((1.0 - pg) * rho0);
    double dssdchi = 1.0 / ((pg - 1.0) * wmrho0);
    double dwmrho0dw = 1.0 / (gamma * gamma);
    double drho0dvsq = -c.D * gamma * 0.5;
    double dssdw = dwmrho0dw * dssdchi;
    double dssdvsq = drho0dvsq * dssdrho + dwmrho0dvsq * dssdchi;
    double dssdwp = dssdw + dssdvsq * dvsq;
    r.df = c.D * dssdwp;
  }
  return r;
}

KZD bool w_oob(double W, const Cons &c, const Resid &r) {  // invert.zig:221-224
  return (W * W * W * (W + 2.0 * c.Bsq) - c.QdotBsq * (2.0 * W + c.Bsq)) <= W * W * (c.Qtsq - c.Bsq * c.Bsq) ||
         !kzm_isfinite(r.f) || !kzm_isfinite(r.df);
}

// invert.zig:235-404 u2pSolverW (pp in/out; returns the reference's code)
KZD int u2p_solver_w(const double *uu, double *pp, const Geo &gm, int etype, double pg) {
  double gi = 1.0 / gm.gdet;
  double alpha = gm.alpha;
  Cons c;
  c.D = uu[RHO] * gi * alpha;
  c.Sc = uu[ENTR] * gi * alpha;
  double Qcov[4] = {(uu[UU] * gi - uu[RHO] * gi) * alpha, uu[VX] * gi * alpha, uu[VY] * gi * alpha,
                    uu[VZ] * gi * alpha};
  double Qcon[4];
  raise4(Qcov, gm, Qcon);
  double Bcon[4] = {0, uu[B1] * gi * alpha, uu[B2] * gi * alpha, uu[B3] * gi * alpha}, Bcov[4];
  lower4(Bcon, gm, Bcov);
  c.Bsq = dot4(Bcon, Bcov);
  double QdotB = dot4(Qcov, Bcon);
  c.QdotBsq = QdotB * QdotB;
  double ncov[4] = {-alpha, 0, 0, 0}, ncon[4];
  raise4(ncov, gm, ncon);
  c.Qn = Qcon[0] * ncov[0];
  double Qt[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    double s = 0;
#pragma unroll
    for (int j = 0; j < 4; ++j) s += ((i == j ? 1.0 : 0.0) + ncon[i] * ncov[j]) * Qcon[j];
    Qt[i] = s;
  }
  double Qtcov[4];
  lower4(Qt, gm, Qtcov);
  c.Qtsq = dot4(Qt, Qtcov);

  double qsq = 0;
#pragma unroll
  for (int i = 1; i < 4; ++i)
#pragma unroll
    for (int j = 1; j < 4; ++j) qsq += pp[VX - 1 + i] * pp[VX - 1 + j] * gm.g[i][j];
  double W = (pp[RHO] + pg * pp[UU]) * (1.0 + qsq);
  const double conv = 1.0e-10;

  int inc = 0;
  for (;;) {
    Resid r = f_u2p(etype, W - c.D, c, pg);
    if (w_oob(W, c, r) && inc < 50) {
      W *= 10.0;
      inc += 1;
      continue;
    }
    break;
  }
  if (inc >= 50) return -150;
  int iter = 0;
  while (iter < 50) {
    double Wprev = W;
    iter += 1;
    Resid r = f_u2p(etype, W - c.D, c, pg);
    double f0 = r.f, dfdW = r.df, errr = r.err;
    if (errr < conv) break;
    if (dfdW == 0.0) {
      W *= 1.1;
      continue;
    }
    double Wnew = W - f0 / dfdW;
    int idump = 0;
    double dumpfac = 1.0;
    for (;;) {
      Resid rt = f_u2p(etype, Wnew - c.D, c, pg);
      if (w_oob(Wnew, c, rt) && idump < 100) {
        idump += 1;
        dumpfac /= 2.0;
        Wnew = W - dumpfac * f0 / dfdW;
        continue;
      }
      break;
    }
    if (idump >= 100) return -101;
    W = Wnew;
    if (kzm_fabs(W) > 1.0 / 1.0e-80) return -103;
    if (kzm_fabs((W - Wprev) / Wprev) < conv && errr < 1.0e-1) break;
  }
  if (iter >= 50) return -102;
  if (!kzm_isfinite(W)) return -103;

  double wsq = W * W;
  double xsq = (c.Bsq + W) * (c.Bsq + W);
  double v2 = (wsq * c.Qtsq + c.QdotBsq * (c.Bsq + 2.0 * W)) / (wsq * xsq);
  double gamma2 = 1.0 / (1.0 - v2);
  double gamma = kzm_sqrt(gamma2);
  double rho = c.D / gamma;
  double entr = c.Sc / gamma;
  double uint = 1.0 / pg * (W / gamma2 - rho);
  double ut[4];
#pragma unroll
  for (int i = 1; i < 4; ++i) ut[i] = gamma / (W + c.Bsq) * (Qt[i] + QdotB * Bcon[i] / W);
  if (!kzm_isfinite(ut[1])) return -120;
  if (uint < 0.0 || gamma2 < 0.0 || W != W || !kzm_isfinite(W)) return -104;
  pp[RHO] = rho;
  pp[UU] = uint;
  pp[VX] = ut[1];
  pp[VY] = ut[2];
  pp[VZ] = ut[3];
  if (rho < 0.0) return -105;
  pp[ENTR] = entr;
  pp[B1] = uu[B1] * gi;
  pp[B2] = uu[B2] * gi;
  pp[B3] = uu[B3] * gi;
  return 0;
}

// invert.zig:422-467 u2pMhd; returns ret (0/-1/-2/-3), sets *entropy_used
KZD int u2p_mhd(const double *uu_in, double *pp, const Geo &gm, double gam, const kz_floor_params &f,
                int *entropy_used) {
  double uu[NV], pm[NV];
#pragma unroll
  for (int i = 0; i < NV; ++i) {
    uu[i] = uu_in[i];
    pm[i] = pp[i];
  }
  int ret = 0;
  *entropy_used = 0;
  if (uu[RHO] * gm.alpha / gm.gdet < 0.0) {
    pm[RHO] = f.rhofloor;
    uu[RHO] = f.rhofloor * gm.gdet / gm.alpha;
    ret = -2;
  }
  int r = u2p_solver_w(uu, pm, gm, 0, gam);
  if (r < 0) {
    ret = -1;
    *entropy_used = 1;
    r = u2p_solver_w(uu, pm, gm, 1, gam);
    if (r < 0) ret = -2;
  }
  if (r >= 0) {
#pragma unroll
    for (int i = 0; i < NV; ++i) pp[i] = pm[i];
  } else {
    ret = -3;  // pp keeps the incoming primitives (ppbak)
  }
  return ret;
}

// invert.zig:474-715 checkFloorsMhd; *guard counts the guard-ladder
// recoveries (n_guard_recoveries).  ZAMO selects the b^2/rho floor in the
// ZAMO frame (b2rhofloorframe 1) at compile time, so that the kernels of the
// default drift frame carry none of its code.
template <bool ZAMO = false>
KZD int check_floors_mhd(double *pp, const Geo &gm, double pg, const kz_floor_params &fp, int *guard) {
  int ret = 0;
  // invert.zig:485-492: the ZAMO floor adds its delta to the conserveds of
  // the primitives as they enter, before any floor
  double uu_entry[NV];
  if (ZAMO) p2u_mhd(pp, uu_entry, gm, pg);
  if (pp[RHO] < fp.rhofloor) {
    pp[RHO] = fp.rhofloor;
    ret = -1;
  }
  if (pp[UU] < fp.uurhoratiomin * pp[RHO]) {
    pp[UU] = fp.uurhoratiomin * pp[RHO];
    ret = -1;
  }
  if (pp[UU] > fp.uurhoratiomax * pp[RHO]) {
    pp[UU] = fp.uurhoratiomax * pp[RHO];
    ret = -1;
  }
  {
    double ucon[4], ucov[4];
    ucon_ucov(pp, gm, ucon, ucov);
    double bcon[4], bcov[4], bsq;
    bcon_bcov_bsq(pp[B1], pp[B2], pp[B3], ucon, ucov, gm, bcon, bcov, &bsq);
    double f = 1.0, fuu = 1.0;
    bool floored = false;
    if (bsq > fp.b2rhoratiomax * pp[RHO]) {
      f = bsq / (fp.b2rhoratiomax * pp[RHO]);
      floored = true;
    }
    if (!ZAMO && fp.b2rhofloorframe == 0 && fp.b2uufloor && bsq > fp.b2uuratiomax * pp[UU]) {
      fuu = bsq / (fp.b2uuratiomax * pp[UU]);
      floored = true;
    }
    if (ZAMO && floored) {
      // invert.zig:649-684: the new mass drho = rho (f - 1) moves with the
      // normal observer n^mu = -alpha g^{mu 0}; its conserved vector is added
      // to the entry conserveds through B3 and the primitives recovered by the
      // hot, then the entropy solver; if both fail the fluid-frame backup
      // scales rho and u by f
      double pporg[NV];
#pragma unroll
      for (int iv = 0; iv < NV; ++iv) pporg[iv] = pp[iv];
      const double drho = pp[RHO] * (f - 1.0);
      double etacon[4], etarel[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) etacon[i] = -gm.alpha * gm.G[i][0];
      vel4_to_velr(etacon, gm, true, etarel);
      double dpp[NV], duu[NV], uu[NV];
#pragma unroll
      for (int iv = 0; iv < NV; ++iv) dpp[iv] = 0.0;
      dpp[RHO] = drho;
      dpp[VX] = etarel[1];
      dpp[VY] = etarel[2];
      dpp[VZ] = etarel[3];
      p2u_mhd(dpp, duu, gm, pg);
#pragma unroll
      for (int iv = 0; iv < NV; ++iv) uu[iv] = iv <= B3 ? uu_entry[iv] + duu[iv] : 0.0;
      int rt = u2p_solver_w(uu, pp, gm, 0, pg);
      if (rt < 0) rt = u2p_solver_w(uu, pp, gm, 1, pg);
      if (rt < 0) {
#pragma unroll
        for (int iv = 0; iv < NV; ++iv) pp[iv] = pporg[iv];
        pp[RHO] *= f;
        pp[UU] *= f;
      }
      ret = -1;
    } else if (floored) {
      double wold = pp[RHO] + pp[UU] * pg;
      pp[RHO] *= f;
      pp[UU] *= fp.isentropic_b2rhofloors ? f : fuu;
      double wnew = pp[RHO] + pp[UU] * pg;
      bool guarded = false;
      if (gm.x1 >= fp.horizon_x1) {
        do {
          double Bcon[4] = {0, pp[B1], pp[B2], pp[B3]}, Bcov[4];
          lower4(Bcon, gm, Bcov);
          double Bsq = dot4(Bcon, Bcov);
          double Bmag = kzm_sqrt(Bsq);
          if (!(bsq > 1.0e-10 * wnew && Bsq > 1.0e-10 * wnew)) {
            guarded = true;
            break;
          }
          double betapar = -bcon[0] / (bsq * ucon[0]);
          double betasq = betapar * betapar * bsq;
          double betasqmax = 1.0 - 1.0 / (fp.gammamaxhd * fp.gammamaxhd);
          if (betasq > betasqmax) betasq = betasqmax;
          double gammapar = 1.0 / kzm_sqrt(1.0 - betasq);
          double ucondr[4];
#pragma unroll
          for (int iv = 0; iv < 4; ++iv) ucondr[iv] = gammapar * (ucon[iv] + betapar * bcon[iv]);
          if (!kzm_isfinite(ucondr[0]) || kzm_fabs(ucondr[0]) < 1.0e-10) {
            guarded = true;
            break;
          }
          double udotB = dot4(ucon, Bcov);
          double QdotB = udotB * wold * ucon[0];
          double xx = 2.0 * QdotB / (Bmag * wnew * ucondr[0]);
          if (!(xx <= 1.0e10)) {
            xx = 1.0e10;
            guarded = true;
          }
          if (!(xx >= -1.0e10)) {
            xx = -1.0e10;
            guarded = true;
          }
          double vpar = xx / (ucondr[0] * (1.0 + kzm_sqrt(1.0 + xx * xx)));
          double vcon[4];
          vcon[0] = 1.0;
#pragma unroll
          for (int iv = 1; iv < 4; ++iv) vcon[iv] = vpar * Bcon[iv] / Bmag + ucondr[iv] / ucondr[0];
          for (int idamp = 0; idamp < 50; ++idamp) {
            double bv = 0.0, cv = 0.0;
#pragma unroll
            for (int iv = 1; iv < 4; ++iv) {
              bv += vcon[iv] * gm.g[0][iv];
#pragma unroll
              for (int jv = 1; jv < 4; ++jv) cv += vcon[iv] * vcon[jv] * gm.g[iv][jv];
            }
            if (gm.g[0][0] + 2.0 * bv + cv < -1.0 / (fp.gammamaxhd * fp.gammamaxhd)) break;
            guarded = true;
            vpar *= 0.5;
#pragma unroll
            for (int iv = 1; iv < 4; ++iv) vcon[iv] = vpar * Bcon[iv] / Bmag + ucondr[iv] / ucondr[0];
          }
          double ucont[4];
          if (vel3_to_velr(vcon, gm, ucont) != E_OK) {
            guarded = true;
            break;
          }
          pp[VX] = ucont[1];
          pp[VY] = ucont[2];
          pp[VZ] = ucont[3];
        } while (0);
      }
      if (guarded) *guard += 1;
      ret = -1;
    }
  }
  {
    double qsq = 0;
#pragma unroll
    for (int i = 1; i < 4; ++i)
#pragma unroll
      for (int j = 1; j < 4; ++j) qsq += pp[VX - 1 + i] * pp[VX - 1 + j] * gm.g[i][j];
    double gamma2 = 1.0 + qsq;
    if (gamma2 > fp.gammamaxhd * fp.gammamaxhd) {
      double qsqmax = fp.gammamaxhd * fp.gammamaxhd - 1.0;
      double a = kzm_sqrt(qsqmax / qsq);
#pragma unroll
      for (int j = 1; j < 4; ++j) pp[VX - 1 + j] *= a;
      ret = -1;
    }
  }
  if (ret < 0) pp[ENTR] = s_from_u(pp[RHO], pp[UU], pg);
  return ret;
}

// ---- invert_rad.zig ----------------------------------------------------------
constexpr double GSL = 1.0 - 1e-10;

KZD double m1_gammarel2(const Geo &gm, const double av[4]) {  // :66-124
  double gn11 = gm.G[0][0], gn12 = gm.G[0][1], gn13 = gm.G[0][2], gn14 = gm.G[0][3];
  double gn22 = gm.G[1][1], gn23 = gm.G[1][2], gn24 = gm.G[1][3], gn33 = gm.G[2][2], gn34 = gm.G[2][3],
         gn44 = gm.G[3][3];
  double rdtt = av[0], rdtx = av[1], rdty = av[2], rdtz = av[3];
  double mix = gn12 * rdtx + gn13 * rdty + gn14 * rdtz;
  double sq = kzm_sqrt(4.0 * (gn11 * gn11) * (rdtt * rdtt) + mix * mix +
                       gn11 * (8.0 * gn12 * rdtt * rdtx + 3.0 * gn22 * (rdtx * rdtx) + 8.0 * gn13 * rdtt * rdty +
                               6.0 * gn23 * rdtx * rdty + 3.0 * gn33 * (rdty * rdty) + 8.0 * gn14 * rdtt * rdtz +
                               6.0 * gn24 * rdtx * rdtz + 6.0 * gn34 * rdty * rdtz + 3.0 * gn44 * (rdtz * rdtz)));
  double den = gn11 * (rdtt * rdtt) + 2.0 * gn12 * rdtt * rdtx + gn22 * (rdtx * rdtx) + 2.0 * gn13 * rdtt * rdty +
               2.0 * gn23 * rdtx * rdty + gn33 * (rdty * rdty) + 2.0 * (gn14 * rdtt + gn24 * rdtx + gn34 * rdty) * rdtz +
               gn44 * (rdtz * rdtz);
  double g2a = (-0.25 * (2.0 * (gn11 * gn11) * (rdtt * rdtt) + mix * (mix + sq) +
                         gn11 * (4.0 * gn12 * rdtt * rdtx + gn22 * (rdtx * rdtx) + 2.0 * gn23 * rdtx * rdty +
                                 gn33 * (rdty * rdty) + 2.0 * gn24 * rdtx * rdtz + 2.0 * gn34 * rdty * rdtz +
                                 gn44 * (rdtz * rdtz) + rdtt * (4.0 * gn13 * rdty + 4.0 * gn14 * rdtz + sq)))) /
               den;
  double g2;
  if (g2a < GSL || !kzm_isfinite(g2a)) {
    g2 = (0.25 * (-2.0 * (gn11 * gn11) * (rdtt * rdtt) -
                  1.0 * gn11 * (4.0 * gn12 * rdtt * rdtx + gn22 * (rdtx * rdtx) +
                                rdty * (4.0 * gn13 * rdtt + 2.0 * gn23 * rdtx + gn33 * rdty) +
                                2.0 * (2.0 * gn14 * rdtt + gn24 * rdtx + gn34 * rdty) * rdtz + gn44 * (rdtz * rdtz)) +
                  gn11 * rdtt * sq + mix * (-1.0 * gn12 * rdtx - 1.0 * gn13 * rdty - 1.0 * gn14 * rdtz + sq))) /
         den;
  } else {
    g2 = g2a;
  }
  double gr2 = g2 * gm.alpha * gm.alpha;
  if (gr2 > GSL && gr2 < 1.0) gr2 = 1.0;
  return gr2;
}

KZD void m1_cold(const Geo &gm, const double in[4], double gr2, double av[4], double *erf_o, double ur[4]) {
  double gn11 = gm.G[0][0], gn12 = gm.G[0][1], gn13 = gm.G[0][2], gn14 = gm.G[0][3];  // :140-185
  double gn22 = gm.G[1][1], gn23 = gm.G[1][2], gn24 = gm.G[1][3], gn33 = gm.G[2][2], gn34 = gm.G[2][3],
         gn44 = gm.G[3][3];
  double alpha = gm.alpha, rdtx = in[1], rdty = in[2], rdtz = in[3];
  double utsq = gr2 / (alpha * alpha);
#pragma unroll
  for (int i = 0; i < 4; ++i) av[i] = in[i];
  double mix = gn13 * rdty + gn14 * rdtz;
  double disc = kzm_sqrt(((gn12 * gn12) * (rdtx * rdtx) + 2.0 * gn12 * rdtx * mix + mix * mix -
                          1.0 * gn11 * (gn22 * (rdtx * rdtx) + 2.0 * gn23 * rdtx * rdty + gn33 * (rdty * rdty) +
                                        2.0 * gn24 * rdtx * rdtz + 2.0 * gn34 * rdty * rdtz + gn44 * (rdtz * rdtz))) *
                         utsq * (gn11 + utsq) * ((gn11 + 4.0 * utsq) * (gn11 + 4.0 * utsq)));
  av[0] = (0.25 * (-4.0 * (gn12 * rdtx + gn13 * rdty + gn14 * rdtz) * utsq * (gn11 + utsq) + disc)) /
          (gn11 * utsq * (gn11 + utsq));
  double erf = (0.75 * disc) / (utsq * (gn11 + utsq) * (gn11 + 4.0 * utsq));
  double avc[4];
  raise4(av, gm, avc);
  double gr = kzm_sqrt(gr2);
#pragma unroll
  for (int i = 0; i < 4; ++i) ur[i] = 0;
  if (erf > 0.0) {
#pragma unroll
    for (int jj = 1; jj < 4; ++jj)
      ur[jj] = alpha * (avc[jj] + 1.0 / 3.0 * erf * gm.G[0][jj] * (4.0 * gr2 - 1.0)) / (4.0 / 3.0 * erf * gr);
  }
  *erf_o = erf;
}

// invert_rad.zig:189-261 u2pRad; returns ret, *corrected
KZD int u2p_rad(const double *uu, double *pp, const Geo &gm, const kz_rad_params &rp, int *corrected) {
  double av[4] = {uu[EE] / gm.gdet, uu[FX] / gm.gdet, uu[FY] / gm.gdet, uu[FZ] / gm.gdet}, avc[4];
  raise4(av, gm, avc);
  double gr2 = m1_gammarel2(gm, av);
  double erf = 3.0 * avc[0] * gm.alpha * gm.alpha / (4.0 * gr2 - 1.0);
  double gmax = rp.gammamaxrad;
  bool ok = gr2 >= 1.0 && erf > rp.eradfloor && gr2 <= gmax * gmax / GSL / GSL;
  double uc[4];
  *corrected = 0;
  if (ok) {
    double gr = kzm_sqrt(gr2);
#pragma unroll
    for (int ii = 0; ii < 4; ++ii)
      uc[ii] = gm.alpha * (avc[ii] + 1.0 / 3.0 * erf * gm.G[0][ii] * (4.0 * gr2 - 1.0)) / (4.0 / 3.0 * erf * gr);
  } else {
    const double one_eps = 1.0 + 10.0 * 2.220446049250313080847e-16;
    double sa[4], se, su[4], fa[4], fe, fu[4];
    m1_cold(gm, av, one_eps * one_eps, sa, &se, su);
    m1_cold(gm, av, gmax * gmax, fa, &fe, fu);
    if (kzm_fabs(sa[0] - av[0]) > kzm_fabs(fa[0] - av[0])) {
#pragma unroll
      for (int i = 0; i < 4; ++i) uc[i] = fu[i];
      gr2 = gmax * gmax;
      erf = fe;
    } else {
#pragma unroll
      for (int i = 0; i < 4; ++i) uc[i] = su[i];
      gr2 = one_eps * one_eps;
      erf = se;
    }
    *corrected = 1;
    if (!kzm_isfinite(erf) || !kzm_isfinite(gr2) || !kzm_isfinite(uc[1]) || !kzm_isfinite(uc[2]) ||
        !kzm_isfinite(uc[3]))
      return -1;
  }
  pp[EE] = erf;
  pp[FX] = uc[1];
  pp[FY] = uc[2];
  pp[FZ] = uc[3];
  return 0;
}

// invert_rad.zig:266-303 checkFloorsRad
KZD int check_floors_rad(double *pp, const Geo &gm, const kz_rad_params &rp) {
  int ret = 0;
  double ehat = -calc_ff_rtt(pp, gm);
  double eratio = pp[EE] / ehat;
  if (pp[EE] < rp.eradfloor) {
    pp[EE] = rp.eradfloor;
    ret = -1;
  }
  if (ehat < rp.eerhoratiomin * pp[RHO]) {
    pp[EE] = eratio * rp.eerhoratiomin * pp[RHO];
    ret = -1;
  }
  if (ehat > rp.eerhoratiomax * pp[RHO]) {
    pp[RHO] = 1.0 / rp.eerhoratiomax * ehat;
    ret = -1;
  }
  if (ehat < rp.eeuuratiomin * pp[UU]) {
    pp[EE] = eratio * rp.eeuuratiomin * pp[UU];
    ret = -1;
  }
  if (ehat > rp.eeuuratiomax * pp[UU]) {
    pp[UU] = 1.0 / rp.eeuuratiomax * ehat;
    ret = -1;
  }
  return ret;
}

// ---- implicit.zig --------------------------------------------------------------
KZD double my_sign(double x) { return x > 0.0 ? 1.0 : (x < 0.0 ? -1.0 : 1.0); }  // :167-172

KZD double abs_t(double x) { return kzm_fabs(x); }
KZD float abs_t(float x) { return __builtin_fabsf(x); }
// implicit.zig:176-224 invert4, on doubles (the reference) or on floats (the
// mixed-precision Newton's LU; same pivoting and elimination order)
template <class T>
KZD void invert4_t(const T a[4][4], T inv[4][4]) {
  T lu[4][4];
  int perm[4] = {0, 1, 2, 3};
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) lu[i][j] = a[i][j];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    int piv = k;
    T pmax = abs_t(lu[k][k]);
#pragma unroll
    for (int r = k + 1; r < 4; ++r)
      if (abs_t(lu[r][k]) > pmax) {
        pmax = abs_t(lu[r][k]);
        piv = r;
      }
    // row swap expressed as selects so the matrix stays in registers
#pragma unroll
    for (int r = k + 1; r < 4; ++r) {
      bool sw = piv == r;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        T tk = lu[k][j], tr = lu[r][j];
        lu[k][j] = sw ? tr : tk;
        lu[r][j] = sw ? tk : tr;
      }
      int pk = perm[k], pr = perm[r];
      perm[k] = sw ? pr : pk;
      perm[r] = sw ? pk : pr;
    }
#pragma unroll
    for (int r = k + 1; r < 4; ++r) {
      lu[r][k] /= lu[k][k];
#pragma unroll
      for (int cc = k + 1; cc < 4; ++cc) lu[r][cc] -= lu[r][k] * lu[k][cc];
    }
  }
#pragma unroll
  for (int col = 0; col < 4; ++col) {
    T y[4], x[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      y[r] = (perm[r] == col) ? T(1) : T(0);
#pragma unroll
      for (int cc = 0; cc < r; ++cc) y[r] -= lu[r][cc] * y[cc];
    }
#pragma unroll
    for (int r = 3; r >= 0; --r) {
      x[r] = y[r];
#pragma unroll
      for (int cc = r + 1; cc < 4; ++cc) x[r] -= lu[r][cc] * x[cc];
      x[r] /= lu[r][r];
    }
#pragma unroll
    for (int rr = 0; rr < 4; ++rr) inv[rr][col] = x[rr];
  }
}
KZD void invert4(const double a[4][4], double inv[4][4]) { invert4_t<double>(a, inv); }

struct ImpCtx {
  const Geo *gm;
  double gam;
  const kz_rad_params *rp;
  const kz_opac_params *op;
  const kz_implicit_params *ip;
  double opdamp;
};

// implicit.zig:240-306 applyConstraints.  WP (the primitive block the
// Newton solves for) is a template parameter so that every state index is
// a compile-time constant and the 13-slot states stay in registers.
template <int WP>
KZD int apply_constraints(double *pp, double *uu, const double *uu0, const ImpCtx &cx) {
  const Geo &gm = *cx.gm;
  double gdetu = gm.gdet, gi = 1.0 / gdetu;
  if (WP == WP_MHD) {
    double ucon[4], ucov[4];
    ucon_ucov(pp, gm, ucon, ucov);
    pp[RHO] = uu0[RHO] * gi / ucon[0];
    pp[ENTR] = s_from_u(pp[RHO], pp[UU], cx.gam);
    p2u_mhd(pp, uu, gm, cx.gam);
#pragma unroll
    for (int k = 0; k < 4; ++k) uu[EE + k] = uu0[EE + k] - (uu[UU + k] - uu0[UU + k]);
    int corr;
    int r = u2p_rad(uu, pp, gm, *cx.rp, &corr);
    if (corr) r = cx.ip->allow_rad_ceiling ? -1 : -2;
    return r;
  }
  p2u_rad(pp, uu, gm);
  uu[RHO] = uu0[RHO];
#pragma unroll
  for (int k = 0; k < 4; ++k) uu[UU + k] = uu0[UU + k] - (uu[EE + k] - uu0[EE + k]);
#pragma unroll
  for (int k = 0; k < 3; ++k) uu[B1 + k] = gdetu * pp[B1 + k];
  int rt = u2p_solver_w(uu, pp, gm, 0, cx.gam);
  if (rt < 0 && cx.ip->allow_entr_in_4dprim) rt = u2p_solver_w(uu, pp, gm, 1, cx.gam);
  int ret = rt < 0 ? -2 : 0;
  double ucon[4], ucov[4];
  ucon_ucov(pp, gm, ucon, ucov);
  pp[ENTR] = s_from_u(pp[RHO], pp[UU], cx.gam);
  uu[ENTR] = pp[ENTR] * gdetu * ucon[0];
  return ret;
}

// reference-state quantities the residual reads (implicit.zig:345-347,377,397)
struct State0 {
  double ehat, uint, sgas;
};
// the reference-state opacities lagOpac substitutes (implicit.zig:631-640)
struct LagOpac {
  double kappaes;
  Opac op;
};

// implicit.zig:337-427 residualG (scalar) on a filled state
template <int WP>
KZD double residual(const double *uu, const double *pp, const RadState &st, const double *uu0, const State0 &s0,
                    double dt, const ImpCtx &cx, int we, int wf, double f[4]) {
  const Geo &gm = *cx.gm;
  double gdetu = gm.gdet;
  double dtau = dt / st.ucon[0];
  double giff0, gilab[4], gi[4];
  calc_gi(st, *cx.op, cx.opdamp, &giff0, gilab);
  lower4(gilab, gm, gi);
  double err[4] = {0, 0, 0, 0};
  const double SMALLI = 1.0e-80;
  constexpr int base = (WP == WP_MHD) ? UU : EE;
#pragma unroll
  for (int k = 1; k < 4; ++k) {
    // MHD: uu - uu0 - dt*gdet*G;  RAD: uu - uu0 + dt*gdet*G
    f[k] = (WP == WP_MHD) ? uu[base + k] - uu0[base + k] - dt * gdetu * gi[k]
                          : uu[base + k] - uu0[base + k] + dt * gdetu * gi[k];
    err[k] = (kzm_fabs(f[k]) > SMALLI) ? kzm_fabs(f[k]) / (1.0e-20 * uu[base] + kzm_fabs(uu[base + k]) +
                                                           kzm_fabs(uu0[base + k]) + kzm_fabs(dt * gdetu * gi[k]))
                                       : 0.0;
  }
  if (wf == WF_LAB) {
    if (WP == WP_RAD) {
      f[0] = uu[EE] - uu0[EE] + dt * gdetu * gi[0];
      err[0] = (kzm_fabs(f[0]) > SMALLI)
                   ? kzm_fabs(f[0]) / (kzm_fabs(uu[EE]) + kzm_fabs(uu0[EE]) + kzm_fabs(dt * gdetu * gi[0]))
                   : 0.0;
    } else {
      f[0] = uu[UU] - uu0[UU] - dt * gdetu * gi[0];
      err[0] = (kzm_fabs(f[0]) > SMALLI)
                   ? kzm_fabs(f[0]) / (kzm_fabs(uu[UU]) + kzm_fabs(uu0[UU]) + kzm_fabs(dt * gdetu * gi[0]))
                   : 0.0;
    }
  } else if (we == WE_ENERGY) {
    if (WP == WP_RAD) {
      f[0] = st.ehat - s0.ehat + dtau * giff0;
      err[0] = kzm_fabs(f[0]) / (kzm_fabs(st.ehat) + kzm_fabs(s0.ehat) + kzm_fabs(dtau * giff0));
    } else {
      f[0] = pp[UU] - s0.uint - dtau * giff0;
      err[0] = kzm_fabs(f[0]) / (kzm_fabs(pp[UU]) + kzm_fabs(s0.uint) + kzm_fabs(dtau * giff0));
    }
  } else {
    f[0] = st.tgas * (st.sgas - s0.sgas) - dtau * giff0;
    err[0] = kzm_fabs(f[0]) / (kzm_fabs(st.tgas * st.sgas) + kzm_fabs(st.tgas * s0.sgas) + kzm_fabs(dtau * giff0));
  }
  double e0 = 0;
#pragma unroll
  for (int k = 0; k < 4; ++k) e0 = err[k] > e0 ? err[k] : e0;
  return e0;
}

// The residual at pp: fill the radiation state (slim for the energy
// equations, full for entropy), apply lagOpac, evaluate.  The reference
// fills the state right after each applyConstraints and consumes it only
// here, with pp unchanged in between, so filling at the point of use gives
// the same values and keeps the state a temporary.
template <int WP>
KZD double eval_residual(const double *uu, const double *pp, const LagOpac &lag, const double *uu0,
                         const State0 &s0, double dt, const ImpCtx &cx, int we, int wf, double f[4]) {
  RadState st;
  fill_radstate(pp, *cx.gm, cx.gam, *cx.op, we == WE_ENTROPY, st);
  if (cx.ip->lag_opac) {
    st.kappaes = lag.kappaes;
    st.op = lag.op;
  }
  return residual<WP>(uu, pp, st, uu0, s0, dt, cx, we, wf, f);
}

// implicit.zig:431-467 f1dErr (energy equation, undamped opacities)
KZD double f1d_err(double xxx, const double *pp0, double dtau, const ImpCtx &cx, int wp, double totenergy,
                   double ratio) {
  double ugas, ehat;
  if (wp == WP_RAD) {
    ehat = xxx * ratio;
    ugas = totenergy - ehat;
  } else {
    ugas = xxx;
    ehat = totenergy - ugas;
  }
  double pp[NV];
#pragma unroll
  for (int i = 0; i < NV; ++i) pp[i] = pp0[i];
  pp[UU] = ugas;
  pp[EE] = ehat / ratio;
  RadState st;
  fill_radstate(pp, *cx.gm, cx.gam, *cx.op, false, st);
  double giff0, gilab[4];
  calc_gi(st, *cx.op, cx.op->opdamp, &giff0, gilab);
  if (wp == WP_RAD) return ehat - pp0[EE] / ratio + dtau * giff0;
  return ugas - pp0[UU] - dtau * giff0;
}

// implicit.zig:473-574 solve1dPrim: bisection warm start on the fluid-frame
// energy equation; reads Ehat, u^t and T_gas of the reference state.
KZD void solve_1d_prim(const double *pp00, double st_ehat, double st_ut, double st_tgas, double dt, const ImpCtx &cx,
                       double *pp0out) {
#pragma unroll
  for (int i = 0; i < NV; ++i) pp0out[i] = pp00[i];
  double ehat = st_ehat;
  double dtau = dt / st_ut;
  double tot = ehat + pp00[UU];
  int wp = (ehat < 1.0e-3 * pp00[UU]) ? WP_RAD : WP_MHD;
  double tgas = st_tgas;
  double ratio = ehat / pp00[EE];
  const double fac = 3.0, fac1 = fac + 1.0;
  double pu = pp00[UU], pe = pp00[EE];
  if (tgas > 1.0e14) {
    pu = pp00[UU] * 1.0e14 / tgas;
    pe = (tot - pu) / ratio;
  }
  double xlo, xhi;
  if (wp == WP_MHD) {
    xlo = pu / 1.05;
    xhi = kzm_fmin(pu * 1.05, (pu + fac * tot) / fac1);
  } else {
    xlo = pe / 1.05;
    xhi = kzm_fmin(pe * 1.05, (pe + fac * tot / ratio) / fac1);
  }
  double elo = f1d_err(xlo, pp00, dtau, cx, wp, tot, ratio);
  double ehi = f1d_err(xhi, pp00, dtau, cx, wp, tot, ratio);
  int iter = 0;
  while (elo * ehi > 0.0) {
    iter += 1;
    if (elo > 0.0) {
      xhi = xlo;
      ehi = elo;
      xlo /= fac;
      elo = f1d_err(xlo, pp00, dtau, cx, wp, tot, ratio);
    } else {
      xlo = xhi;
      elo = ehi;
      if (wp == WP_RAD) xhi = kzm_fmin(xhi * fac, (xhi + fac * tot / ratio) / fac1);
      else xhi = kzm_fmin(xhi * fac, (xhi + fac * tot) / fac1);
      ehi = f1d_err(xhi, pp00, dtau, cx, wp, tot, ratio);
    }
    if (elo != elo || ehi != ehi) iter = 50;
    if (iter >= 50) break;
  }
  if (iter >= 50) return;
  double xmid;
  for (;;) {
    iter += 1;
    xmid = 0.5 * (xlo + xhi);
    double emid = f1d_err(xmid, pp00, dtau, cx, wp, tot, ratio);
    if (emid * elo > 0.0) {
      xlo = xmid;
      elo = emid;
    } else {
      xhi = xmid;
      ehi = emid;
    }
    if (xmid != xmid || emid != emid) return;
    if (!(kzm_fabs((xhi - xlo) / xmid) > 1.0e-3)) break;
  }
  if (wp == WP_RAD) {
    pp0out[EE] = xmid;
    pp0out[UU] = tot - xmid * ratio;
  } else {
    pp0out[UU] = xmid;
    pp0out[EE] = (tot - xmid) / ratio;
  }
}

// Element j (0..3, not a compile-time constant) of a 4-block, by selects,
// so that register-resident arrays are never indexed dynamically.
KZD double get4(const double *a, int j) { return j == 0 ? a[0] : (j == 1 ? a[1] : (j == 2 ? a[2] : a[3])); }
KZD void set4(double *a, int j, double v) {
  if (j == 0) a[0] = v;
  else if (j == 1) a[1] = v;
  else if (j == 2) a[2] = v;
  else a[3] = v;
}
KZD double gdiag(const Geo &gm, int j) { return j == 1 ? gm.g[1][1] : (j == 2 ? gm.g[2][2] : gm.g[3][3]); }

// implicit.zig:586-591 fdDel on the block b = ppp + sh
KZD double fd_del(int j, double sign, const double *b, double eps, const Geo &gm) {
  if (j == 0) return sign * eps * b[0];
  double bj = get4(b, j);
  double veleps = eps * my_sign(bj) * kzm_fmax(1.0e-8 / kzm_sqrt(gdiag(gm, j)), kzm_fabs(bj));
  return bj >= 0.0 ? sign * veleps : -sign * veleps;
}

// implicit.zig:642-922 solve4dPrimCore (scalar Jacobian), as a start and a
// one-iteration step so that a persistent kernel can interleave cells at
// different iterations: newton_step runs one pass of the reference's
// iteration loop and returns 0 to go on, 1 when the loop ends solved (n.pp
// and n.iter hold the result) and -1 when it ends failed.
struct NewtonState {
  double uu[NV], pp[NV];
  int iter, fret;
};
template <int WP>
KZD void newton_start(NewtonState &n, const double *uu00, const double *pp0in, const ImpCtx &cx) {
#pragma unroll
  for (int i = 0; i < NV; ++i) {
    n.uu[i] = uu00[i];
    n.pp[i] = pp0in[i];
  }
  n.fret = apply_constraints<WP>(n.pp, n.uu, uu00, cx);
  n.iter = 0;
}
template <int WP>
KZD int newton_step(NewtonState &n, const double *uu00, const LagOpac &lag, const State0 &s0, double dt,
                    const ImpCtx &cx, int we, int wf) {
  const kz_implicit_params &ip = *cx.ip;
  const Geo &gm = *cx.gm;
  double *uu = n.uu, *pp = n.pp;
  double ppp[NV];
  constexpr int sh = (WP == WP_MHD) ? UU : EE;
  const double conv = (we == WE_ENTROPY) ? ip.entrconv : ip.conv;
  double f1[4], f2[4], jac[4][4];
  bool failed = false;
  int &fret = n.fret;
  const int iter = ++n.iter;
#pragma unroll
  for (int i = 0; i < NV; ++i) ppp[i] = pp[i];
  if (fret < -1) return -1;
  double errbase = eval_residual<WP>(uu, pp, lag, uu00, s0, dt, cx, we, wf, f1);
  if (errbase < conv) return iter > ip.maxiter ? -1 : 1;
#pragma unroll 1
  for (int j = 0; j < 4; ++j) {
    double sign = -1.0;
    for (;;) {
      double del = fd_del(j, sign, ppp + sh, ip.eps, gm);
      set4(pp + sh, j, get4(ppp + sh, j) + del);
      fret = apply_constraints<WP>(pp, uu, uu00, cx);
      if (fret < -1) {
        if (sign > 0.0) {
          failed = true;
          break;
        }
        set4(pp + sh, j, get4(ppp + sh, j));
        sign *= -1.0;
        continue;
      }
      eval_residual<WP>(uu, pp, lag, uu00, s0, dt, cx, we, wf, f2);
#pragma unroll
      for (int i = 0; i < 4; ++i) set4(jac[i], j, (f2[i] - f1[i]) / del);
      set4(pp + sh, j, get4(ppp + sh, j));
      break;
    }
    if (failed) break;
  }
  if (failed) return -1;
  double sv[4] = {pp[sh], 1.0, 1.0, 1.0}, isv[4];
  if (ip.scale_jacobian) {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      if (sv[j] == 0.0) sv[j] = 1.0;
      isv[j] = 1.0 / sv[j];
    }
  }
  double ij[4][4];
  if (ip.jacobian_fp32) {
    // mixed precision: FP64 quotients rounded to FP32, scaled and
    // factored in FP32, the inverse applied in FP64
    float jf[4][4], ijf[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) jf[i][j] = (float)jac[i][j];
    if (ip.scale_jacobian) {
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
          if (i != j) jf[i][j] = jf[i][j] * ((float)sv[j] * (float)isv[i]);
    }
    invert4_t<float>(jf, ijf);
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) ij[i][j] = (double)ijf[i][j];
  } else {
    if (ip.scale_jacobian) {
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
          if (i != j) jac[i][j] = jac[i][j] * (sv[j] * isv[i]);
    }
    invert4(jac, ij);
  }
  double xiapp = 1.0;
  double xxx[4];
  for (;;) {
#pragma unroll
    for (int k = 0; k < 4; ++k) xxx[k] = ppp[k + sh];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        if (ip.scale_jacobian) xxx[i] -= xiapp * (ij[i][j] * sv[i] * isv[j]) * f1[j];
        else xxx[i] -= xiapp * ij[i][j] * f1[j];
      }
#pragma unroll
    for (int k = 0; k < 4; ++k) pp[k + sh] = xxx[k];
    fret = apply_constraints<WP>(pp, uu, uu00, cx);
    bool ok = false;
    double edenmin = ppp[sh] / ip.max_en_change_down;
    double edenmax = ppp[sh] * ip.max_en_change_up;
    if (xxx[0] > edenmin && xxx[0] < edenmax && fret >= -1) ok = true;
    double tgo = t_from_urho(*cx.op, ppp[UU], ppp[RHO], cx.gam);
    double tgn = t_from_urho(*cx.op, pp[UU], pp[RHO], cx.gam);
    if (ok && (tgn > tgo * ip.max_tgas_change || tgn < tgo / ip.max_tgas_change)) ok = false;
    double trn = lte_t_from_e(*cx.op, pp[EE]);
    double tro = lte_t_from_e(*cx.op, ppp[EE]);
    if (ok && (trn > tro * ip.max_trad_change || trn < tro / ip.max_trad_change)) ok = false;
    if (ok) break;
    if (xxx[0] <= 0.0) {
      xiapp *= ppp[sh] / (ppp[sh] + kzm_fabs(xxx[0]));
      xiapp *= 1.0e-1;
    } else {
      xiapp /= ip.damping_factor;
    }
    if (xiapp < ip.max_damping) {
      failed = true;
      break;
    }
  }
  if (!failed) {
    double f3[4];
    f3[0] = kzm_fabs((pp[sh] - ppp[sh]) / ppp[sh]);
#pragma unroll
    for (int i = 1; i < 4; ++i) {
      double d = pp[i + sh] - ppp[i + sh];
      f3[i] = kzm_fabs(d / kzm_fmax(ip.eps, kzm_fabs(ppp[i + sh])));
    }
    double cr = (we == WE_ENTROPY) ? ip.convrelentr : ip.convrel;
    double cre = (we == WE_ENTROPY) ? ip.convrelentrerr : ip.convrelerr;
    bool cc = true;
#pragma unroll
    for (int ib = 0; ib < 4; ++ib)
      if (f3[ib] > cr) cc = false;
    if (cc || errbase <= cre) return iter > ip.maxiter ? -1 : 1;
  }
  if (iter > ip.maxiter || failed) return -1;
  return 0;
}
template <int WP>
__device__ static __forceinline__ bool solve_4d_prim_core(const double *uu00, const LagOpac &lag, const State0 &s0,
                                                          const double *pp0in, double dt, const ImpCtx &cx, int we,
                                                          int wf, double *pp_out, int *iters_out) {
  NewtonState n;
  newton_start<WP>(n, uu00, pp0in, cx);
  int r;
  do {
    r = newton_step<WP>(n, uu00, lag, s0, dt, cx, we, wf);
  } while (r == 0);
  if (r < 0) return false;
#pragma unroll
  for (int i = 0; i < NV; ++i) pp_out[i] = n.pp[i];
  *iters_out = n.iter;
  return true;
}

// implicit.zig:927-1004 solveImplicitLab; returns kz_implicit_result
KZD kz_implicit_result solve_implicit_lab(double *uu, double *pp, const Geo &gm, double dt, double gam,
                                          const kz_rad_params &rp, const kz_opac_params &op,
                                          const kz_implicit_params &ip) {
  kz_implicit_result fail = {0, -1, 0, WP_RAD, WE_ENERGY, WF_LAB};
  ImpCtx cx{&gm, gam, &rp, &op, &ip, 1.0};
  double pp00[NV], uu00[NV];
#pragma unroll
  for (int i = 0; i < NV; ++i) pp00[i] = pp[i];
  p2u(pp00, uu00, gm, gam);
  double ehat = -calc_ff_rtt(pp00, gm);
  State0 s0;
  LagOpac lag;
  double pp0[NV];
  {
    RadState st00;
    fill_radstate(pp00, gm, gam, op, true, st00);
    s0 = State0{st00.ehat, st00.uint, st00.sgas};
    lag.kappaes = st00.kappaes;
    lag.op = st00.op;
    if (ip.start_with_bisect) {
      solve_1d_prim(pp00, st00.ehat, st00.ucon[0], st00.tgas, dt, cx, pp0);
    } else {
#pragma unroll
      for (int i = 0; i < NV; ++i) pp0[i] = pp00[i];
    }
  }
  const int sw = (ehat < ip.threshold * pp00[UU]) ? WP_RAD : WP_MHD;
  const int ot = (sw == WP_RAD) ? WP_MHD : WP_RAD;
  for (int level = 0; level <= ip.opdamp_maxlevels; ++level) {
    double od = 1.0;
    for (int i = 0; i < level; ++i) od /= ip.opdamp_factor;
    cx.opdamp = od;
    for (int ir = 0; ir < 6; ++ir) {
      int wp = (ir == 0 || ir == 1 || ir == 4) ? sw : ot;
      int we = (ir >= 4) ? WE_ENTROPY : WE_ENERGY;
      int wf = (ir == 0 || ir == 2) ? WF_LAB : WF_FF;
      double ppw[NV];
      int iters = 0;
      bool ok = (wp == WP_MHD) ? solve_4d_prim_core<WP_MHD>(uu00, lag, s0, pp0, dt, cx, we, wf, ppw, &iters)
                               : solve_4d_prim_core<WP_RAD>(uu00, lag, s0, pp0, dt, cx, we, wf, ppw, &iters);
      if (ok) {
#pragma unroll
        for (int i = 0; i < NV; ++i) pp[i] = ppw[i];
        p2u(ppw, uu, gm, gam);
        kz_implicit_result r = {1, ir, iters, wp, we, wf};
        return r;
      }
    }
  }
  return fail;
}

}  // namespace kzd
