// kz_host.cpp — caller-side C++ for the koral step: everything koral-zig
// computes once on the host before handing state to the per-step hot path.
//
//   * MetricCache (koral/metric/precompute.zig:121-526): MKS2 metric blocks at
//     cell centres and faces, Christoffels with the sqrt(-g) trace correction
//     (precompute.zig:30-73), built with forward-mode dual numbers
//     (koral/math/dual.zig) and cofactor det/inverse (koral/math/linalg.zig).
//   * coordinate maps and Jacobians (koral/metric/coco.zig).
//   * tables for the PUFFY outer-radial boundary (puffy.zig:1089-1166):
//     BL geometry and the MKS2<->BL Jacobians at every xhi ghost cell.
//   * the PUFFY limotorus initializer (puffy.zig:330-808), calc_BfromA
//     (koral/sim/ct.zig:184-334) and the seeded init perturbation.
//
// Host transcendentals are the C library's (this code runs once per run);
// the per-step physics lives in the HIP kernels.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "../../include/kz_types.h"

#define KZM_QUAL static inline
#include "../csrc/kz_math.h"

namespace {

constexpr double PI = 3.14159265358979323846264338327950288;

// ---- Dual(3) (koral/math/dual.zig:19-120) --------------------------------
struct D3 {
  double v;
  double d[3];
};
inline D3 dc(double x) { return D3{x, {0, 0, 0}}; }
inline D3 dvar(double x, int s) {
  D3 r = dc(x);
  r.d[s] = 1;
  return r;
}
inline D3 add(D3 a, D3 b) { return D3{a.v + b.v, {a.d[0] + b.d[0], a.d[1] + b.d[1], a.d[2] + b.d[2]}}; }
inline D3 sub(D3 a, D3 b) { return D3{a.v - b.v, {a.d[0] - b.d[0], a.d[1] - b.d[1], a.d[2] - b.d[2]}}; }
inline D3 mul(D3 a, D3 b) {
  D3 r{a.v * b.v, {0, 0, 0}};
  for (int i = 0; i < 3; ++i) r.d[i] = a.d[i] * b.v + a.v * b.d[i];
  return r;
}
inline D3 dvd(D3 a, D3 b) {
  double inv = 1.0 / b.v;
  double q = a.v * inv;
  D3 r{q, {0, 0, 0}};
  for (int i = 0; i < 3; ++i) r.d[i] = (a.d[i] - q * b.d[i]) * inv;
  return r;
}
inline D3 scale(D3 a, double c) { return D3{a.v * c, {a.d[0] * c, a.d[1] * c, a.d[2] * c}}; }
inline D3 addc(D3 a, double c) { return D3{a.v + c, {a.d[0], a.d[1], a.d[2]}}; }
inline D3 sq(D3 a) { return mul(a, a); }
inline D3 chain(D3 a, double v, double dv) { return D3{v, {a.d[0] * dv, a.d[1] * dv, a.d[2] * dv}}; }
inline D3 dsin(D3 a) { return chain(a, std::sin(a.v), std::cos(a.v)); }
inline D3 dcos(D3 a) { return chain(a, std::cos(a.v), -std::sin(a.v)); }
inline D3 dtan(D3 a) {
  double t = std::tan(a.v);
  return chain(a, t, 1.0 + t * t);
}
inline D3 dexp(D3 a) {
  double e = std::exp(a.v);
  return chain(a, e, e);
}

struct MP {
  double a, r0, h0;
};

// ---- metric forms (koral/metric/forms.zig:50-141) ------------------------
typedef D3 M4[4][4];

void gcov_ks(const D3 x[4], const MP &mp, M4 g) {
  double a = mp.a;
  D3 r = x[1];
  D3 sth = dsin(x[2]), cth = dcos(x[2]);
  D3 s2 = sq(sth);
  D3 sigma = add(sq(r), scale(sq(cth), a * a));
  D3 two_r_sigma = dvd(scale(r, 2.0), sigma);
  D3 one_p = addc(two_r_sigma, 1.0);
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) g[i][j] = dc(0);
  g[0][0] = addc(two_r_sigma, -1.0);
  g[0][1] = two_r_sigma;
  g[0][3] = dvd(scale(mul(r, s2), -2.0 * a), sigma);
  g[1][1] = one_p;
  g[1][3] = scale(mul(one_p, s2), -a);
  g[2][2] = sigma;
  g[3][3] = mul(s2, add(sigma, scale(mul(one_p, s2), a * a)));
  g[1][0] = g[0][1];
  g[3][0] = g[0][3];
  g[3][1] = g[1][3];
}

void gcov_bl(const D3 x[4], const MP &mp, M4 g) {
  double a = mp.a;
  D3 r = x[1];
  D3 sth = dsin(x[2]), cth = dcos(x[2]);
  D3 s2 = sq(sth);
  D3 sigma = add(sq(r), scale(sq(cth), a * a));
  D3 delta = addc(mul(addc(r, -2.0), r), a * a);
  D3 two_r_sigma = dvd(scale(r, 2.0), sigma);
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) g[i][j] = dc(0);
  g[0][0] = addc(two_r_sigma, -1.0);
  g[0][3] = dvd(scale(mul(r, s2), -2.0 * a), sigma);
  g[1][1] = dvd(sigma, delta);
  g[2][2] = sigma;
  g[3][3] = mul(s2, add(addc(sq(r), a * a), dvd(scale(mul(r, s2), 2.0 * a * a), sigma)));
  g[3][0] = g[0][3];
}

double mks2_cot(double h0) { return 1.0 / std::tan(0.5 * PI * h0); }
D3 mks2_theta(D3 x2, double h0) {
  D3 u = scale(addc(x2, -0.5), h0 * PI);
  return addc(scale(dtan(u), 0.5 * PI * mks2_cot(h0)), 0.5 * PI);
}
D3 mks2_theta_metric(D3 x2, double h0) {
  D3 u = scale(addc(x2, -0.5), h0 * PI);
  return scale(addc(scale(dtan(u), mks2_cot(h0)), 1.0), 0.5 * PI);
}
D3 mks2_dtheta_dx2_metric(D3 x2, double h0) {
  D3 u = scale(addc(x2, -0.5), h0 * PI);
  D3 t = dtan(u);
  return scale(addc(sq(t), 1.0), 0.5 * h0 * PI * PI * mks2_cot(h0));
}

void gcov_mks2(const D3 x[4], const MP &mp, M4 g) {
  D3 j1 = dexp(x[1]);
  D3 r = addc(j1, mp.r0);
  D3 th = mks2_theta_metric(x[2], mp.h0);
  D3 j2 = mks2_dtheta_dx2_metric(x[2], mp.h0);
  D3 xk[4] = {x[0], r, th, x[3]};
  M4 gks;
  gcov_ks(xk, mp, gks);
  D3 jf[4] = {dc(1), j1, j2, dc(1)};
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) g[i][j] = mul(mul(gks[i][j], jf[i]), jf[j]);
}

// ---- linalg (koral/math/linalg.zig:1-60) ----------------------------------
const int complement[4][3] = {{1, 2, 3}, {0, 2, 3}, {0, 1, 3}, {0, 1, 2}};
D3 minor3(M4 m, const int r[3], const int c[3]) {
  D3 t0 = mul(m[r[0]][c[0]], sub(mul(m[r[1]][c[1]], m[r[2]][c[2]]), mul(m[r[1]][c[2]], m[r[2]][c[1]])));
  D3 t1 = mul(m[r[0]][c[1]], sub(mul(m[r[1]][c[0]], m[r[2]][c[2]]), mul(m[r[1]][c[2]], m[r[2]][c[0]])));
  D3 t2 = mul(m[r[0]][c[2]], sub(mul(m[r[1]][c[0]], m[r[2]][c[1]]), mul(m[r[1]][c[1]], m[r[2]][c[0]])));
  return add(sub(t0, t1), t2);
}
D3 det4(M4 m) {
  D3 d = dc(0);
  double sign = 1.0;
  for (int j = 0; j < 4; ++j) {
    D3 minor = minor3(m, complement[0], complement[j]);
    d = add(d, scale(mul(m[0][j], minor), sign));
    sign = -sign;
  }
  return d;
}
void inv4(M4 m, D3 det, M4 out) {
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) {
      D3 minor = minor3(m, complement[j], complement[i]);
      double sign = ((i + j) % 2 == 0) ? 1.0 : -1.0;
      out[i][j] = dvd(scale(minor, sign), det);
    }
}

// ---- metric.zig:34-124 ----------------------------------------------------
enum Coords { C_MINK = 0, C_BL = 1, C_KS = 2, C_MKS2 = 3 };

struct CoordData {
  double gcov[4][4], gcon[4][4], gdet, dlgdet[3], kris[4][4][4], gttpert;
};

void gcov_dual(int coords, const MP &mp, const double x[4], M4 g) {
  D3 xd[4] = {dc(x[0]), dvar(x[1], 0), dvar(x[2], 1), dvar(x[3], 2)};
  if (coords == C_BL) gcov_bl(xd, mp, g);
  else if (coords == C_KS) gcov_ks(xd, mp, g);
  else if (coords == C_MKS2) gcov_mks2(xd, mp, g);
  else {
    for (int i = 0; i < 4; ++i)
      for (int j = 0; j < 4; ++j) g[i][j] = dc(0);
    g[0][0] = dc(-1); g[1][1] = dc(1); g[2][2] = dc(1); g[3][3] = dc(1);
  }
}

double gttpert_of(int coords, const MP &mp, const double x[4]) {
  if (coords == C_MINK) return 0.0;
  if (coords == C_BL || coords == C_KS) {
    double r = x[1], c = std::cos(x[2]);
    return 2.0 * r / (r * r + mp.a * mp.a * c * c);
  }
  double r = std::exp(x[1]) + mp.r0;
  double th = mks2_theta(dc(x[2]), mp.h0).v;
  double c = std::cos(th);
  return 2.0 * r / (r * r + mp.a * mp.a * c * c);
}

inline double dpart(M4 g, int a, int b, int m) { return m == 0 ? 0.0 : g[a][b].d[m - 1]; }

CoordData compute(int coords, const MP &mp, const double x[4]) {
  M4 g, ginv;
  gcov_dual(coords, mp, x, g);
  D3 det = det4(g);
  inv4(g, det, ginv);
  CoordData o;
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) {
      o.gcov[i][j] = g[i][j].v;
      o.gcon[i][j] = ginv[i][j].v;
    }
  o.gdet = std::sqrt(-det.v);
  for (int i = 0; i < 3; ++i) o.dlgdet[i] = 0.5 * det.d[i] / det.v;
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j)
      for (int k = j; k < 4; ++k) {
        double s = 0;
        for (int l = 0; l < 4; ++l) {
          double t = dpart(g, l, k, j) + dpart(g, l, j, k) - dpart(g, j, k, l);
          s += o.gcon[i][l] * t;
        }
        o.kris[i][j][k] = 0.5 * s;
        o.kris[i][k][j] = 0.5 * s;
      }
  o.gttpert = gttpert_of(coords, mp, x);
  return o;
}

double gdet_at(int coords, const MP &mp, const double x[4]) {
  M4 g;
  gcov_dual(coords, mp, x, g);
  return std::sqrt(-det4(g).v);
}

// ---- grid (koral/grid.zig:127-159) ----------------------------------------
struct Grid {
  kz_grid g;
  int ngx, ngy, ngz, sx, sy, sz;
  double dx, dy, dz;
  explicit Grid(const kz_grid &gr) : g(gr) {
    ngx = gr.nx > 1 ? gr.ng : 0;
    ngy = gr.ny > 1 ? gr.ng : 0;
    ngz = gr.nz > 1 ? gr.ng : 0;
    sx = gr.nx + 2 * ngx;
    sy = gr.ny + 2 * ngy;
    sz = gr.nz + 2 * ngz;
    dx = (gr.maxx - gr.minx) / (double)gr.nx;
    dy = (gr.maxy - gr.miny) / (double)gr.ny;
    dz = (gr.maxz - gr.minz) / (double)gr.nz;
  }
  double xl(long i) const { return g.minx + (double)i * dx; }
  double yl(long i) const { return g.miny + (double)i * dy; }
  double zl(long i) const { return g.minz + (double)i * dz; }
  double xc(long i) const { return 0.5 * (xl(i) + xl(i + 1)); }
  double yc(long i) const { return 0.5 * (yl(i) + yl(i + 1)); }
  double zc(long i) const { return 0.5 * (zl(i) + zl(i + 1)); }
  double csize(long i, int d) const {
    return d == 0 ? xl(i + 1) - xl(i) : (d == 1 ? yl(i + 1) - yl(i) : zl(i + 1) - zl(i));
  }
  long cidx(long ix, long iy, long iz) const {
    return (ix + ngx) + (iy + ngy) * (long)sx + (iz + ngz) * (long)sy * sx;
  }
  long ncell() const { return (long)sx * sy * sz; }
  long nface(int d) const {
    return d == 0 ? (long)(sx + 1) * sy * sz : (d == 1 ? (long)sx * (sy + 1) * sz : (long)sx * sy * (sz + 1));
  }
  long fidx(int d, long ix, long iy, long iz) const {
    long jx = ix + ngx, jy = iy + ngy, jz = iz + ngz;
    if (d == 0) return jx + jy * (sx + 1) + jz * (long)sy * (sx + 1);
    if (d == 1) return jx + jy * sx + jz * (long)(sy + 1) * sx;
    return jx + jy * sx + jz * (long)sy * sx;
  }
};

// precompute.zig:30-73 applyKrisCorrection
void apply_kris_correction(int coords, const MP &mp, const Grid &g, CoordData &d, long ix, long iy, long iz) {
  double xc[4] = {0, g.xc(ix), g.yc(iy), g.zc(iz)};
  double gdet_c = d.gdet;
  double dxs[3] = {g.dx, g.dy, g.dz};
  for (int kappa = 1; kappa < 4; ++kappa) {
    double ck = 0, sk = 1.0e-300;
    for (int mu = 0; mu < 4; ++mu) {
      ck += d.kris[mu][kappa][mu];
      sk += std::fabs(d.kris[mu][kappa][mu]);
    }
    double xh[4] = {xc[0], xc[1], xc[2], xc[3]}, xlo[4] = {xc[0], xc[1], xc[2], xc[3]};
    if (kappa == 1) { xh[1] = g.xl(ix + 1); xlo[1] = g.xl(ix); }
    if (kappa == 2) { xh[2] = g.yl(iy + 1); xlo[2] = g.yl(iy); }
    if (kappa == 3) { xh[3] = g.zl(iz + 1); xlo[3] = g.zl(iz); }
    double gh = gdet_at(coords, mp, xh), gl = gdet_at(coords, mp, xlo);
    double dk = (gh - gl) / (dxs[kappa - 1] * gdet_c);
    for (int mu = 0; mu < 4; ++mu) {
      double wk = std::fabs(d.kris[mu][kappa][mu]) / sk;
      d.kris[mu][kappa][mu] += (dk - ck) * wk;
      d.kris[mu][mu][kappa] = d.kris[mu][kappa][mu];
    }
  }
}

void store_blocks(double *g20, double *G20, const CoordData &d) {
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) {
      g20[i * 5 + j] = d.gcov[i][j];
      G20[i * 5 + j] = d.gcon[i][j];
    }
  for (int j = 0; j < 3; ++j) g20[j * 5 + 4] = d.dlgdet[j];
  g20[3 * 5 + 4] = d.gdet;
  for (int i = 0; i < 3; ++i) G20[i * 5 + 4] = 0;
  G20[3 * 5 + 4] = d.gttpert;
}

// ---- coco.zig -------------------------------------------------------------
void mks2_to_ks(const double x[4], const MP &mp, double o[4]) {
  o[0] = x[0];
  o[1] = std::exp(x[1]) + mp.r0;
  o[2] = mks2_theta(dc(x[2]), mp.h0).v;
  o[3] = x[3];
}
// time shift is not needed for any quantity on the path (x0 is unused).
void coco_mks2_to_bl(const double x[4], const MP &mp, double o[4]) {
  mks2_to_ks(x, mp, o);
  o[0] = 0.0;
}
void eye(double m[4][4]) {
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) m[i][j] = i == j ? 1.0 : 0.0;
}
void matmul4(double a[4][4], double b[4][4], double c[4][4]) {
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) {
      double s = 0;
      for (int k = 0; k < 4; ++k) s += a[i][k] * b[k][j];
      c[i][j] = s;
    }
}
// dxdx(x, mks2, bl) = dxdxKsToBl(mks2ToKs(x)) * dxdxMks2ToKs(x)
void dxdx_mks2_to_bl(const double x[4], const MP &mp, double out[4][4]) {
  double a[4][4], b[4][4], xk[4];
  mks2_to_ks(x, mp, xk);
  eye(a);
  double r = xk[1];
  double delta = r * r - 2.0 * r + mp.a * mp.a;
  a[0][1] = -2.0 * r / delta;
  a[3][1] = -mp.a / delta;
  eye(b);
  b[1][1] = std::exp(x[1]);
  b[2][2] = mks2_dtheta_dx2_metric(dc(x[2]), mp.h0).v;
  matmul4(a, b, out);
}
// dxdx(x, bl, mks2) = dxdxKsToMks2(blToKs(x)) * dxdxBlToKs(x)
void dxdx_bl_to_mks2(const double x[4], const MP &mp, double out[4][4]) {
  double a[4][4], b[4][4];
  double h0 = mp.h0;
  eye(a);
  a[1][1] = 1.0 / (x[1] - mp.r0);
  double v = (2.0 / PI) * (x[2] - 0.5 * PI) * std::tan(0.5 * PI * h0);
  a[2][2] = 2.0 * std::tan(0.5 * PI * h0) / (h0 * PI * PI * (1.0 + v * v));
  eye(b);
  double r = x[1];
  double delta = r * r - 2.0 * r + mp.a * mp.a;
  b[0][1] = 2.0 * r / delta;
  b[3][1] = mp.a / delta;
  matmul4(a, b, out);
}

// precompute.zig:77-100 geometryAt
kz_geom geometry_at(int coords, const MP &mp, const double x[4]) {
  CoordData d = compute(coords, mp, x);
  kz_geom o;
  memset(&o, 0, sizeof o);
  o.xx[0] = 0; o.xx[1] = x[1]; o.xx[2] = x[2]; o.xx[3] = x[3];
  for (int i = 0; i < 4; ++i) {
    for (int j = 0; j < 4; ++j) {
      o.gg[i][j] = d.gcov[i][j];
      o.GG[i][j] = d.gcon[i][j];
    }
    o.gg[i][4] = i < 3 ? d.dlgdet[i] : d.gdet;
    o.GG[i][4] = i == 3 ? d.gttpert : 0;
  }
  o.gdet = d.gdet;
  o.gttpert = d.gttpert;
  o.alpha = std::sqrt(-1.0 / o.GG[0][0]);
  return o;
}

kz_geom geometry_bl_at(const Grid &g, const MP &mp, long ix, long iy, long iz) {
  double xx[4] = {0.0, g.xc(ix), g.yc(iy), g.zc(iz)}, xb[4];
  coco_mks2_to_bl(xx, mp, xb);
  return geometry_at(C_BL, mp, xb);
}

// ---- relele helpers on kz_geom (init only) --------------------------------
void raise_g(const double a[4], const kz_geom &g, double o[4]) {
  double r[4] = {0, 0, 0, 0};
  for (int i = 0; i < 4; ++i)
    for (int k = 0; k < 4; ++k) r[i] += a[k] * g.GG[i][k];
  for (int i = 0; i < 4; ++i) o[i] = r[i];
}
void lower_g(const double a[4], const kz_geom &g, double o[4]) {
  double r[4] = {0, 0, 0, 0};
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) r[i] += a[j] * g.gg[i][j];
  for (int i = 0; i < 4; ++i) o[i] = r[i];
}
double dot4(const double a[4], const double b[4]) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2] + a[3] * b[3]; }
double alpha_gamma(const double u[4], const kz_geom &g) {
  double qsq = 0;
  for (int i = 1; i < 4; ++i)
    for (int j = 1; j < 4; ++j) qsq += u[i] * u[j] * g.gg[i][j];
  double ag2 = (-1.0 / g.GG[0][0]) * (1.0 + qsq);
  return ag2 < 0 ? 1.0 : std::sqrt(ag2);
}
void velr_to_vel4(const double u[4], const kz_geom &g, double o[4]) {
  double ag = alpha_gamma(u, g);
  double o0 = -ag * g.GG[0][0], o1 = u[1] - ag * g.GG[0][1], o2 = u[2] - ag * g.GG[0][2], o3 = u[3] - ag * g.GG[0][3];
  o[0] = o0; o[1] = o1; o[2] = o2; o[3] = o3;
}
// VEL4 -> VELR with u^t trusted (relele.zig convert, .trust_ut)
void vel4_to_velr_trust(const double u[4], const kz_geom &g, double o[4]) {
  double ut = u[0];
  double r1 = u[1] - ut * g.GG[0][1] / g.GG[0][0];
  double r2 = u[2] - ut * g.GG[0][2] / g.GG[0][0];
  double r3 = u[3] - ut * g.GG[0][3] / g.GG[0][0];
  o[0] = ut; o[1] = r1; o[2] = r2; o[3] = r3;
}
void normal_obs_velr(const kz_geom &g, double o[4]) {
  double alp = 1.0 / std::sqrt(-g.GG[0][0]);
  double w[4] = {-alp, 0, 0, 0}, ucon[4];
  raise_g(w, g, ucon);
  o[0] = ucon[0];
  for (int i = 1; i < 4; ++i) o[i] = ucon[i] - ucon[0] * g.GG[0][i] / g.GG[0][0];
}
void tvec4(double J[4][4], const double u[4], double o[4]) {
  double r[4] = {0, 0, 0, 0};
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) r[i] += J[i][j] * u[j];
  for (int i = 0; i < 4; ++i) o[i] = r[i];
}

// invert_rad.zig:189-261 u2pRad restated for the init path (pradFf2Lab)
double m1_gammarel2(const kz_geom &gm, const double av[4]) {
  double gn11 = gm.GG[0][0], gn12 = gm.GG[0][1], gn13 = gm.GG[0][2], gn14 = gm.GG[0][3];
  double gn22 = gm.GG[1][1], gn23 = gm.GG[1][2], gn24 = gm.GG[1][3], gn33 = gm.GG[2][2], gn34 = gm.GG[2][3],
         gn44 = gm.GG[3][3];
  double rdtt = av[0], rdtx = av[1], rdty = av[2], rdtz = av[3];
  double mix = gn12 * rdtx + gn13 * rdty + gn14 * rdtz;
  double sqv = std::sqrt(4.0 * (gn11 * gn11) * (rdtt * rdtt) + mix * mix +
                         gn11 * (8.0 * gn12 * rdtt * rdtx + 3.0 * gn22 * (rdtx * rdtx) + 8.0 * gn13 * rdtt * rdty +
                                 6.0 * gn23 * rdtx * rdty + 3.0 * gn33 * (rdty * rdty) + 8.0 * gn14 * rdtt * rdtz +
                                 6.0 * gn24 * rdtx * rdtz + 6.0 * gn34 * rdty * rdtz + 3.0 * gn44 * (rdtz * rdtz)));
  double den = gn11 * (rdtt * rdtt) + 2.0 * gn12 * rdtt * rdtx + gn22 * (rdtx * rdtx) + 2.0 * gn13 * rdtt * rdty +
               2.0 * gn23 * rdtx * rdty + gn33 * (rdty * rdty) + 2.0 * (gn14 * rdtt + gn24 * rdtx + gn34 * rdty) * rdtz +
               gn44 * (rdtz * rdtz);
  double g2a = (-0.25 * (2.0 * (gn11 * gn11) * (rdtt * rdtt) + mix * (mix + sqv) +
                         gn11 * (4.0 * gn12 * rdtt * rdtx + gn22 * (rdtx * rdtx) + 2.0 * gn23 * rdtx * rdty +
                                 gn33 * (rdty * rdty) + 2.0 * gn24 * rdtx * rdtz + 2.0 * gn34 * rdty * rdtz +
                                 gn44 * (rdtz * rdtz) + rdtt * (4.0 * gn13 * rdty + 4.0 * gn14 * rdtz + sqv)))) /
               den;
  const double gsl = 1.0 - 1e-10;
  double g2;
  if (g2a < gsl || !std::isfinite(g2a)) {
    g2 = (0.25 * (-2.0 * (gn11 * gn11) * (rdtt * rdtt) -
                  1.0 * gn11 * (4.0 * gn12 * rdtt * rdtx + gn22 * (rdtx * rdtx) +
                                rdty * (4.0 * gn13 * rdtt + 2.0 * gn23 * rdtx + gn33 * rdty) +
                                2.0 * (2.0 * gn14 * rdtt + gn24 * rdtx + gn34 * rdty) * rdtz + gn44 * (rdtz * rdtz)) +
                  gn11 * rdtt * sqv + mix * (-1.0 * gn12 * rdtx - 1.0 * gn13 * rdty - 1.0 * gn14 * rdtz + sqv))) /
         den;
  } else {
    g2 = g2a;
  }
  double gr2 = g2 * gm.alpha * gm.alpha;
  if (gr2 > gsl && gr2 < 1.0) gr2 = 1.0;
  return gr2;
}
// invert_rad.zig:140-185 m1Cold
void m1_cold(const kz_geom &gm, const double in[4], double gr2, double av[4], double *erf_o, double ur[4]) {
  double gn11 = gm.GG[0][0], gn12 = gm.GG[0][1], gn13 = gm.GG[0][2], gn14 = gm.GG[0][3];
  double gn22 = gm.GG[1][1], gn23 = gm.GG[1][2], gn24 = gm.GG[1][3], gn33 = gm.GG[2][2], gn34 = gm.GG[2][3],
         gn44 = gm.GG[3][3];
  double alpha = gm.alpha, rdtx = in[1], rdty = in[2], rdtz = in[3];
  double utsq = gr2 / (alpha * alpha);
  for (int i = 0; i < 4; ++i) av[i] = in[i];
  double mix = gn13 * rdty + gn14 * rdtz;
  double disc = std::sqrt(((gn12 * gn12) * (rdtx * rdtx) + 2.0 * gn12 * rdtx * mix + mix * mix -
                           1.0 * gn11 * (gn22 * (rdtx * rdtx) + 2.0 * gn23 * rdtx * rdty + gn33 * (rdty * rdty) +
                                         2.0 * gn24 * rdtx * rdtz + 2.0 * gn34 * rdty * rdtz + gn44 * (rdtz * rdtz))) *
                          utsq * (gn11 + utsq) * ((gn11 + 4.0 * utsq) * (gn11 + 4.0 * utsq)));
  av[0] = (0.25 * (-4.0 * (gn12 * rdtx + gn13 * rdty + gn14 * rdtz) * utsq * (gn11 + utsq) + disc)) /
          (gn11 * utsq * (gn11 + utsq));
  double erf = (0.75 * disc) / (utsq * (gn11 + utsq) * (gn11 + 4.0 * utsq));
  double avc[4];
  raise_g(av, gm, avc);
  double gr = std::sqrt(gr2);
  for (int i = 0; i < 4; ++i) ur[i] = 0;
  if (erf > 0.0)
    for (int jj = 1; jj < 4; ++jj)
      ur[jj] = alpha * (avc[jj] + 1.0 / 3.0 * erf * gm.GG[0][jj] * (4.0 * gr2 - 1.0)) / (4.0 / 3.0 * erf * gr);
  *erf_o = erf;
}

// invert_rad.zig:189-261 u2pRad restated for the init path (pradFf2Lab)
void u2p_rad_init(const double uu[4], double *pp, const kz_geom &gm, const kz_rad_params &rp) {
  double av[4] = {uu[0] / gm.gdet, uu[1] / gm.gdet, uu[2] / gm.gdet, uu[3] / gm.gdet}, avc[4];
  raise_g(av, gm, avc);
  double gr2 = m1_gammarel2(gm, av);
  double erf = 3.0 * avc[0] * gm.alpha * gm.alpha / (4.0 * gr2 - 1.0);
  const double gsl = 1.0 - 1e-10;
  double gmax = rp.gammamaxrad;
  bool ok = gr2 >= 1.0 && erf > rp.eradfloor && gr2 <= gmax * gmax / gsl / gsl;
  double uc[4];
  if (ok) {
    double gr = std::sqrt(gr2);
    for (int ii = 0; ii < 4; ++ii)
      uc[ii] = gm.alpha * (avc[ii] + 1.0 / 3.0 * erf * gm.GG[0][ii] * (4.0 * gr2 - 1.0)) / (4.0 / 3.0 * erf * gr);
  } else {
    double one_eps = 1.0 + 10.0 * 2.220446049250313080847e-16;
    double sa[4], se, su[4], fa[4], fe, fu[4];
    m1_cold(gm, av, one_eps * one_eps, sa, &se, su);
    m1_cold(gm, av, gmax * gmax, fa, &fe, fu);
    if (std::fabs(sa[0] - av[0]) > std::fabs(fa[0] - av[0])) {
      for (int i = 0; i < 4; ++i) uc[i] = fu[i];
      gr2 = gmax * gmax;
      erf = fe;
    } else {
      for (int i = 0; i < 4; ++i) uc[i] = su[i];
      gr2 = one_eps * one_eps;
      erf = se;
    }
    if (!std::isfinite(erf) || !std::isfinite(gr2) || !std::isfinite(uc[1]) || !std::isfinite(uc[2]) ||
        !std::isfinite(uc[3]))
      return;
  }
  pp[KZ_EE] = erf;
  pp[KZ_FX] = uc[1];
  pp[KZ_FY] = uc[2];
  pp[KZ_FZ] = uc[3];
}

// ---- PUFFY limotorus (puffy.zig:330-609) ----------------------------------
const double LT_XI = 0.995, LT_R1 = 20.0, LT_R2 = 350.0, LT_GAMMA = 4.0 / 3.0, LT_RIN = 35.0;
const double XACC = 5.0 * 2.220446049250313080847e-16;
const double PI_2 = PI / 2.0;

struct Gd {
  double gdtt, gdtp, gdpp;
};
Gd compute_gd(double r, double th, double a) {
  double ac = a * std::cos(th);
  double sigma = ac * ac + r * r;
  double s2 = std::sin(th) * std::sin(th);
  double tmp = 2.0 * r * s2 / sigma;
  return Gd{-1.0 + 2.0 * r / sigma, -a * tmp, (r * r + a * a * (1.0 + tmp)) * s2};
}
double lK(double r, double a) {
  double cf = 1.0 - 2.0 * a / std::pow(r, 1.5) + (a / r) * (a / r);
  double cg = 1.0 - 2.0 / r + a / std::pow(r, 1.5);
  return std::sqrt(r) * cf / cg;
}
double l3d(double lam, double a, double lb1, double lb2, double xi) {
  double arg = lam <= lb1 ? lb1 : (lam >= lb2 ? lb2 : lam);
  return xi * lK(arg, a);
}
template <class F>
double rtbis(const F &f, double x1, double x2, double xacc) {
  double fv = f(x1);
  double fmid = f(x2);
  (void)fmid;
  double dx, rtb;
  if (fv < 0.0) {
    dx = x2 - x1;
    rtb = x1;
  } else {
    dx = x1 - x2;
    rtb = x2;
  }
  for (int j = 1; j <= 100; ++j) {
    dx *= 0.5;
    double xmid = rtb + dx;
    fmid = f(xmid);
    if (fmid <= 0.0) rtb = xmid;
    if (std::fabs(dx) < xacc || fmid == 0.0) return rtb;
  }
  return 0.0;
}
double lamBL(double R, Gd gd, double a, double lb1, double lb2, double xi) {
  auto f = [&](double lam) {
    double l = l3d(lam, a, lb1, lb2, xi);
    return lam * lam + l * (l * gd.gdtp + gd.gdpp) / (l * gd.gdtt + gd.gdtp);
  };
  return rtbis(f, R, 10.0 * R, XACC);
}
double omega3d(double l, Gd gd) { return -(gd.gdtt * l + gd.gdtp) * (1.0 / (gd.gdpp + gd.gdtp * l)); }
double compute_agrav(double om, Gd gd) {
  return std::sqrt(std::fabs(1.0 / (gd.gdtt + 2.0 * om * gd.gdtp + om * om * gd.gdpp)));
}
double find_rml(double lam, double a, double lb1, double lb2, double xi) {
  double lamsq = lam * lam;
  auto f = [&](double x) {
    Gd gd = compute_gd(x, PI_2, a);
    double lx = lamBL(x, gd, a, lb1, lb2, xi);
    return lamsq - lx * lx;
  };
  return rtbis(f, 6.0, 10.0 * lamsq, XACC);
}
struct Lnf {
  double a, lb1, lb2, xi;
  double operator()(double r) const {
    double r2 = r * r, r3 = r2 * r, a2 = a * a;
    Gd gd = compute_gd(r, PI_2, a);
    double lam = lamBL(r, gd, a, lb1, lb2, xi);
    double lam2 = lam * lam;
    double l = l3d(lam, a, lb1, lb2, xi);
    double term1 = (r - 2.0) * l, term2 = 2.0 * a, term3 = r3 + a2 * (r + 2.0), term4 = term2 * l;
    double omn = term1 + term2, omd = term3 - term4;
    double om = omn / omd;
    double dl_dr = 0.0;
    if (lam <= lb1 || lam >= lb2) {
      dl_dr = 0.0;
    } else {
      double oneplusx = 1.0 - 2.0 / lam + a * std::pow(lam, -1.5);
      double dx_dlam = 2.0 * std::pow(lam, -2.0) - 1.5 * a * std::pow(lam, -2.5);
      double lamroot = std::sqrt(lam);
      double dlk_dlam = (oneplusx * 0.5 * (1.0 / lamroot) + (a - lamroot) * dx_dlam) / (oneplusx * oneplusx);
      double dd = term3 - 2.0 * term4 - lam2 * (r - 2.0);
      double ee = l * (3.0 * r2 + a2 - lam2);
      dl_dr = ee / (2.0 * lam * omn / (xi * dlk_dlam) - dd);
    }
    double dom_dr = (omd * (l + (r - 2.0) * dl_dr) - omn * (3.0 * r2 + a2 + 2.0 * a * dl_dr)) / (omd * omd);
    return -l / (1.0 - om * l) * dom_dr;
  }
};

// koral/math/quad.zig:1-130 adaptive Gauss-Kronrod 21
const double XGK[11] = {0.995657163025808080735527280689003, 0.973906528517171720077964012084452,
                        0.930157491355708226001207180059508, 0.865063366688984510732096688423493,
                        0.780817726586416897063717578345042, 0.679409568299024406234327365114874,
                        0.562757134668604683339000099272694, 0.433395394129247190799265943165784,
                        0.294392862701460198131126603103866, 0.148874338981631210884826001129720, 0.0};
const double WGK[11] = {0.011694638867371874278064396062192, 0.032558162307964727478818972459390,
                        0.054755896574351996031381300244580, 0.075039674810919952767043140916190,
                        0.093125454583697605535065465083366, 0.109387158802297641899210590325805,
                        0.123491976262065851077958109831074, 0.134709217311473325928054001771707,
                        0.142775938577060080797094273138717, 0.147739104901338491374841515972068,
                        0.149445554002916905664936468389821};
const double WG[5] = {0.066671344308688137593568809893332, 0.149451349150580593145776339657697,
                      0.219086362515982043995534934228163, 0.269266719309996355091226921569469,
                      0.295524224714752870173892994651338};
template <class F>
void gk21(const F &f, double a, double b, double *val, double *err) {
  double c = 0.5 * (a + b), h = 0.5 * (b - a);
  double fc = f(c);
  double resg = 0.0, resk = WGK[10] * fc;
  for (int j = 0; j < 5; ++j) {
    int jtw = 2 * j + 1;
    double x = h * XGK[jtw];
    double fsum = f(c - x) + f(c + x);
    resg += WG[j] * fsum;
    resk += WGK[jtw] * fsum;
  }
  for (int j = 0; j < 5; ++j) {
    int jtw = 2 * j;
    double x = h * XGK[jtw];
    resk += WGK[jtw] * (f(c - x) + f(c + x));
  }
  *val = resk * h;
  *err = std::fabs((resk - resg) * h);
}
template <class F>
bool integrate(const F &f, double a, double b, double epsabs, double epsrel, double *value) {
  if (a == b) {
    *value = 0.0;
    return true;
  }
  struct Seg {
    double a, b, v, e;
  };
  std::vector<Seg> segs(512);
  int nseg = 1;
  double v0, e0;
  gk21(f, a, b, &v0, &e0);
  segs[0] = Seg{a, b, v0, e0};
  for (;;) {
    double total = 0.0, toterr = 0.0;
    int worst = 0;
    for (int i = 0; i < nseg; ++i) {
      total += segs[i].v;
      toterr += segs[i].e;
      if (segs[i].e > segs[worst].e) worst = i;
    }
    if (toterr <= std::fmax(epsabs, epsrel * std::fabs(total))) {
      *value = total;
      return true;
    }
    if (nseg == 512) {
      *value = total;
      return false;
    }
    Seg s = segs[worst];
    double mid = 0.5 * (s.a + s.b), lv, le, rv, re;
    gk21(f, s.a, mid, &lv, &le);
    gk21(f, mid, s.b, &rv, &re);
    segs[worst] = Seg{s.a, mid, lv, le};
    segs[nseg] = Seg{mid, s.b, rv, re};
    nseg += 1;
  }
}

struct TorusConsts {
  double lb1, lb2, lamin, lin, omin, agravin;
};
TorusConsts torus_consts(double a) {
  TorusConsts t;
  Gd gd = compute_gd(LT_R1, PI_2, a);
  t.lb1 = lamBL(LT_R1, gd, a, 0.0, 200000.0, LT_XI);
  gd = compute_gd(LT_R2, PI_2, a);
  t.lb2 = lamBL(LT_R2, gd, a, t.lb1, 200000.0, LT_XI);
  gd = compute_gd(LT_RIN, PI_2, a);
  t.lamin = lamBL(LT_RIN, gd, a, t.lb1, t.lb2, LT_XI);
  t.lin = l3d(t.lamin, a, t.lb1, t.lb2, LT_XI);
  t.omin = omega3d(t.lin, gd);
  t.agravin = compute_agrav(t.omin, gd);
  return t;
}

struct DsVels {
  double rho, uu, ell;
};
DsVels init_dsandvels(double r, double th, double a, const TorusConsts &tc, double kappa) {
  double R = r * std::sin(th);
  if (R < LT_RIN) return DsVels{-1.0, 0.0, 0.0};
  Gd gd = compute_gd(r, th, a);
  double lam = lamBL(R, gd, a, tc.lb1, tc.lb2, LT_XI);
  double l = l3d(lam, a, tc.lb1, tc.lb2, LT_XI);
  double om = omega3d(l, gd);
  double agrav = compute_agrav(om, gd);
  double rml = find_rml(lam, a, tc.lb1, tc.lb2, LT_XI);
  Lnf integrand{a, tc.lb1, tc.lb2, LT_XI};
  double q;
  bool conv = integrate(integrand, LT_RIN, rml, 0.0, 1.0e-12, &q);
  double lnf3d = conv ? q : -INFINITY;
  double f3d = std::exp(-lnf3d);
  double hh = 1.0 * agrav / (f3d * tc.agravin);
  double eps = (-1.0 + hh) * (1.0 / LT_GAMMA);
  double rho = eps < 0.0 ? -1.0 : std::pow((-1.0 + LT_GAMMA) * eps * (1.0 / kappa), 1.0 / (-1.0 + LT_GAMMA));
  return DsVels{rho, kappa * std::pow(rho, LT_GAMMA) / (LT_GAMMA - 1.0), l};
}

// prepinit.c:92-93 tFromPtot (puffy.zig:665-678)
double t_from_ptot(double P, double aaa, double bbb) {
  const double third = 0.3333333333333333, twothird = 0.6666666666666666;
  double naw1 = std::cbrt(9.0 * aaa * (bbb * bbb) -
                          std::sqrt(3.0) * std::sqrt(27.0 * (aaa * aaa) * std::pow(bbb, 4.0) +
                                                     256.0 * std::pow(aaa, 3.0) * std::pow(P, 3.0)));
  double c23 = std::pow(twothird, third), c2 = std::pow(2.0, third), c3 = std::pow(3.0, twothird);
  return -std::sqrt((-4.0 * c23 * P) / naw1 + naw1 / (c2 * c3 * aaa)) / 2.0 +
         std::sqrt((4.0 * c23 * P) / naw1 - naw1 / (c2 * c3 * aaa) +
                   (2.0 * bbb) / (aaa * std::sqrt((-4.0 * c23 * P) / naw1 + naw1 / (c2 * c3 * aaa)))) /
             2.0;
}

// frames.zig lorentzFromPair + boost22Ff2Lab (pradFf2Lab)
void lorentz_pair(const double ucon[4], const double ucov[4], const double wcon[4], const double wcov[4],
                  double l[4][4]) {
  double om[4][4];
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) om[i][j] = ucon[i] * wcov[j] - wcon[i] * ucov[j];
  double gam = -dot4(wcon, ucov);
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) {
      double omsum = 0;
      for (int k = 0; k < 4; ++k) omsum += om[i][k] * om[k][j];
      l[i][j] = (i == j ? 1.0 : 0.0) + 1.0 / (1.0 + gam) * omsum + om[i][j];
    }
}

// radiation.zig:146-171 pradFf2Lab
void prad_ff2lab(double *pp, const kz_geom &g, const kz_rad_params &rp) {
  double urf[4], v[4] = {0, pp[KZ_FX], pp[KZ_FY], pp[KZ_FZ]};
  velr_to_vel4(v, g, urf);
  double erf = pp[KZ_EE];
  double rff[4][4];
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) rff[i][j] = 4.0 / 3.0 * erf * urf[i] * urf[j] + 1.0 / 3.0 * erf * g.GG[i][j];
  // lorentzFf2Lab: w = gas (VELR -> con/cov), u = normal observer
  double vg[4] = {0, pp[KZ_VX], pp[KZ_VY], pp[KZ_VZ]}, wcon[4], wcov[4];
  velr_to_vel4(vg, g, wcon);
  lower_g(wcon, g, wcov);
  double alpha = std::sqrt(-1.0 / g.GG[0][0]);
  double ncov[4] = {-alpha, 0, 0, 0}, ncon[4];
  raise_g(ncov, g, ncon);
  double L[4][4];
  lorentz_pair(ncon, ncov, wcon, wcov, L);
  double rlab[4][4];
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) {
      double s = 0;
      for (int k = 0; k < 4; ++k)
        for (int l = 0; l < 4; ++l) s += L[i][k] * L[j][l] * rff[k][l];
      rlab[i][j] = s;
    }
  double rud[4][4];
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) {
      double s = 0;
      for (int k = 0; k < 4; ++k) s += rlab[i][k] * g.gg[k][j];
      rud[i][j] = s;
    }
  double uu[4] = {g.gdet * rud[0][0], g.gdet * rud[0][1], g.gdet * rud[0][2], g.gdet * rud[0][3]};
  u2p_rad_init(uu, pp, g, rp);
}

// frames.zig transPmhdCocoJ + transPradCoco (BL -> MKS2 at init)
void trans_pall_bl_to_mks2(double *pp, const kz_geom &g1, const kz_geom &g2, double J[4][4]) {
  double v[4] = {0, pp[KZ_VX], pp[KZ_VY], pp[KZ_VZ]}, u1c[4], u1v[4];
  velr_to_vel4(v, g1, u1c);
  lower_g(u1c, g1, u1v);
  double bcon[4];
  bcon[0] = pp[KZ_B1] * u1v[1] + pp[KZ_B2] * u1v[2] + pp[KZ_B3] * u1v[3];
  double inv_ut = 1.0 / u1c[0];
  for (int j = 1; j < 4; ++j) bcon[j] = (pp[KZ_B1 + j - 1] + bcon[0] * u1c[j]) * inv_ut;
  tvec4(J, bcon, bcon);
  double uc[4];
  tvec4(J, u1c, uc);
  vel4_to_velr_trust(uc, g2, uc);
  pp[KZ_VX] = uc[1]; pp[KZ_VY] = uc[2]; pp[KZ_VZ] = uc[3];
  double v2[4] = {0, pp[KZ_VX], pp[KZ_VY], pp[KZ_VZ]}, u2c[4];
  velr_to_vel4(v2, g2, u2c);
  for (int j = 1; j < 4; ++j) pp[KZ_B1 + j - 1] = bcon[j] * u2c[0] - bcon[0] * u2c[j];
  // radiation
  double vr[4] = {0, pp[KZ_FX], pp[KZ_FY], pp[KZ_FZ]}, ur[4];
  velr_to_vel4(vr, g1, ur);
  tvec4(J, ur, ur);
  vel4_to_velr_trust(ur, g2, ur);
  pp[KZ_FX] = ur[1]; pp[KZ_FY] = ur[2]; pp[KZ_FZ] = ur[3];
}

}  // namespace

// =====================================================================
// exported C-ABI (declared in include/kz_host.h)
// =====================================================================
extern "C" {

long kz_host_cell_count(const kz_grid *gr) { return Grid(*gr).ncell(); }
long kz_host_face_count(const kz_grid *gr, int dim) { return Grid(*gr).nface(dim); }

// MetricCache.init (precompute.zig:175-213, fillCenters/fillFaces):
// g, gcon: ncell*20; kris: ncell*64; gb[d], gconb[d]: nface(d)*20.
int kz_host_metric_tables(const kz_grid *gr, double a, double mksr0, double mksh0, double *g, double *gcon,
                        double *kris, double *gb0, double *gconb0, double *gb1, double *gconb1, double *gb2,
                        double *gconb2) {
  Grid G(*gr);
  MP mp{a, mksr0, mksh0};
  // every point is independent: rows of cells (and of faces below) run in
  // parallel, each point evaluated exactly as a serial loop would
  const long ncy = gr->ny + 2 * G.ngy, ncz = gr->nz + 2 * G.ngz;
#pragma omp parallel for schedule(dynamic, 1)
  for (long row = 0; row < ncy * ncz; ++row) {
    const long iz = row / ncy - G.ngz, iy = row % ncy - G.ngy;
    for (long ix = -G.ngx; ix < gr->nx + G.ngx; ++ix) {
      double x[4] = {0, G.xc(ix), G.yc(iy), G.zc(iz)};
      CoordData d = compute(C_MKS2, mp, x);
      apply_kris_correction(C_MKS2, mp, G, d, ix, iy, iz);
      long ci = G.cidx(ix, iy, iz);
      store_blocks(g + ci * 20, gcon + ci * 20, d);
      for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j)
          for (int k = 0; k < 4; ++k) kris[ci * 64 + i * 16 + j * 4 + k] = d.kris[i][j][k];
    }
  }
  double *gb[3] = {gb0, gb1, gb2}, *gcb[3] = {gconb0, gconb1, gconb2};
  long nxh = gr->nx + G.ngx, nyh = gr->ny + G.ngy, nzh = gr->nz + G.ngz;
  const long nfy = nyh + G.ngy + 1, nfz = nzh + G.ngz + 1;
#pragma omp parallel for schedule(dynamic, 1)
  for (long row = 0; row < nfy * nfz; ++row) {
    const long iz = row / nfy - G.ngz, iy = row % nfy - G.ngy;
    for (long ix = -G.ngx; ix <= nxh; ++ix) {
      if (iy < nyh && iz < nzh) {
        double x[4] = {0, G.xl(ix), G.yc(iy), G.zc(iz)};
        CoordData d = compute(C_MKS2, mp, x);
        long fi = G.fidx(0, ix, iy, iz);
        store_blocks(gb[0] + fi * 20, gcb[0] + fi * 20, d);
      }
      if (ix < nxh && iz < nzh) {
        double x[4] = {0, G.xc(ix), G.yl(iy), G.zc(iz)};
        CoordData d = compute(C_MKS2, mp, x);
        long fi = G.fidx(1, ix, iy, iz);
        store_blocks(gb[1] + fi * 20, gcb[1] + fi * 20, d);
      }
      if (ix < nxh && iy < nyh) {
        double x[4] = {0, G.xc(ix), G.yc(iy), G.zl(iz)};
        CoordData d = compute(C_MKS2, mp, x);
        long fi = G.fidx(2, ix, iy, iz);
        store_blocks(gb[2] + fi * 20, gcb[2] + fi * 20, d);
      }
    }
  }
  return 0;
}

// metric point evaluation for the metric goldens: out101 = gg16 GG16 gdet
// dlgdet3 Kr64 gttpert (coords: 1 BL, 2 KS, 3 MKS2, 0 Minkowski).
int kz_host_metric_point(int coords, double a, double mksr0, double mksh0, const double *x4, double *out101) {
  MP mp{a, mksr0, mksh0};
  CoordData d = compute(coords, mp, x4);
  int k = 0;
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) out101[k++] = d.gcov[i][j];
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) out101[k++] = d.gcon[i][j];
  out101[k++] = d.gdet;
  for (int i = 0; i < 3; ++i) out101[k++] = d.dlgdet[i];
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j)
      for (int l = 0; l < 4; ++l) out101[k++] = d.kris[i][j][l];
  out101[k++] = d.gttpert;
  return 0;
}

// kz_geom records (MetricCache.fillGeometry, precompute.zig:499-507) at n
// arbitrary points x4[4 n]: the 20-double blocks of storeBlocks, gdet,
// alpha = sqrt(-1/G^00), gttpert and the position.
int kz_host_geom_batch(int coords, double a, double mksr0, double mksh0, int64_t n, const double *x4,
                       kz_geom *out) {
  MP mp{a, mksr0, mksh0};
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) {
    CoordData d = compute(coords, mp, x4 + 4 * i);
    kz_geom &o = out[i];
    store_blocks(&o.gg[0][0], &o.GG[0][0], d);
    o.gdet = d.gdet;
    o.alpha = std::sqrt(-1.0 / d.gcon[0][0]);
    o.gttpert = d.gttpert;
    for (int k = 0; k < 4; ++k) o.xx[k] = x4[4 * i + k];
  }
  return 0;
}

// Jacobians for the metric goldens: dir 0 = MKS2->BL, 1 = BL->MKS2.
int kz_host_dxdx(int dir, double a, double mksr0, double mksh0, const double *x4, double *out16) {
  MP mp{a, mksr0, mksh0};
  double J[4][4];
  if (dir == 0) dxdx_mks2_to_bl(x4, mp, J);
  else dxdx_bl_to_mks2(x4, mp, J);
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) out16[i * 4 + j] = J[i][j];
  return 0;
}

// Outer radial (xhi) boundary tables for Bc.calc (puffy.zig:1089-1166), one
// 66-double record per ghost (iz, iy, i = ix - nx):
//   gBL[16] GBL[16] r_ghost r_bound J(MKS2->BL at the ghost)[16] J(BL->MKS2 at its BL point)[16]
int kz_host_bc_xhi_tables(const kz_grid *gr, double a, double mksr0, double mksh0, double *out) {
  Grid G(*gr);
  MP mp{a, mksr0, mksh0};
  long nx = gr->nx;
  for (long iz = 0; iz < gr->nz; ++iz)
    for (long iy = 0; iy < gr->ny; ++iy) {
      kz_geom bd = geometry_bl_at(G, mp, nx - 1, iy, iz);
      for (long i = 0; i < gr->ng; ++i) {
        long ix = nx + i;
        double *rec = out + (((iz * gr->ny) + iy) * gr->ng + i) * 66;
        kz_geom gbl = geometry_bl_at(G, mp, ix, iy, iz);
        for (int r = 0; r < 4; ++r)
          for (int c = 0; c < 4; ++c) {
            rec[r * 4 + c] = gbl.gg[r][c];
            rec[16 + r * 4 + c] = gbl.GG[r][c];
          }
        rec[32] = gbl.xx[1];
        rec[33] = bd.xx[1];
        double xx[4] = {0.0, G.xc(ix), G.yc(iy), G.zc(iz)};
        double J12[4][4], J21[4][4];
        dxdx_mks2_to_bl(xx, mp, J12);
        dxdx_bl_to_mks2(gbl.xx, mp, J21);
        for (int r = 0; r < 4; ++r)
          for (int c = 0; c < 4; ++c) {
            rec[34 + r * 4 + c] = J12[r][c];
            rec[50 + r * 4 + c] = J21[r][c];
          }
      }
    }
  return 0;
}

// PUFFY prepinit + init over the domain (puffy.zig:715-866): writes the
// padded AoS primitives p (ghosts untouched) with A_phi in the B3 slot and
// ENTR = sFromU.  Physics constants come from op (thermo.Consts) and rp.
int kz_host_puffy_prepinit(const kz_grid *gr, double a, double mksr0, double mksh0, double gam,
                           const kz_opac_params *op, const kz_rad_params *rp, double mass, double rhoatmmin,
                           double atm_tgas, double atm_trad_init, double lt_kappa, const double *g20,
                           const double *gcon20, double *p) {
  Grid G(*gr);
  MP mp{a, mksr0, mksh0};
  TorusConsts tc = torus_consts(a);
  double uintatmmin = (op->kb_over_mugas_mp * rhoatmmin * atm_tgas) / (gam - 1.0);
  double t3 = atm_trad_init;
  double eradatmmin = op->four_sigmarad * t3 * t3 * t3 * t3 / 10.0 * 6.62 / mass;
  // cells are independent (each runs its own torus quadrature): rows in
  // parallel.  The torus, the atmosphere and the metric do not depend on phi,
  // so in 3D the phi = first slice is computed and copied to the others
  // (identical bits: every input of a cell but its phi coordinate, which no
  // expression here reads, is the same in every slice).
  const long nrows = gr->ny;
#pragma omp parallel for schedule(dynamic, 1)
  for (long row = 0; row < nrows; ++row)
    for (long ix = 0; ix < gr->nx; ++ix) {
      const long iz = 0, iy = row;
      {
        long ci = G.cidx(ix, iy, iz);
        kz_geom geom;
        memset(&geom, 0, sizeof geom);
        for (int i = 0; i < 4; ++i)
          for (int j = 0; j < 5; ++j) {
            geom.gg[i][j] = g20[ci * 20 + i * 5 + j];
            geom.GG[i][j] = gcon20[ci * 20 + i * 5 + j];
          }
        geom.gdet = geom.gg[3][4];
        geom.gttpert = geom.GG[3][4];
        geom.alpha = std::sqrt(-1.0 / geom.GG[0][0]);
        geom.xx[1] = G.xc(ix); geom.xx[2] = G.yc(iy); geom.xx[3] = G.zc(iz);
        kz_geom gbl = geometry_bl_at(G, mp, ix, iy, iz);
        double r = gbl.xx[1], th = gbl.xx[2];
        DsVels dv = init_dsandvels(r, th, a, tc, lt_kappa);
        double pp[KZ_NV] = {0}, ppback[KZ_NV] = {0};
        // atmospheres (relele.c:518/:718 atmtype 0) in MKS2 at the cell
        auto atmos = [&](double *q) {
          double uc[4];
          normal_obs_velr(geom, uc);
          q[KZ_VX] = uc[1]; q[KZ_VY] = uc[2]; q[KZ_VZ] = uc[3];
          double xb[4];
          coco_mks2_to_bl(geom.xx, mp, xb);
          double rr = xb[1];
          q[KZ_RHO] = rhoatmmin * std::pow(rr / 2.0, -1.5);
          q[KZ_UU] = uintatmmin * std::pow(rr / 2.0, -2.5);
          q[KZ_B1] = 0.0; q[KZ_B2] = 0.0; q[KZ_B3] = 0.0;
          q[KZ_EE] = eradatmmin;
          q[KZ_FX] = uc[1]; q[KZ_FY] = uc[2]; q[KZ_FZ] = uc[3];
        };
        if (dv.rho < 0.0) {
          atmos(pp);
          pp[KZ_ENTR] = -1.0;
        } else {
          pp[KZ_ENTR] = 1.0;
          atmos(ppback);
          double uint = lt_kappa * std::pow(dv.rho, LT_GAMMA) / (LT_GAMMA - 1.0);
          double ell = -dv.ell;
          const auto &GB = gbl.GG;
          double ulph = std::sqrt(-1.0 / (GB[0][0] / ell / ell + 2.0 / ell * GB[0][3] + GB[3][3]));
          double ult = ulph / ell;
          double ucov[4] = {ult, 0.0, 0.0, ulph}, ucon[4];
          raise_g(ucov, gbl, ucon);
          vel4_to_velr_trust(ucon, gbl, ucon);
          pp[KZ_RHO] = std::fmax(dv.rho, ppback[KZ_RHO]);
          pp[KZ_UU] = std::fmax(uint, ppback[KZ_UU]);
          pp[KZ_VX] = ucon[1]; pp[KZ_VY] = ucon[2]; pp[KZ_VZ] = ucon[3];
          double P = (gam - 1.0) * uint;
          double aaa = op->four_sigmarad / 3.0;
          double bbb = op->kb_over_mugas_mp * dv.rho;
          double t4 = t_from_ptot(P, aaa, bbb);
          double E = op->four_sigmarad * t4 * t4 * t4 * t4;
          uint = (op->kb_over_mugas_mp * dv.rho * t4) / (gam - 1.0);
          pp[KZ_UU] = std::fmax(uint, ppback[KZ_UU]);
          pp[KZ_EE] = std::fmax(E, ppback[KZ_EE]);
          pp[KZ_FX] = 0.0; pp[KZ_FY] = 0.0; pp[KZ_FZ] = 0.0;
          prad_ff2lab(pp, gbl, *rp);
          double J[4][4];
          dxdx_bl_to_mks2(gbl.xx, mp, J);
          trans_pall_bl_to_mks2(pp, gbl, geom, J);
          double base = pp[KZ_RHO] * gbl.xx[1] / 4.0e-22;
          double acov3 = std::fmax(base * base - 0.02, 0.0) * std::sqrt(1.0e-23);
          acov3 *= std::sin((PI / 2.0 - gbl.xx[2]) / 0.1);
          pp[KZ_B1] = 0.0; pp[KZ_B2] = 0.0; pp[KZ_B3] = acov3;
        }
        {
          double gammam1 = gam - 1.0, indexn = 1.0 / gammam1, pre = gammam1 * pp[KZ_UU];
          pp[KZ_ENTR] = pp[KZ_RHO] * kzm_log(kzm_pow(pre, indexn) / kzm_pow(pp[KZ_RHO], indexn + 1.0));
        }
        memcpy(p + ci * KZ_NV, pp, sizeof pp);
      }
    }
  for (long iz = 1; iz < gr->nz; ++iz)
    for (long iy = 0; iy < gr->ny; ++iy)
      memcpy(p + G.cidx(0, iy, iz) * KZ_NV, p + G.cidx(0, iy, 0) * KZ_NV, (size_t)gr->nx * KZ_NV * sizeof(double));
  return 0;
}

// ct.zig:184-334 calcBfromA with overwrite: cell-centred A in the B slots of
// the padded p -> corner average -> curl -> B in every cell's B slots (ghost
// cells get 0, as the reference's zeroed scratch gives them).
int kz_host_calc_b_from_a(const kz_grid *gr, const double *g20, double *p) {
  Grid G(*gr);
  const long nx = gr->nx, ny = gr->ny, nz = gr->nz;
  if (ny <= 1) return KZ_E_UNSUPPORTED;
  std::vector<double> vp((size_t)G.ncell() * 6, 0.0);
  auto A = [&](int c, long ix, long iy, long iz) { return p[G.cidx(ix, iy, iz) * KZ_NV + KZ_B1 + c]; };
  // ct.zig:228-264 cornerAverageA
  for (long iz = 0; iz <= nz; ++iz)
    for (long iy = 0; iy <= ny; ++iy)
      for (long ix = 0; ix <= nx; ++ix) {
        if (nz == 1 && iz > 0) continue;
        for (int c = 0; c < 3; ++c)
          vp[G.cidx(ix, iy, iz) * 6 + c] =
              nz == 1 ? 1.0 / 4.0 * (A(c, ix, iy, iz) + A(c, ix, iy - 1, iz) + A(c, ix - 1, iy, iz) +
                                     A(c, ix - 1, iy - 1, iz))
                      : 1.0 / 8.0 * (A(c, ix, iy, iz) + A(c, ix, iy - 1, iz) + A(c, ix - 1, iy, iz) +
                                     A(c, ix - 1, iy - 1, iz) + A(c, ix, iy, iz - 1) + A(c, ix, iy - 1, iz - 1) +
                                     A(c, ix - 1, iy, iz - 1) + A(c, ix - 1, iy - 1, iz - 1));
      }
  // ct.zig:279-331 calcBfromACore
  auto S = [&](int c, long ix, long iy, long iz) { return vp[G.cidx(ix, iy, iz) * 6 + c]; };
  for (long iz = 0; iz < nz; ++iz)
    for (long iy = 0; iy < ny; ++iy)
      for (long ix = 0; ix < nx; ++ix) {
        double dA[4][4] = {{0}};
        for (int c = 0; c < 3; ++c) {
          if (nz == 1) {
            dA[c + 1][2] = -(S(c, ix, iy, iz) - S(c, ix, iy + 1, iz) + S(c, ix + 1, iy, iz) -
                             S(c, ix + 1, iy + 1, iz)) /
                           (2.0 * G.csize(iy, 1));
            dA[c + 1][1] = -(S(c, ix, iy, iz) + S(c, ix, iy + 1, iz) - S(c, ix + 1, iy, iz) -
                             S(c, ix + 1, iy + 1, iz)) /
                           (2.0 * G.csize(ix, 0));
            dA[c + 1][3] = 0;
          } else {
            dA[c + 1][1] = (S(c, ix + 1, iy, iz) - S(c, ix, iy, iz) + S(c, ix + 1, iy + 1, iz) -
                            S(c, ix, iy + 1, iz) + S(c, ix + 1, iy, iz + 1) - S(c, ix, iy, iz + 1) +
                            S(c, ix + 1, iy + 1, iz + 1) - S(c, ix, iy + 1, iz + 1)) /
                           (4.0 * G.csize(ix, 0));
            dA[c + 1][2] = (S(c, ix, iy + 1, iz) - S(c, ix, iy, iz) + S(c, ix + 1, iy + 1, iz) -
                            S(c, ix + 1, iy, iz) + S(c, ix, iy + 1, iz + 1) - S(c, ix, iy, iz + 1) +
                            S(c, ix + 1, iy + 1, iz + 1) - S(c, ix + 1, iy, iz + 1)) /
                           (4.0 * G.csize(iy, 1));
            dA[c + 1][3] = (S(c, ix, iy, iz + 1) - S(c, ix, iy, iz) + S(c, ix + 1, iy, iz + 1) -
                            S(c, ix + 1, iy, iz) + S(c, ix, iy + 1, iz + 1) - S(c, ix, iy + 1, iz) +
                            S(c, ix + 1, iy + 1, iz + 1) - S(c, ix + 1, iy + 1, iz)) /
                           (4.0 * G.csize(iz, 2));
          }
        }
        const long ci = G.cidx(ix, iy, iz);
        const double gdet = g20[ci * 20 + 19];
        vp[ci * 6 + 3] = (dA[2][3] - dA[3][2]) / gdet;
        vp[ci * 6 + 4] = (dA[3][1] - dA[1][3]) / gdet;
        vp[ci * 6 + 5] = (dA[1][2] - dA[2][1]) / gdet;
      }
  for (long c = 0; c < G.ncell(); ++c) {
    p[c * KZ_NV + KZ_B1] = vp[c * 6 + 3];
    p[c * KZ_NV + KZ_B2] = vp[c * 6 + 4];
    p[c * KZ_NV + KZ_B3] = vp[c * 6 + 5];
  }
  return 0;
}

// Seeded perturbation of a domain state: every domain primitive except
// B1..B3 is multiplied by (1 + amp * U(-1,1)), U from a splitmix64 hash of
// (seed, padded cell index, variable) mapped as puffy.zig:696-702 maps it;
// ENTR is then re-synced from (rho, u) with kz_math.
static uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}
double kz_host_uniform(uint64_t seed, uint64_t cell, uint64_t var) {
  uint64_t u = mix64(mix64(mix64(seed) ^ cell) ^ var);
  return 2.0 * ((double)(u >> 11) * 0x1.0p-53) - 1.0;
}
int kz_host_perturb(const kz_grid *gr, uint64_t seed, double amp, double gam, int rho_u_only, double *p) {
  Grid G(*gr);
  for (long iz = 0; iz < gr->nz; ++iz)
    for (long iy = 0; iy < gr->ny; ++iy)
      for (long ix = 0; ix < gr->nx; ++ix) {
        long ci = G.cidx(ix, iy, iz);
        double *pp = p + ci * KZ_NV;
        for (int iv = 0; iv < KZ_NV; ++iv) {
          if (iv >= KZ_B1 && iv <= KZ_B3) continue;
          if (iv == KZ_ENTR) continue;
          if (rho_u_only && iv != KZ_RHO && iv != KZ_UU) continue;
          pp[iv] *= 1.0 + amp * kz_host_uniform(seed, (uint64_t)ci, (uint64_t)iv);
        }
        double gammam1 = gam - 1.0, indexn = 1.0 / gammam1, pre = gammam1 * pp[KZ_UU];
        pp[KZ_ENTR] = pp[KZ_RHO] * kzm_log(kzm_pow(pre, indexn) / kzm_pow(pp[KZ_RHO], indexn + 1.0));
      }
  return 0;
}

// geometry of one point in BL (for tests of the xhi tables / init)
// The BL sidecar of MetricCache (precompute.zig:361-380, :457-475) over every
// padded cell: 52 doubles per cell = gBL[16] GBL[16] alpha_BL gdet_BL r theta
// J_mks2->bl[16].  The radiative viscosity reads r, the dynamo all of it.
int kz_host_bl_tables(const kz_grid *gr, double a, double mksr0, double mksh0, double *out) {
  Grid G(*gr);
  MP mp{a, mksr0, mksh0};
  const long nrows = (long)G.sz * G.sy;
#pragma omp parallel for schedule(static)
  for (long row = 0; row < nrows; ++row) {
    long iz = row / G.sy - G.ngz, iy = row % G.sy - G.ngy;
    for (long ix = -G.ngx; ix < gr->nx + G.ngx; ++ix) {
      double *o = out + G.cidx(ix, iy, iz) * 52;
      kz_geom b = geometry_bl_at(G, mp, ix, iy, iz);
      for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) {
          o[i * 4 + j] = b.gg[i][j];
          o[16 + i * 4 + j] = b.GG[i][j];
        }
      o[32] = b.alpha;
      o[33] = b.gdet;
      o[34] = b.xx[1];
      o[35] = b.xx[2];
      double x[4] = {0.0, G.xc(ix), G.yc(iy), G.zc(iz)}, J[4][4];
      dxdx_mks2_to_bl(x, mp, J);
      for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) o[36 + i * 4 + j] = J[i][j];
    }
  }
  return 0;
}

// scalars.zig:47-50 radiusBL: the BL radius of each radial index, at the
// centre of the first theta and phi cells
int kz_host_radius_bl(const kz_grid *gr, double a, double mksr0, double mksh0, double *r) {
  Grid G(*gr);
  MP mp{a, mksr0, mksh0};
  for (long ix = 0; ix < gr->nx; ++ix) {
    double x[4] = {0.0, G.xc(ix), G.yc(0), G.zc(0)}, o[4];
    coco_mks2_to_bl(x, mp, o);
    r[ix] = o[1];
  }
  return 0;
}

// sim.zig:481-495 (REDUCEORDERATBH): the BL radius of a cell centre grows
// with the radial index alone, so the cells inside the horizon r_H(a) =
// 1 + sqrt(1 - a^2) (metric.zig:20-23) form a prefix of the radial indices;
// the largest index of that prefix over the full radial extent, ghosts
// included, at the domain's middle theta row.  INT32_MIN when no centre is
// inside.
int32_t kz_host_reduce_order_ix_max(const kz_grid *gr, double a, double mksr0, double mksh0) {
  Grid G(*gr);
  MP mp{a, mksr0, mksh0};
  const double r_h = 1.0 + std::sqrt(1.0 - a * a);
  const double y_mid = G.yc(gr->ny / 2);
  int32_t ix_max = INT32_MIN;
  for (long jx = -gr->ng; jx < gr->nx + gr->ng; ++jx) {
    double x[4] = {0.0, G.xc(jx), y_mid, 0.0}, o[4];
    coco_mks2_to_bl(x, mp, o);
    if (o[1] < r_h) ix_max = (int32_t)jx;
  }
  return ix_max;
}

// scalars.zig:188-204 blThetaWidth / blPhiWidth at radial index ix, for every
// (iz, iy) of the shell: |theta_BL| across the two y faces at the x centre,
// |phi_BL| across the two z faces at the cell centre; row-major [nz][ny]
int kz_host_shell_widths_bl(const kz_grid *gr, double a, double mksr0, double mksh0, long ix, double *dth,
                            double *dph) {
  Grid G(*gr);
  MP mp{a, mksr0, mksh0};
  for (long iz = 0; iz < gr->nz; ++iz)
    for (long iy = 0; iy < gr->ny; ++iy) {
      double xa[4] = {0.0, G.xc(ix), G.yl(iy), G.zc(iz)}, xb[4] = {0.0, G.xc(ix), G.yl(iy + 1), G.zc(iz)};
      double oa[4], ob[4];
      coco_mks2_to_bl(xa, mp, oa);
      coco_mks2_to_bl(xb, mp, ob);
      dth[iz * gr->ny + iy] = std::fabs(ob[2] - oa[2]);
      double za[4] = {0.0, G.xc(ix), G.yc(iy), G.zl(iz)}, zb[4] = {0.0, G.xc(ix), G.yc(iy), G.zl(iz + 1)};
      coco_mks2_to_bl(za, mp, oa);
      coco_mks2_to_bl(zb, mp, ob);
      dph[iz * gr->ny + iy] = std::fabs(ob[3] - oa[3]);
    }
  return 0;
}

int kz_host_geometry_bl(const kz_grid *gr, double a, double mksr0, double mksh0, long ix, long iy, long iz,
                        kz_geom *out) {
  Grid G(*gr);
  MP mp{a, mksr0, mksh0};
  *out = geometry_bl_at(G, mp, ix, iy, iz);
  return 0;
}

}  // extern "C"
