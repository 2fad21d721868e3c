// kz_step.hip — the koral per-step hot path on one MI355X (gfx950).
//
// Device-resident state of one run (the reference's Sim(config.puffy),
// koral/sim.zig:93-1900) and the kernels of its passes:
//   wavespeeds + CFL fold       sim.zig:747-886
//   fused face flux             sim.zig:1209-1503 (sweep + fluxesAtFaces:
//                               PPM, face floors, fFluxPrime, p2u, LAXF in
//                               one thread per face; no face-array round trip)
//   flux-CT                     sim/ct.zig:38-177 (EMF, then B-flux rebuild)
//   conserved update            sim.zig:1140-1175, 1542-1582
//   u2p + fixups                sim.zig:924-1133
//   implicit four-force solve   sim.zig:1715-1776
//   boundary conditions         sim/bc.zig:40-377 + puffy.zig:1080-1194
//   polar-axis overwrite        sim/polaraxis.zig:84-147
//   RK2IMEX stage arithmetic    sim.zig:1586-1899
// Fields are SoA planes (field[iv * ncell + cell]) over the padded grid, the
// metric blocks are SoA planes of full 4x4 g, G plus gdet, gttpert, alpha.
// Reductions are max/min/integer sums only (order-insensitive: the result
// is bit-identical to the serial reference loop).
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../include/kz_step.h"
#include "kz_dev.h"
#include "kz_dev_solve.h"

using namespace kzd;

#define GEO_NC 35  // g[16] G[16] gdet gttpert alpha
#define BCX_REC 66

struct GridD {
  int nx, ny, nz, ng, ngx, ngy, ngz, sx, sy, sz;
  long ncell;
  double minx, miny, minz, dx, dy, dz;
  int fsx[3], fsy[3];
  long nface[3];
};

struct Dev {
  double *p, *u, *ut0, *ut1, *ut2, *dut1, *dut2, *drt1, *drt2, *uforget;
  double *flb[3];
  double *scal;  // planes: ahd_m[3], arad_m[3], tstepden
  int32_t *flags;
  double *emf[3];
  double *gc, *gf[3], *kris;
  double *bcx;
  unsigned long long *cnt;  // [0] n_fail [1] n_iters [2] n_solves [3] n_guard
  unsigned long long *tsd;  // [0] max bits [1] min bits
  int *err;
  double *ss;  // step scalars (see k_begin_step)
  double *stage;  // staging for AoS <-> SoA
  // split implicit solve: per domain cell (index t = ix + nx*(iy + ny*iz))
  double *imp_uu00;  // [13][nd] conserved state of the incoming primitives
  double *imp_pp0;   // [13][nd] Newton start (after the bisection warm start)
  double *imp_aux;   // [IMP_NAUX][nd] reference-state scalars, see ImpAux
  int32_t *imp_slow;           // cells that need rungs past the first
  unsigned long long *imp_nslow;
};
// imp_aux planes
enum { IA_EHAT = 0, IA_UINT, IA_SGAS, IA_KES, IA_KGA, IA_KRA, IA_KRR, IA_SW, IMP_NAUX };

enum { SC_AHD = 0, SC_ARAD = 3, SC_TSTEPDEN = 6, SC_N = 7 };
enum { SS_DT = 0, SS_DTG, SS_INV_DTG, SS_NINV_DTG, SS_INV_DT, SS_NINV_DT, SS_C12, SS_HALF, SS_T, SS_OWN_DT, SS_USER, SS_N };

static const double GAM_IMEX = 0x1.2bec333018867p-2;  // 1 - 1/sqrt(2), correctly rounded

// ---------------------------------------------------------------------------
// indexing helpers
__device__ __forceinline__ long cidx(const GridD &g, long ix, long iy, long iz) {
  return (ix + g.ngx) + (iy + g.ngy) * (long)g.sx + (iz + g.ngz) * (long)g.sy * g.sx;
}
__device__ __forceinline__ long fidx(const GridD &g, int d, long ix, long iy, long iz) {
  long jx = ix + g.ngx, jy = iy + g.ngy, jz = iz + g.ngz;
  return (jz * g.fsy[d] + jy) * (long)g.fsx[d] + jx;
}
__device__ __forceinline__ double xl_(const GridD &g, long i) { return g.minx + (double)i * g.dx; }
__device__ __forceinline__ double yl_(const GridD &g, long i) { return g.miny + (double)i * g.dy; }
__device__ __forceinline__ double zl_(const GridD &g, long i) { return g.minz + (double)i * g.dz; }
__device__ __forceinline__ double xc_(const GridD &g, long i) { return 0.5 * (xl_(g, i) + xl_(g, i + 1)); }
__device__ __forceinline__ double yc_(const GridD &g, long i) { return 0.5 * (yl_(g, i) + yl_(g, i + 1)); }
__device__ __forceinline__ double csize(const GridD &g, long i, int d) {
  return d == 0 ? xl_(g, i + 1) - xl_(g, i) : (d == 1 ? yl_(g, i + 1) - yl_(g, i) : zl_(g, i + 1) - zl_(g, i));
}
__device__ __forceinline__ void load_geo(const double *__restrict__ t, long stride, long i, double x1, Geo &gm) {
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      gm.g[a][b] = t[(a * 4 + b) * stride + i];
      gm.G[a][b] = t[(16 + a * 4 + b) * stride + i];
    }
  gm.gdet = t[32 * stride + i];
  gm.gttpert = t[33 * stride + i];
  gm.alpha = t[34 * stride + i];
  gm.x1 = x1;
}
__device__ __forceinline__ void geo_c(const GridD &g, const Dev &d, long ix, long iy, long iz, Geo &gm) {
  load_geo(d.gc, g.ncell, cidx(g, ix, iy, iz), xc_(g, ix), gm);
}
__device__ __forceinline__ void geo_f(const GridD &g, const Dev &d, int dim, long ix, long iy, long iz, Geo &gm) {
  load_geo(d.gf[dim], g.nface[dim], fidx(g, dim, ix, iy, iz), dim == 0 ? xl_(g, ix) : xc_(g, ix), gm);
}
__device__ __forceinline__ void ld(const double *__restrict__ f, long n, long c, double *pp) {
#pragma unroll
  for (int iv = 0; iv < NV; ++iv) pp[iv] = f[iv * n + c];
}
__device__ __forceinline__ void st(double *__restrict__ f, long n, long c, const double *pp) {
#pragma unroll
  for (int iv = 0; iv < NV; ++iv) f[iv * n + c] = pp[iv];
}
__device__ __forceinline__ void set_err(int *err, int code) {
  if (code) atomicCAS(err, 0, code);
}
__device__ __forceinline__ bool polar_row(const kz_physics &ph, const GridD &g, long iy) {
  if (!ph.correct_polaraxis) return false;
  long nc = ph.nccorrectpolar;
  return iy < nc || iy > g.ny - nc - 1;
}

#define TID (blockIdx.x * (long)blockDim.x + threadIdx.x)

// ---------------------------------------------------------------------------
// p2u over domain cells (Sim.initCell / postinit refresh)
__global__ void k_p2u_domain(GridD g, Dev d, double gam) {
  long t = TID;
  if (t >= (long)g.nx * g.ny * g.nz) return;
  long ix = t % g.nx, iy = (t / g.nx) % g.ny, iz = t / ((long)g.nx * g.ny);
  long c = cidx(g, ix, iy, iz);
  Geo gm;
  geo_c(g, d, ix, iy, iz, gm);
  double pp[NV], uu[NV];
  ld(d.p, g.ncell, c, pp);
  p2u(pp, uu, gm, gam);
  st(d.u, g.ncell, c, uu);
}

// ---------------------------------------------------------------------------
// boundary conditions
__device__ void tvec(const double *J, const double u[4], double o[4]) {
  double r[4] = {0, 0, 0, 0};
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) r[i] += J[i * 4 + j] * u[j];
#pragma unroll
  for (int i = 0; i < 4; ++i) o[i] = r[i];
}
// frames.zig transPmhdCocoJ
__device__ int trans_pmhd(double *pp, const Geo &g1, const Geo &g2, const double *J) {
  double u1c[4], u1v[4];
  ucon_ucov(pp, g1, u1c, u1v);
  double bcon[4], bcov[4], bsq;
  bcon_bcov_bsq(pp[B1], pp[B2], pp[B3], u1c, u1v, g1, bcon, bcov, &bsq);
  tvec(J, bcon, bcon);
  double uc[4];
  tvec(J, u1c, uc);
  vel4_to_velr(uc, g2, true, uc);
  pp[VX] = uc[1];
  pp[VY] = uc[2];
  pp[VZ] = uc[3];
  double u2c[4], u2v[4];
  ucon_ucov(pp, g2, u2c, u2v);
#pragma unroll
  for (int j = 1; j < 4; ++j) pp[B1 + j - 1] = bcon[j] * u2c[0] - bcon[0] * u2c[j];
  return E_OK;
}
__device__ void geo_from_rec(const double *rec, Geo &o) {
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      o.g[i][j] = rec[i * 4 + j];
      o.G[i][j] = rec[16 + i * 4 + j];
    }
  o.gdet = 0;
  o.alpha = kzm_sqrt(-1.0 / o.G[0][0]);
  o.gttpert = 0;
  o.x1 = 0;
}
// puffy.zig:1089-1166 (xhi) / :1167-1171 (xlo); t < nz*ny*ngx*2
__device__ void bc_x_cell(const GridD &g, const Dev &d, const kz_physics &ph, long t) {
  int side = t % 2;
  long i = (t / 2) % g.ngx + 1;
  long iy = (t / (2 * g.ngx)) % g.ny;
  long iz = t / (2L * g.ngx * g.ny);
  long nx = g.nx;
  long ix = side ? nx - 1 + i : -i;
  double pp[NV];
  Geo geom;
  geo_c(g, d, ix, iy, iz, geom);
  if (side == 0) {
    ld(d.p, g.ncell, cidx(g, 0, iy, iz), pp);
  } else {
    ld(d.p, g.ncell, cidx(g, nx - 1, iy, iz), pp);
    const double *rec = d.bcx + (((iz * g.ny) + iy) * g.ng + (ix - nx)) * BCX_REC;
    Geo gbl;
    geo_from_rec(rec, gbl);
    double rghost = rec[32], rbound = rec[33];
    const double *J12 = rec + 34, *J21 = rec + 50;
    trans_pmhd(pp, geom, gbl, J12);
    double s1 = rbound * rbound / rghost / rghost, s2 = rbound / rghost;
    pp[RHO] *= s1; pp[UU] *= s1;
    pp[B1] *= s1; pp[B2] *= s2; pp[B3] *= s2;
    pp[VY] *= s1; pp[VZ] *= s1;
    pp[EE] *= s1; pp[FY] *= s1; pp[FZ] *= s1;
    trans_pmhd(pp, gbl, geom, J21);
    double uc[4] = {0.0, pp[VX], pp[VY], pp[VZ]};
    velr_to_vel4(uc, geom, uc);
    tvec(J12, uc, uc);
    if (uc[1] < 0.0) {
      uc[1] = 0.0;
      tvec(J21, uc, uc);
      int e = vel4_to_velr(uc, geom, false, uc);
      if (e) { set_err(d.err, e); return; }
      pp[VX] = uc[1]; pp[VY] = uc[2]; pp[VZ] = uc[3];
    }
    double ur[4] = {0.0, pp[FX], pp[FY], pp[FZ]};
    velr_to_vel4(ur, geom, ur);
    tvec(J12, ur, ur);
    if (ur[1] < 0.0) {
      ur[1] = 0.0;
      tvec(J21, ur, ur);
      int e = vel4_to_velr(ur, geom, false, ur);
      if (e) { set_err(d.err, e); return; }
      pp[FX] = ur[1]; pp[FY] = ur[2]; pp[FZ] = ur[3];
    }
  }
  double uu[NV];
  p2u(pp, uu, geom, ph.gam);
  long c = cidx(g, ix, iy, iz);
  st(d.p, g.ncell, c, pp);
  st(d.u, g.ncell, c, uu);
}
// puffy.zig:1172-1187 polar reflection; t < nz*nx*ngy*2
__device__ void bc_y_cell(const GridD &g, const Dev &d, const kz_physics &ph, long t) {
  int side = t % 2;
  long i = (t / 2) % g.ngy + 1;
  long ix = (t / (2 * g.ngy)) % g.nx;
  long iz = t / (2L * g.ngy * g.nx);
  long ny = g.ny;
  long iy = side ? ny - 1 + i : -i;
  long iiy = side ? ny - (iy - ny) - 1 : -iy - 1;
  double pp[NV], uu[NV];
  ld(d.p, g.ncell, cidx(g, ix, iiy, iz), pp);
  pp[VY] = -pp[VY];
  pp[B2] = -pp[B2];
  pp[FY] = -pp[FY];
  Geo geom;
  geo_c(g, d, ix, iy, iz, geom);
  p2u(pp, uu, geom, ph.gam);
  long c = cidx(g, ix, iy, iz);
  st(d.p, g.ncell, c, pp);
  st(d.u, g.ncell, c, uu);
}
// x and y ghost cells in one launch: the x pass reads only domain cells of
// rows 0..ny-1 and writes x ghosts, the y pass reads only domain cells and
// writes y ghosts, so the two are independent (bc.zig:40-73 runs them in
// sequence).
__global__ void __launch_bounds__(64) k_bc_xy(GridD g, Dev d, kz_physics ph) {
  long t = TID;
  long nxg = (long)g.nz * g.ny * g.ngx * 2;
  long nyg = (long)g.nz * g.nx * g.ngy * 2;
  if (t < nxg) bc_x_cell(g, d, ph, t);
  else if (t < nxg + nyg) bc_y_cell(g, d, ph, t - nxg);
}

// bc.zig:286-377 fillCorners2d.  Per corner quadrant the reference copies
// 2(ng-1) cells from x/y ghosts, averages the diagonal cell next to the
// domain from an x and a y ghost, then averages the outer diagonal cell from
// two of the copied cells.  The first group reads no cell of its own group,
// so it runs one thread per cell; the outer diagonals follow a barrier.
// Quadrants never read each other's cells.  One block of 64 threads.
__device__ void cp_p2u(const GridD &g, const Dev &d, double gam, long dx_, long dy_, long sx_, long sy_) {
  double pp[NV], uu[NV];
  ld(d.p, g.ncell, cidx(g, sx_, sy_, 0), pp);
  Geo gm;
  geo_c(g, d, dx_, dy_, 0, gm);
  p2u(pp, uu, gm, gam);
  long c = cidx(g, dx_, dy_, 0);
  st(d.p, g.ncell, c, pp);
  st(d.u, g.ncell, c, uu);
}
__device__ void avg_p2u(const GridD &g, const Dev &d, double gam, long dx_, long dy_, long ax, long ay, long bx,
                        long by) {
  double pa[NV], pb[NV], pp[NV], uu[NV];
  ld(d.p, g.ncell, cidx(g, ax, ay, 0), pa);
  ld(d.p, g.ncell, cidx(g, bx, by, 0), pb);
#pragma unroll
  for (int iv = 0; iv < NV; ++iv) pp[iv] = 0.5 * (pa[iv] + pb[iv]);
  Geo gm;
  geo_c(g, d, dx_, dy_, 0, gm);
  p2u(pp, uu, gm, gam);
  long c = cidx(g, dx_, dy_, 0);
  st(d.p, g.ncell, c, pp);
  st(d.u, g.ncell, c, uu);
}
__global__ void __launch_bounds__(64) k_bc_corners2d(GridD g, Dev d, kz_physics ph) {
  const long nx = g.nx, ny = g.ny, ng = g.ng;
  const double gam = ph.gam;
  const int per = 2 * (g.ng - 1) + 1;  // copies + inner diagonal per quadrant
  const int tid = threadIdx.x;
  if (tid < 4 * per) {
    const int q = tid / per, k = tid % per;
    // quadrant q: bit 0 = high y side, bit 1 = high x side
    const bool hx = q & 2, hy = q & 1;
    const long xin = hx ? nx - 1 : 0, yin = hy ? ny - 1 : 0;    // edge domain index
    const long xg1 = hx ? nx : -1, yg1 = hy ? ny : -1;          // first ghost index
    if (k < 2 * (g.ng - 1)) {
      const long i = k / 2;
      if (k % 2 == 0) {
        // ghost column beyond x, first ghost row in y: copy from the edge row
        const long x = hx ? nx + i + 1 : -ng + i;
        cp_p2u(g, d, gam, x, yg1, x, yin);
      } else {
        const long y = hy ? ny + i + 1 : -ng + i;
        cp_p2u(g, d, gam, xg1, y, xin, y);
      }
    } else {
      avg_p2u(g, d, gam, xg1, yg1, xg1, yin, xin, yg1);
    }
  }
  __syncthreads();
  if (tid < 4) {
    const bool hx = tid & 2, hy = tid & 1;
    const long xg1 = hx ? nx : -1, yg1 = hy ? ny : -1;
    const long xg2 = hx ? nx + 1 : -2, yg2 = hy ? ny + 1 : -2;
    avg_p2u(g, d, gam, xg2, yg2, xg2, yg1, xg1, yg2);
  }
}

// ---------------------------------------------------------------------------
// u2p pass (sim.zig:940-1000)
__global__ void __launch_bounds__(128) k_u2p(GridD g, Dev d, kz_physics ph) {
  long t = TID;
  if (t >= (long)g.nx * g.ny * g.nz) return;
  long ix = t % g.nx, iy = (t / g.nx) % g.ny, iz = t / ((long)g.nx * g.ny);
  long c = cidx(g, ix, iy, iz);
  long n = g.ncell;
  double uu[NV], pp[NV];
  ld(d.u, n, c, uu);
  ld(d.p, n, c, pp);
  d.flags[KZ_FLAG_ENTROPY * n + c] = 0;
  d.flags[KZ_FLAG_ENTROPY2 * n + c] = 0;
  Geo gm;
  geo_c(g, d, ix, iy, iz, gm);
  if (polar_row(ph, g, iy)) {
    double gi = 1.0 / gm.gdet;
    d.p[B1 * n + c] = uu[B1] * gi;
    d.p[B2 * n + c] = uu[B2] * gi;
    d.p[B3 * n + c] = uu[B3] * gi;
    d.flags[KZ_FLAG_HD_FIXUP * n + c] = 0;
    d.flags[KZ_FLAG_RAD_FIXUP * n + c] = 0;
    return;
  }
  int ent;
  int ret = u2p_mhd(uu, pp, gm, ph.gam, ph.floors, &ent);
  if (ent) d.flags[KZ_FLAG_ENTROPY * n + c] = 1;
  int corr;
  u2p_rad(uu, pp, gm, ph.rad, &corr);
  int guard = 0;
  check_floors_mhd(pp, gm, ph.gam, ph.floors, &guard);
  check_floors_rad(pp, gm, ph.rad);
  st(d.p, n, c, pp);
  d.flags[KZ_FLAG_HD_FIXUP * n + c] = ret < -1 ? 1 : 0;
  d.flags[KZ_FLAG_RAD_FIXUP * n + c] = corr ? -1 : 0;
  if (guard) atomicAdd(&d.cnt[3], (unsigned long long)guard);
}

// cell_fixup (sim.zig:1027-1133), in place: a flagged cell reads only
// unflagged neighbours, which no thread writes, so the reference's
// copy-to-backup staging is unnecessary.
__global__ void __launch_bounds__(256) k_fixup(GridD g, Dev d, kz_physics ph, int which) {
  long t = TID;
  if (t >= (long)g.nx * g.ny * g.nz) return;
  long ix = t % g.nx, iy = (t / g.nx) % g.ny, iz = t / ((long)g.nx * g.ny);
  long n = g.ncell;
  long c = cidx(g, ix, iy, iz);
  if (d.flags[which * n + c] == 0) return;
  if (polar_row(ph, g, iy)) return;
  long nx = g.nx, ny = g.ny, nz = g.nz;
  long nb[6][3] = {{ix - 1, iy, iz}, {ix + 1, iy, iz}, {ix, iy - 1, iz}, {ix, iy + 1, iz}, {ix, iy, iz - 1},
                   {ix, iy, iz + 1}};
  long ci[6];
  int in_n = 0;
  for (int k = 0; k < 6; ++k) {
    if (nb[k][0] < 0 || nb[k][0] >= nx || nb[k][1] < 0 || nb[k][1] >= ny || nb[k][2] < 0 || nb[k][2] >= nz) continue;
    long cc = cidx(g, nb[k][0], nb[k][1], nb[k][2]);
    if (d.flags[which * n + cc] != 0) continue;
    ci[in_n++] = cc;
  }
  bool enough = (nz == 1 && ny == 1 && in_n >= 1) || (nz == 1 && in_n >= 2) || (ny == 1 && in_n >= 2) || in_n >= 3;
  if (!enough) return;
  double pp[NV], uu[NV];
  ld(d.p, n, c, pp);
  for (int iv = 0; iv < NV; ++iv) {
    bool fix;
    if (which == KZ_FLAG_HD_FIXUP) fix = iv != RHO && iv < B1;
    else if (which == KZ_FLAG_RAD_FIXUP) fix = iv >= EE && iv <= FZ;
    else fix = iv != RHO && (iv < B1 || (iv >= EE && iv <= FZ));
    if (fix) {
      double s = 0;
      for (int k = 0; k < in_n; ++k) s += d.p[iv * n + ci[k]];
      pp[iv] = s / (double)in_n;
    }
  }
  Geo gm;
  geo_c(g, d, ix, iy, iz, gm);
  p2u(pp, uu, gm, ph.gam);
  st(d.p, n, c, pp);
  st(d.u, n, c, uu);
}

// ---------------------------------------------------------------------------
// wavespeeds + CFL denominators (sim.zig:747-886)
__device__ __forceinline__ long long dbits(double x) { return __double_as_longlong(x); }
__global__ void __launch_bounds__(128) k_wavespeeds(GridD g, Dev d, kz_physics ph) {
  long t = TID;
  int act[3] = {g.nx > 1, g.ny > 1, g.nz > 1};
  long lx = act[0], ly = act[1], lz = act[2];
  long wx = g.nx + 2 * lx, wy = g.ny + 2 * ly, wz = g.nz + 2 * lz;
  if (t >= wx * wy * wz) return;
  long ix = t % wx - lx, iy = (t / wx) % wy - ly, iz = t / (wx * wy) - lz;
  long n = g.ncell;
  long c = cidx(g, ix, iy, iz);
  double pp[NV];
  ld(d.p, n, c, pp);
  Geo gm;
  geo_c(g, d, ix, iy, iz, gm);
  double aaa[6], av[12], rad0[6], radl[6];
  gas_wavespeeds(pp, gm, ph.gam, act, aaa);
  double tau[3] = {0, 0, 0};
  if (ph.has_opac) {
    double chi = calc_chi_slim(pp, gm, ph.gam, ph.opac);
    long idx3[3] = {ix, iy, iz};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      double sg = kzm_sqrt(gm.g[k + 1][k + 1]);
      double dxp = kzm_fmax(csize(g, idx3[k], k) * sg, csize(g, idx3[k] + 1, k) * sg);
      tau[k] = chi * dxp;
    }
  }
  if (ph.dampradwavespeednearaxis > 0) {
    long nc = ph.dampradwavespeednearaxis;
    if (iy < nc || iy >= g.ny - nc) tau[0] = tau[1] = tau[2] = 0;
  }
  rad_wavespeeds(pp, gm, tau, act, av);
#pragma unroll
  for (int i = 0; i < 6; ++i) {
    rad0[i] = av[i];
    radl[i] = av[6 + i];
  }
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    if (radl[2 * k] > 0.0) radl[2 * k] = 0.0;
    if (radl[2 * k + 1] < 0.0) radl[2 * k + 1] = 0.0;
    if (rad0[2 * k] > 0.0) rad0[2 * k] = 0.0;
    if (rad0[2 * k + 1] < 0.0) rad0[2 * k + 1] = 0.0;
  }
  double ax = kzm_fmax(kzm_fabs(aaa[0]), kzm_fabs(aaa[1]));
  double ay = kzm_fmax(kzm_fabs(aaa[2]), kzm_fabs(aaa[3]));
  double az = kzm_fmax(kzm_fabs(aaa[4]), kzm_fabs(aaa[5]));
  d.scal[(SC_AHD + 0) * n + c] = ax;
  d.scal[(SC_AHD + 1) * n + c] = ay;
  d.scal[(SC_AHD + 2) * n + c] = az;
#pragma unroll
  for (int k = 0; k < 3; ++k)
    d.scal[(SC_ARAD + k) * n + c] = kzm_fmax(kzm_fabs(radl[2 * k]), kzm_fabs(radl[2 * k + 1]));
  double wsx = kzm_fmax(ax, kzm_fmax(kzm_fabs(rad0[0]), kzm_fabs(rad0[1])));
  double wsy = kzm_fmax(ay, kzm_fmax(kzm_fabs(rad0[2]), kzm_fabs(rad0[3])));
  double wsz = kzm_fmax(az, kzm_fmax(kzm_fabs(rad0[4]), kzm_fabs(rad0[5])));
  if (!(ix >= 0 && ix < g.nx && iy >= 0 && iy < g.ny && iz >= 0 && iz < g.nz)) return;
  double dx = csize(g, ix, 0), dy = csize(g, iy, 1), dz = csize(g, iz, 2);
  double tsd;
  if (g.nz > 1 && g.ny > 1) tsd = wsx / dx + wsy / dy + wsz / dz;
  else if (g.nz == 1 && g.ny > 1) tsd = wsx / dx + wsy / dy;
  else if (g.ny == 1 && g.nz > 1) tsd = wsx / dx + wsz / dz;
  else tsd = wsx / dx;
  tsd /= ph.tsteplim;
  d.scal[SC_TSTEPDEN * n + c] = tsd;
  if (tsd == tsd) {
    atomicMax((long long *)&d.tsd[0], dbits(tsd));
    atomicMin((long long *)&d.tsd[1], dbits(tsd));
  }
}

// ---------------------------------------------------------------------------
// fused face flux: PPM of the two adjacent cells, face floors, fFluxPrime,
// p2u, LAXF (sim.zig:1284-1503).  One thread per face i in [0, n] of
// dimension `dim`, cross indices over [-1, n_c + 1) in active dimensions.
__global__ void __launch_bounds__(256) k_face_flux(GridD g, Dev d, kz_physics ph, int dim) {
  long t = TID;
  long nd = dim == 0 ? g.nx : (dim == 1 ? g.ny : g.nz);
  int c0d = dim == 0 ? 1 : 0, c1d = dim == 2 ? 1 : 2;
  long n0 = c0d == 1 ? g.ny : g.nx, n1 = c1d == 1 ? g.ny : g.nz;
  long l0 = n0 > 1 ? -1 : 0, h0 = n0 > 1 ? n0 + 1 : n0;
  long l1 = n1 > 1 ? -1 : 0, h1 = n1 > 1 ? n1 + 1 : n1;
  long w0 = h0 - l0, w1 = h1 - l1, wf = nd + 1;
  if (t >= wf * w0 * w1) return;
  // x-innermost thread order (coalesced plane loads)
  long i, c0, c1;
  if (dim == 0) {
    i = t % wf;
    c0 = (t / wf) % w0 + l0;
    c1 = t / (wf * w0) + l1;
  } else {
    c0 = t % w0 + l0;
    i = (t / w0) % wf;
    c1 = t / (w0 * wf) + l1;
  }
  long cell[3];
  cell[dim] = i;
  cell[c0d] = c0;
  cell[c1d] = c1;
  long n = g.ncell;
  long stride = dim == 0 ? 1 : (dim == 1 ? g.sx : (long)g.sx * g.sy);
  long cf = cidx(g, cell[0], cell[1], cell[2]);
  double dxs[6];
#pragma unroll
  for (int k = 0; k < 6; ++k) dxs[k] = csize(g, i - 3 + k, dim);
  double pl[NV], pr[NV];  // pl: left state at face i (cell i-1's right); pr: right state (cell i's left)
#pragma unroll
  for (int iv = 0; iv < NV; ++iv) {
    const double *f = d.p + iv * n;
    double s[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) s[k] = f[cf + (k - 3) * stride];
    double l0_, r0_, l1_, r1_;
    ppm(s, dxs, &l0_, &r0_);          // cell i-1: stencil i-3..i+1
    ppm(s + 1, dxs + 1, &l1_, &r1_);  // cell i:   stencil i-2..i+2
    pl[iv] = r0_;
    pr[iv] = l1_;
  }
  Geo gm;
  geo_f(g, d, dim, cell[0], cell[1], cell[2], gm);
  int guard = 0;
  check_floors_mhd(pl, gm, ph.gam, ph.floors, &guard);
  check_floors_mhd(pr, gm, ph.gam, ph.floors, &guard);
  double fll[NV], flr[NV], ul[NV], ur[NV];
  int e1 = flux_prime(pl, dim, gm, ph.gam, fll);
  int e2 = flux_prime(pr, dim, gm, ph.gam, flr);
  if (e1 | e2) set_err(d.err, E_NANFLUX);
  p2u(pl, ul, gm, ph.gam);
  p2u(pr, ur, gm, ph.gam);
  long cm = cf - stride;
  double ag = kzm_fmax(d.scal[(SC_AHD + dim) * n + cf], d.scal[(SC_AHD + dim) * n + cm]);
  double agr = kzm_fmax(d.scal[(SC_ARAD + dim) * n + cf], d.scal[(SC_ARAD + dim) * n + cm]);
  long fi = fidx(g, dim, cell[0], cell[1], cell[2]);
  long nfd = g.nface[dim];
  double *fo = d.flb[dim];
#pragma unroll
  for (int iv = 0; iv < NV; ++iv) fo[iv * nfd + fi] = laxf(fll[iv], flr[iv], ul[iv], ur[iv], iv >= EE ? agr : ag);
  if (guard) atomicAdd(&d.cnt[3], (unsigned long long)guard);
}

// ---------------------------------------------------------------------------
// flux-CT (ct.zig:38-177)
__device__ __forceinline__ long pin(long n, long i) { return n == 1 ? 0 : i; }
__device__ __forceinline__ double fb(const GridD &g, const Dev &d, int dim, int iv, long ix, long iy, long iz) {
  return d.flb[dim][iv * g.nface[dim] + fidx(g, dim, ix, iy, iz)];
}
__device__ __forceinline__ long eidx(const GridD &g, long ix, long iy, long iz) {
  return (iz * (g.ny + 1) + iy) * (long)(g.nx + 1) + ix;
}
__global__ void k_emf(GridD g, Dev d) {
  long t = TID;
  long nx = g.nx, ny = g.ny, nz = g.nz;
  if (t >= (nx + 1) * (ny + 1) * (nz + 1)) return;
  long ix = t % (nx + 1), iy = (t / (nx + 1)) % (ny + 1), iz = t / ((nx + 1) * (ny + 1));
  double c1 = (ny > 1 && nz > 1) ? 0.25 : 0.5, c2 = (nx > 1 && nz > 1) ? 0.25 : 0.5,
         c3 = (nx > 1 && ny > 1) ? 0.25 : 0.5;
  long e = eidx(g, ix, iy, iz);
  if (ny > 1 || nz > 1) {
    double s = 0;
    if (ny > 1) s += fb(g, d, 1, B3, pin(nx, ix), pin(ny, iy), pin(nz, iz)) + fb(g, d, 1, B3, pin(nx, ix), pin(ny, iy), pin(nz, iz - 1));
    if (nz > 1) s -= fb(g, d, 2, B2, pin(nx, ix), pin(ny, iy), pin(nz, iz)) + fb(g, d, 2, B2, pin(nx, ix), pin(ny, iy - 1), pin(nz, iz));
    d.emf[0][e] = c1 * s;
  } else {
    d.emf[0][e] = 0;
  }
  if (nx > 1 || nz > 1) {
    double s = 0;
    if (nz > 1) s += fb(g, d, 2, B1, pin(nx, ix), pin(ny, iy), pin(nz, iz)) + fb(g, d, 2, B1, pin(nx, ix - 1), pin(ny, iy), pin(nz, iz));
    if (nx > 1) s -= fb(g, d, 0, B3, ix, pin(ny, iy), pin(nz, iz)) + fb(g, d, 0, B3, ix, pin(ny, iy), pin(nz, iz - 1));
    d.emf[1][e] = c2 * s;
  } else {
    d.emf[1][e] = 0;
  }
  if (nx > 1 || ny > 1) {
    double s = 0;
    if (nx > 1) s += fb(g, d, 0, B2, ix, pin(ny, iy), pin(nz, iz)) + fb(g, d, 0, B2, ix, pin(ny, iy - 1), pin(nz, iz));
    if (ny > 1) s -= fb(g, d, 1, B1, pin(nx, ix), pin(ny, iy), pin(nz, iz)) + fb(g, d, 1, B1, pin(nx, ix - 1), pin(ny, iy), pin(nz, iz));
    d.emf[2][e] = c3 * s;
  } else {
    d.emf[2][e] = 0;
  }
}
__global__ void k_ct_rebuild(GridD g, Dev d) {
  long t = TID;
  long nx = g.nx, ny = g.ny, nz = g.nz;
  if (t >= (nx + 1) * (ny + 1) * (nz + 1)) return;
  long ix = t % (nx + 1), iy = (t / (nx + 1)) % (ny + 1), iz = t / ((nx + 1) * (ny + 1));
#define EMF(c, a, b, cc) d.emf[(c)-1][eidx(g, a, b, cc)]
  if (nx > 1 && iy < ny && iz < nz) {
    long f = fidx(g, 0, ix, iy, iz), m = g.nface[0];
    d.flb[0][B1 * m + f] = 0;
    d.flb[0][B2 * m + f] = 0.5 * (EMF(3, ix, iy, iz) + EMF(3, ix, iy + 1, iz));
    d.flb[0][B3 * m + f] = -0.5 * (EMF(2, ix, iy, iz) + EMF(2, ix, iy, iz + 1));
  }
  if (ny > 1 && ix < nx && iz < nz) {
    long f = fidx(g, 1, ix, iy, iz), m = g.nface[1];
    d.flb[1][B1 * m + f] = -0.5 * (EMF(3, ix, iy, iz) + EMF(3, ix + 1, iy, iz));
    d.flb[1][B2 * m + f] = 0;
    d.flb[1][B3 * m + f] = 0.5 * (EMF(1, ix, iy, iz) + EMF(1, ix, iy, iz + 1));
  }
  if (nz > 1 && ix < nx && iy < ny) {
    long f = fidx(g, 2, ix, iy, iz), m = g.nface[2];
    d.flb[2][B1 * m + f] = 0.5 * (EMF(2, ix, iy, iz) + EMF(2, ix + 1, iy, iz));
    d.flb[2][B2 * m + f] = -0.5 * (EMF(1, ix, iy, iz) + EMF(1, ix, iy + 1, iz));
    d.flb[2][B3 * m + f] = 0;
  }
#undef EMF
}

// ---------------------------------------------------------------------------
// conserved update: flux divergence + metric source (sim.zig:1140-1175,
// 1542-1582). dt is read from the step scalars (slot `sslot`).
__global__ void __launch_bounds__(128) k_update(GridD g, Dev d, kz_physics ph, int sslot) {
  long t = TID;
  if (t >= (long)g.nx * g.ny * g.nz) return;
  long ix = t % g.nx, iy = (t / g.nx) % g.ny, iz = t / ((long)g.nx * g.ny);
  long n = g.ncell;
  long c = cidx(g, ix, iy, iz);
  double dt = d.ss[sslot];
  Geo gm;
  geo_c(g, d, ix, iy, iz, gm);
  double pp[NV];
  ld(d.p, n, c, pp);
  double tij[4][4], tt[4][4], rup[4][4], rij[4][4];
  {
    double ucon[4], ucov[4];
    ucon_ucov(pp, gm, ucon, ucov);
    double bcon[4], bcov[4], bsq;
    bcon_bcov_bsq(pp[B1], pp[B2], pp[B3], ucon, ucov, gm, bcon, bcov, &bsq);
    double rho = pp[RHO], u = pp[UU];
    double pr = (ph.gam - 1.0) * u;
    double w = rho + u + pr;
    double eta = w + bsq;
    double ptot = pr + 0.5 * bsq;
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b) tij[a][b] = eta * ucon[a] * ucon[b] + ptot * gm.G[a][b] - bcon[a] * bcon[b];
  }
  lower_second(tij, gm, tt);
  calc_rij(pp, gm, rup);
  lower_second(rup, gm, rij);
  double ms[NV];
#pragma unroll
  for (int i = 0; i < NV; ++i) ms[i] = 0;
  const double *kr = d.kris;
  double gdetu = gm.gdet;
#pragma unroll
  for (int k = 0; k < 4; ++k)
#pragma unroll
    for (int l = 0; l < 4; ++l)
#pragma unroll
      for (int nu = 0; nu < 4; ++nu) {
        double K = kr[(l * 16 + nu * 4 + k) * n + c];
        ms[UU + nu] += gdetu * tt[k][l] * K;
        ms[EE + nu] += gdetu * rij[k][l] * K;
      }
  double dx = csize(g, ix, 0), dy = csize(g, iy, 1), dz = csize(g, iz, 2);
  long fx0 = fidx(g, 0, ix, iy, iz), fx1 = fidx(g, 0, ix + 1, iy, iz);
  long fy0 = fidx(g, 1, ix, iy, iz), fy1 = fidx(g, 1, ix, iy + 1, iz);
  long fz0 = fidx(g, 2, ix, iy, iz), fz1 = fidx(g, 2, ix, iy, iz + 1);
  long m0 = g.nface[0], m1 = g.nface[1], m2 = g.nface[2];
#pragma unroll
  for (int iv = 0; iv < NV; ++iv) {
    double du = -(d.flb[0][iv * m0 + fx1] - d.flb[0][iv * m0 + fx0]) * dt / dx -
                (d.flb[1][iv * m1 + fy1] - d.flb[1][iv * m1 + fy0]) * dt / dy -
                (d.flb[2][iv * m2 + fz1] - d.flb[2][iv * m2 + fz0]) * dt / dz;
    double uo = d.u[iv * n + c];
    d.u[iv * n + c] = uo + du + ms[iv] * dt;
  }
}

// ---------------------------------------------------------------------------
// implicit four-force solve (sim.zig:1736-1776 -> implicit.zig:927-1004),
// split in three kernels so that each keeps its state in registers:
//   k_imp_prep     reference state, bisection warm start, starting block
//   k_imp_rung0<W> the first rung (energy equation, lab frame) for the
//                  cells whose starting primitive block is W
//   k_imp_slow     rungs 1..5 (and damped-opacity levels) for the cells the
//                  first rung did not solve, gathered into a list
// The sequence of attempts per cell is the reference ladder's; inputs to
// every rung are the stored reference-state values, so results equal the
// single-pass ladder bit for bit.
__device__ __forceinline__ long dom_t(const GridD &g, long ix, long iy, long iz) {
  return ix + (long)g.nx * (iy + (long)g.ny * iz);
}
__global__ void __launch_bounds__(64) k_imp_prep(GridD g, Dev d, kz_physics ph, int sslot) {
  long t = TID;
  if (t == 0) *d.imp_nslow = 0;  // no appends run concurrently with this kernel
  long nd = (long)g.nx * g.ny * g.nz;
  if (t >= nd) return;
  long ix = t % g.nx, iy = (t / g.nx) % g.ny, iz = t / ((long)g.nx * g.ny);
  if (polar_row(ph, g, iy)) return;
  long n = g.ncell;
  long c = cidx(g, ix, iy, iz);
  double dt = d.ss[sslot];
  double pp00[NV], uu00[NV], pp0[NV];
  ld(d.p, n, c, pp00);
  Geo gm;
  geo_c(g, d, ix, iy, iz, gm);
  p2u(pp00, uu00, gm, ph.gam);
  double ehat = -calc_ff_rtt(pp00, gm);
  RadState st00;
  fill_radstate(pp00, gm, ph.gam, ph.opac, true, st00);
  ImpCtx cx{&gm, ph.gam, &ph.rad, &ph.opac, &ph.implicit, 1.0};
  if (ph.implicit.start_with_bisect) {
    solve_1d_prim(pp00, st00.ehat, st00.ucon[0], st00.tgas, dt, cx, pp0);
  } else {
#pragma unroll
    for (int i = 0; i < NV; ++i) pp0[i] = pp00[i];
  }
  int sw = (ehat < ph.implicit.threshold * pp00[UU]) ? WP_RAD : WP_MHD;
  st(d.imp_uu00, nd, t, uu00);
  st(d.imp_pp0, nd, t, pp0);
  double *a = d.imp_aux;
  a[IA_EHAT * nd + t] = st00.ehat;
  a[IA_UINT * nd + t] = st00.uint;
  a[IA_SGAS * nd + t] = st00.sgas;
  a[IA_KES * nd + t] = st00.kappaes;
  a[IA_KGA * nd + t] = st00.op.gas_abs;
  a[IA_KRA * nd + t] = st00.op.rad_abs;
  a[IA_KRR * nd + t] = st00.op.rad_ross;
  a[IA_SW * nd + t] = (double)sw;
}
__device__ __forceinline__ void imp_load(const Dev &d, long nd, long t, double *uu00, double *pp0, State0 &s0,
                                         LagOpac &lag) {
  ld(d.imp_uu00, nd, t, uu00);
  ld(d.imp_pp0, nd, t, pp0);
  const double *a = d.imp_aux;
  s0.ehat = a[IA_EHAT * nd + t];
  s0.uint = a[IA_UINT * nd + t];
  s0.sgas = a[IA_SGAS * nd + t];
  lag.kappaes = a[IA_KES * nd + t];
  lag.op.gas_abs = a[IA_KGA * nd + t];
  lag.op.rad_abs = a[IA_KRA * nd + t];
  lag.op.rad_ross = a[IA_KRR * nd + t];
}
// store a solved cell (sim.zig:1762-1768)
__device__ __forceinline__ void imp_commit(const GridD &g, const Dev &d, const kz_physics &ph, long c, const Geo &gm,
                                           const double *ppw, int iters) {
  double uu[NV];
  p2u(ppw, uu, gm, ph.gam);
  st(d.u, g.ncell, c, uu);
  st(d.p, g.ncell, c, ppw);
  d.flags[KZ_FLAG_RADIMP_FIXUP * g.ncell + c] = 0;
  atomicAdd(&d.cnt[1], (unsigned long long)iters);
  atomicAdd(&d.cnt[2], 1ULL);
}
template <int WP>
__global__ void __launch_bounds__(64) k_imp_rung0(GridD g, Dev d, kz_physics ph, int sslot) {
  long t = TID;
  long nd = (long)g.nx * g.ny * g.nz;
  if (t >= nd) return;
  long ix = t % g.nx, iy = (t / g.nx) % g.ny, iz = t / ((long)g.nx * g.ny);
  if (polar_row(ph, g, iy)) return;
  if ((int)d.imp_aux[IA_SW * nd + t] != WP) return;
  double uu00[NV], pp0[NV], ppw[NV];
  State0 s0;
  LagOpac lag;
  imp_load(d, nd, t, uu00, pp0, s0, lag);
  Geo gm;
  geo_c(g, d, ix, iy, iz, gm);
  double dt = d.ss[sslot];
  ImpCtx cx{&gm, ph.gam, &ph.rad, &ph.opac, &ph.implicit, 1.0};
  int iters = 0;
  bool ok = ph.implicit.opdamp_maxlevels >= 0 &&
            solve_4d_prim_core<WP>(uu00, lag, s0, pp0, dt, cx, WE_ENERGY, WF_LAB, ppw, &iters);
  if (ok) {
    imp_commit(g, d, ph, cidx(g, ix, iy, iz), gm, ppw, iters);
  } else {
    unsigned long long k = atomicAdd(d.imp_nslow, 1ULL);
    d.imp_slow[k] = (int32_t)t;
  }
}
__global__ void __launch_bounds__(64) k_imp_slow(GridD g, Dev d, kz_physics ph, int sslot) {
  long i = TID;
  if (i >= (long)*d.imp_nslow) return;
  long nd = (long)g.nx * g.ny * g.nz;
  long t = d.imp_slow[i];
  long ix = t % g.nx, iy = (t / g.nx) % g.ny, iz = t / ((long)g.nx * g.ny);
  long c = cidx(g, ix, iy, iz);
  double uu00[NV], pp0[NV];
  State0 s0;
  LagOpac lag;
  imp_load(d, nd, t, uu00, pp0, s0, lag);
  const int sw = (int)d.imp_aux[IA_SW * nd + t];
  const int ot = (sw == WP_RAD) ? WP_MHD : WP_RAD;
  Geo gm;
  geo_c(g, d, ix, iy, iz, gm);
  double dt = d.ss[sslot];
  const kz_implicit_params &ip = ph.implicit;
  ImpCtx cx{&gm, ph.gam, &ph.rad, &ph.opac, &ip, 1.0};
  for (int level = 0; level <= ip.opdamp_maxlevels; ++level) {
    double od = 1.0;
    for (int k = 0; k < level; ++k) od /= ip.opdamp_factor;
    cx.opdamp = od;
    for (int ir = level == 0 ? 1 : 0; ir < 6; ++ir) {
      int wp = (ir == 0 || ir == 1 || ir == 4) ? sw : ot;
      int we = (ir >= 4) ? WE_ENTROPY : WE_ENERGY;
      int wf = (ir == 0 || ir == 2) ? WF_LAB : WF_FF;
      double ppw[NV];
      int iters = 0;
      bool ok = (wp == WP_MHD) ? solve_4d_prim_core<WP_MHD>(uu00, lag, s0, pp0, dt, cx, we, wf, ppw, &iters)
                               : solve_4d_prim_core<WP_RAD>(uu00, lag, s0, pp0, dt, cx, we, wf, ppw, &iters);
      if (ok) {
        imp_commit(g, d, ph, c, gm, ppw, iters);
        return;
      }
    }
  }
  d.flags[KZ_FLAG_RADIMP_FIXUP * g.ncell + c] = -1;
  atomicAdd(&d.cnt[0], 1ULL);
}

// ---------------------------------------------------------------------------
// polar-axis overwrite (polaraxis.zig:84-147)
__global__ void k_polar(GridD g, Dev d, kz_physics ph) {
  long t = TID;
  long nc = ph.nccorrectpolar;
  if (t >= (long)g.nx * nc * 2 * g.nz) return;
  long ix = t % g.nx, ic = (t / g.nx) % nc, side = (t / (g.nx * nc)) % 2, iz = t / (g.nx * nc * 2);
  long ny = g.ny;
  long iy = side ? ny - 1 - ic : ic;
  long iysrc = iy < nc ? nc : ny - 1 - nc;
  double thaxis = yl_(g, iy < nc ? 0 : ny);
  double fac = kzm_fabs((yc_(g, iy) - thaxis) / (yc_(g, iysrc) - thaxis));
  long n = g.ncell;
  long c = cidx(g, ix, iy, iz), cs = cidx(g, ix, iysrc, iz);
  double pp[NV], ps[NV], uu[NV];
  ld(d.p, n, c, pp);
  ld(d.p, n, cs, ps);
  pp[RHO] = ps[RHO]; pp[UU] = ps[UU]; pp[ENTR] = ps[ENTR];
  pp[VX] = ps[VX]; pp[VZ] = ps[VZ]; pp[VY] = fac * ps[VY];
  pp[EE] = ps[EE]; pp[FX] = ps[FX]; pp[FZ] = ps[FZ]; pp[FY] = fac * ps[FY];
  Geo gm;
  geo_c(g, d, ix, iy, iz, gm);
  p2u(pp, uu, gm, ph.gam);
  st(d.p, n, c, pp);
  st(d.u, n, c, uu);
}

// ---------------------------------------------------------------------------
// stage arithmetic over domain cells (exact reference shapes)
__global__ void k_stage_deriv(GridD g, double *dst, const double *a, const double *b, const double *ss, int f1s,
                              int f2s) {
  long t = TID;
  long nd = (long)g.nx * g.ny * g.nz;
  if (t >= nd * NV) return;
  long cellt = t % nd;
  int iv = t / nd;
  long ix = cellt % g.nx, iy = (cellt / g.nx) % g.ny, iz = cellt / ((long)g.nx * g.ny);
  long o = iv * g.ncell + cidx(g, ix, iy, iz);
  dst[o] = ss[f1s] * a[o] + ss[f2s] * b[o];
}
__global__ void k_stage_combine(GridD g, double *dst, const double *a, const double *b, const double *c,
                                const double *ss, int f1s, int f2s) {
  long t = TID;
  long nd = (long)g.nx * g.ny * g.nz;
  if (t >= nd * NV) return;
  long cellt = t % nd;
  int iv = t / nd;
  long ix = cellt % g.nx, iy = (cellt / g.nx) % g.ny, iz = cellt / ((long)g.nx * g.ny);
  long o = iv * g.ncell + cidx(g, ix, iy, iz);
  dst[o] = a[o] + ss[f1s] * b[o] + ss[f2s] * c[o];
}
__global__ void k_copy_entropy_flag(GridD g, Dev d) {
  long t = TID;
  if (t >= (long)g.nx * g.ny * g.nz) return;
  long ix = t % g.nx, iy = (t / g.nx) % g.ny, iz = t / ((long)g.nx * g.ny);
  long c = cidx(g, ix, iy, iz);
  d.flags[KZ_FLAG_ENTROPY3 * g.ncell + c] = d.flags[KZ_FLAG_ENTROPY * g.ncell + c];
}
// u2p.c update_entropy (sim.zig:1683-1709)
__global__ void k_update_entropy(GridD g, Dev d, kz_physics ph) {
  long t = TID;
  if (t >= (long)g.nx * g.ny * g.nz) return;
  long ix = t % g.nx, iy = (t / g.nx) % g.ny, iz = t / ((long)g.nx * g.ny);
  long n = g.ncell;
  long c = cidx(g, ix, iy, iz);
  double pp[NV], uu[NV];
  ld(d.p, n, c, pp);
  ld(d.u, n, c, uu);
  pp[ENTR] = s_from_u(pp[RHO], pp[UU], ph.gam);
  d.p[ENTR * n + c] = pp[ENTR];
  Geo gm;
  geo_c(g, d, ix, iy, iz, gm);
  p2u_mhd(pp, uu, gm, ph.gam);
#pragma unroll
  for (int iv = 0; iv <= B3; ++iv) d.u[iv * n + c] = uu[iv];
}

// step scalars (sim.zig:1792-1830): own_dt = 1/tstepdenmax, dt = forced or
// own, and the stage coefficients with the reference's expression shapes;
// then the CFL accumulators are reset.
__global__ void k_begin_step(Dev d, double forced, int use_forced) {
  double *ss = d.ss;
  double tmax = __longlong_as_double((long long)d.tsd[0]);
  double own = 1.0 / tmax;
  double dt = use_forced ? forced : own;
  const double gm = GAM_IMEX;
  ss[SS_OWN_DT] = own;
  ss[SS_DT] = dt;
  ss[SS_DTG] = dt * gm;
  ss[SS_INV_DTG] = 1.0 / (dt * gm);
  ss[SS_NINV_DTG] = -1.0 / (dt * gm);
  ss[SS_INV_DT] = 1.0 / dt;
  ss[SS_NINV_DT] = -1.0 / dt;
  ss[SS_C12] = dt * (1.0 - 2.0 * gm);
  ss[SS_HALF] = dt / 2.0;
  d.tsd[0] = (unsigned long long)__double_as_longlong(-1.0);
  d.tsd[1] = (unsigned long long)__double_as_longlong(1.0e50);
}
__global__ void k_end_step(Dev d) { d.ss[SS_T] = d.ss[SS_T] + d.ss[SS_DT]; }

// postinit beta normalisation (puffy.zig:892-1024)
__global__ void k_beta_max(GridD g, Dev d, kz_physics ph, unsigned long long *out) {
  long t = TID;
  if (t >= (long)g.nx * g.ny * g.nz) return;
  long ix = t % g.nx, iy = (t / g.nx) % g.ny, iz = t / ((long)g.nx * g.ny);
  long c = cidx(g, ix, iy, iz);
  double pp[NV];
  ld(d.p, g.ncell, c, pp);
  Geo gm;
  geo_c(g, d, ix, iy, iz, gm);
  double ucon[4], ucov[4], bcon[4], bcov[4], bsq;
  ucon_ucov(pp, gm, ucon, ucov);
  bcon_bcov_bsq(pp[B1], pp[B2], pp[B3], ucon, ucov, gm, bcon, bcov, &bsq);
  double pmag = bsq / 2.0;
  double ptot = (ph.gam - 1.0) * pp[UU];
  double ehat = -calc_ff_rtt(pp, gm);
  ptot += ehat / 3.0;
  double b = pmag / ptot;
  if (b > 0.0) atomicMax((long long *)out, __double_as_longlong(b));
}
__global__ void k_beta_scale(GridD g, Dev d, kz_physics ph, double fac) {
  long t = TID;
  if (t >= (long)g.nx * g.ny * g.nz) return;
  long ix = t % g.nx, iy = (t / g.nx) % g.ny, iz = t / ((long)g.nx * g.ny);
  long n = g.ncell;
  long c = cidx(g, ix, iy, iz);
  double pp[NV], uu[NV];
  ld(d.p, n, c, pp);
  pp[B1] *= fac;
  pp[B2] *= fac;
  pp[B3] *= fac;
  Geo gm;
  geo_c(g, d, ix, iy, iz, gm);
  p2u(pp, uu, gm, ph.gam);
  st(d.p, n, c, pp);
  st(d.u, n, c, uu);
}
// initTimestepGuess (sim.zig:693-718)
__global__ void k_init_tsd(GridD g, Dev d, double tsd) {
  long t = TID;
  if (t >= (long)g.nx * g.ny * g.nz) return;
  long ix = t % g.nx, iy = (t / g.nx) % g.ny, iz = t / ((long)g.nx * g.ny);
  d.scal[SC_TSTEPDEN * g.ncell + cidx(g, ix, iy, iz)] = tsd;
  if (t == 0) {
    d.tsd[0] = (unsigned long long)__double_as_longlong(tsd);
    d.tsd[1] = (unsigned long long)__double_as_longlong(tsd);
  }
}

// AoS <-> SoA conversion (the C-ABI edge uses the reference Field layout)
__global__ void k_aos2soa(const double *aos, double *soa, long ncell, int nv) {
  long t = TID;
  if (t >= ncell * nv) return;
  long c = t / nv;
  int iv = t % nv;
  soa[iv * ncell + c] = aos[t];
}
__global__ void k_soa2aos(const double *soa, double *aos, long ncell, int nv) {
  long t = TID;
  if (t >= ncell * nv) return;
  long c = t / nv;
  int iv = t % nv;
  aos[t] = soa[iv * ncell + c];
}
__global__ void k_flags2aos(const int32_t *soa, int32_t *aos, long ncell) {
  long t = TID;
  if (t >= ncell * KZ_NFLAGS) return;
  long c = t / KZ_NFLAGS;
  int f = t % KZ_NFLAGS;
  aos[t] = soa[f * ncell + c];
}

// ===========================================================================
// host side
// ===========================================================================
struct kz_sim {
  GridD g;
  kz_physics ph;
  Dev d;
  hipStream_t stream;
  double min_dx, min_dy, min_dz;
  double t_host;
  uint64_t nstep;
  std::vector<void *> allocs;
  // event timing of one kernel class (kz_sim_time_kernel)
  int timed = -1;
  std::vector<hipEvent_t> ev_pool;  // created events, reused
  size_t ev_used = 0;               // events recorded since the last read
};

// Brackets the launches of one kernel class with events on the sim's stream
// when that class is selected for timing.
struct TimedLaunch {
  kz_sim *s;
  bool on;
  hipEvent_t *ev(kz_sim *x) {
    if (x->ev_used == x->ev_pool.size()) {
      hipEvent_t e;
      if (hipEventCreate(&e) != hipSuccess) return nullptr;
      x->ev_pool.push_back(e);
    }
    return &x->ev_pool[x->ev_used++];
  }
  TimedLaunch(kz_sim *x, int kind) : s(x), on(x->timed == kind) {
    if (on) {
      hipEvent_t *e = ev(s);
      on = e && hipEventRecord(*e, s->stream) == hipSuccess;
    }
  }
  ~TimedLaunch() {
    if (on) {
      hipEvent_t *e = ev(s);
      if (e) (void)hipEventRecord(*e, s->stream);
    }
  }
};

#define HIPCHK(x)                                                                   \
  do {                                                                              \
    hipError_t e_ = (x);                                                            \
    if (e_ != hipSuccess) {                                                         \
      fprintf(stderr, "kz: HIP error %s at %s:%d\n", hipGetErrorString(e_), __FILE__, __LINE__); \
      return KZ_E_HIP;                                                              \
    }                                                                               \
  } while (0)

static int dalloc(kz_sim *s, void **p, size_t bytes) {
  if (hipMalloc(p, bytes) != hipSuccess) return KZ_E_OOM;
  s->allocs.push_back(*p);
  return hipMemset(*p, 0, bytes) == hipSuccess ? 0 : KZ_E_HIP;
}

static inline unsigned nblk(long n, int bs) { return (unsigned)((n + bs - 1) / bs); }

// repack reference 20-blocks into the 35 SoA planes: g[16] G[16] gdet gttpert alpha
static void pack_geo(const double *g20, const double *G20, long n, std::vector<double> &out) {
  out.assign((size_t)n * GEO_NC, 0.0);
  for (long c = 0; c < n; ++c) {
    for (int a = 0; a < 4; ++a)
      for (int b = 0; b < 4; ++b) {
        out[(a * 4 + b) * n + c] = g20[c * 20 + a * 5 + b];
        out[(16 + a * 4 + b) * n + c] = G20[c * 20 + a * 5 + b];
      }
    out[32 * n + c] = g20[c * 20 + 19];
    out[33 * n + c] = G20[c * 20 + 19];
    out[34 * n + c] = __builtin_sqrt(-1.0 / G20[c * 20 + 0]);  // precompute.zig:494
  }
}

extern "C" {

kz_sim *kz_sim_create(const kz_grid *gr, const kz_physics *ph, const double *g, const double *gcon,
                      const double *kris, const double *gb0, const double *gconb0, const double *gb1,
                      const double *gconb1, const double *gb2, const double *gconb2, const double *bcx, int *err) {
  *err = 0;
  if (gr->nz != 1 || gr->ny <= 1 || gr->nx <= 1 || gr->ng < 3) { *err = KZ_E_UNSUPPORTED; return nullptr; }
  if (ph->radviscosity || ph->dynamo || ph->reduceorderatbh || ph->reduceorderafterfixup ||
      ph->opac.synchrotron_bridge || ph->floors.b2rhofloorframe != 0) { *err = KZ_E_UNSUPPORTED; return nullptr; }
  if (ph->correct_polaraxis && gr->ny <= 2 * ph->nccorrectpolar) { *err = KZ_E_INVALID_CONFIG; return nullptr; }
  kz_sim *s = new kz_sim();
  s->ph = *ph;
  GridD &G = s->g;
  G.nx = gr->nx; G.ny = gr->ny; G.nz = gr->nz; G.ng = gr->ng;
  G.ngx = gr->nx > 1 ? gr->ng : 0;
  G.ngy = gr->ny > 1 ? gr->ng : 0;
  G.ngz = gr->nz > 1 ? gr->ng : 0;
  G.sx = gr->nx + 2 * G.ngx; G.sy = gr->ny + 2 * G.ngy; G.sz = gr->nz + 2 * G.ngz;
  G.ncell = (long)G.sx * G.sy * G.sz;
  G.minx = gr->minx; G.miny = gr->miny; G.minz = gr->minz;
  G.dx = (gr->maxx - gr->minx) / (double)gr->nx;
  G.dy = (gr->maxy - gr->miny) / (double)gr->ny;
  G.dz = (gr->maxz - gr->minz) / (double)gr->nz;
  for (int d = 0; d < 3; ++d) {
    G.fsx[d] = G.sx + (d == 0);
    G.fsy[d] = G.sy + (d == 1);
    G.nface[d] = (long)G.fsx[d] * G.fsy[d] * (G.sz + (d == 2));
  }
  if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess) { *err = KZ_E_HIP; delete s; return nullptr; }
  size_t fb = (size_t)G.ncell * NV * sizeof(double);
  Dev &D = s->d;
  double **fl[] = {&D.p, &D.u, &D.ut0, &D.ut1, &D.ut2, &D.dut1, &D.dut2, &D.drt1, &D.drt2, &D.uforget, &D.stage};
  int e = 0;
  for (auto f : fl) e |= dalloc(s, (void **)f, fb);
  for (int d = 0; d < 3; ++d) e |= dalloc(s, (void **)&D.flb[d], (size_t)G.nface[d] * NV * sizeof(double));
  e |= dalloc(s, (void **)&D.scal, (size_t)G.ncell * SC_N * sizeof(double));
  e |= dalloc(s, (void **)&D.flags, (size_t)G.ncell * KZ_NFLAGS * sizeof(int32_t));
  long ncorn = (long)(gr->nx + 1) * (gr->ny + 1) * (gr->nz + 1);
  for (int c = 0; c < 3; ++c) e |= dalloc(s, (void **)&D.emf[c], ncorn * sizeof(double));
  e |= dalloc(s, (void **)&D.gc, (size_t)G.ncell * GEO_NC * sizeof(double));
  for (int d = 0; d < 3; ++d) e |= dalloc(s, (void **)&D.gf[d], (size_t)G.nface[d] * GEO_NC * sizeof(double));
  e |= dalloc(s, (void **)&D.kris, (size_t)G.ncell * 64 * sizeof(double));
  size_t nbcx = (size_t)gr->nz * gr->ny * gr->ng * BCX_REC;
  e |= dalloc(s, (void **)&D.bcx, nbcx * sizeof(double));
  e |= dalloc(s, (void **)&D.cnt, 8 * sizeof(unsigned long long));
  e |= dalloc(s, (void **)&D.tsd, 2 * sizeof(unsigned long long));
  e |= dalloc(s, (void **)&D.err, sizeof(int));
  e |= dalloc(s, (void **)&D.ss, SS_N * sizeof(double));
  long nd = (long)gr->nx * gr->ny * gr->nz;
  e |= dalloc(s, (void **)&D.imp_uu00, (size_t)nd * NV * sizeof(double));
  e |= dalloc(s, (void **)&D.imp_pp0, (size_t)nd * NV * sizeof(double));
  e |= dalloc(s, (void **)&D.imp_aux, (size_t)nd * IMP_NAUX * sizeof(double));
  e |= dalloc(s, (void **)&D.imp_slow, (size_t)nd * sizeof(int32_t));
  e |= dalloc(s, (void **)&D.imp_nslow, sizeof(unsigned long long));
  if (e) { *err = e < 0 ? e : KZ_E_HIP; kz_sim_destroy(s); return nullptr; }
  // metric tables
  std::vector<double> buf;
  pack_geo(g, gcon, G.ncell, buf);
  e |= hipMemcpy(D.gc, buf.data(), buf.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess;
  const double *gbs[3] = {gb0, gb1, gb2}, *gcbs[3] = {gconb0, gconb1, gconb2};
  for (int d = 0; d < 3; ++d) {
    pack_geo(gbs[d], gcbs[d], G.nface[d], buf);
    e |= hipMemcpy(D.gf[d], buf.data(), buf.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess;
  }
  buf.assign((size_t)G.ncell * 64, 0.0);
  for (long c = 0; c < G.ncell; ++c)
    for (int k = 0; k < 64; ++k) buf[(size_t)k * G.ncell + c] = kris[c * 64 + k];
  e |= hipMemcpy(D.kris, buf.data(), buf.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess;
  e |= hipMemcpy(D.bcx, bcx, nbcx * sizeof(double), hipMemcpyHostToDevice) != hipSuccess;
  if (e) { *err = KZ_E_HIP; kz_sim_destroy(s); return nullptr; }
  // set_grid minimum cell sizes incl. the mdx quirk (sim.zig:447-466)
  double mdx = -1, mdy = -1, mdz = -1;
  for (long ix = 0; ix < gr->nx; ++ix)
    for (long iy = 0; iy < gr->ny; ++iy)
      for (long iz = 0; iz < gr->nz; ++iz) {
        double dxv = (G.minx + (double)(ix + 1) * G.dx) - (G.minx + (double)ix * G.dx);
        double dyv = (G.miny + (double)(iy + 1) * G.dy) - (G.miny + (double)iy * G.dy);
        double dzv = (G.minz + (double)(iz + 1) * G.dz) - (G.minz + (double)iz * G.dz);
        if (dxv < mdx || mdx < 0) mdx = dxv;
        if (dyv < mdx || mdy < 0) mdy = dyv;
        if (dzv < mdx || mdz < 0) mdz = dzv;
      }
  s->min_dx = mdx; s->min_dy = mdy; s->min_dz = mdz;
  s->t_host = 0;
  s->nstep = 0;
  unsigned long long tsd0[2] = {(unsigned long long)__builtin_bit_cast(long long, 0.0),
                                (unsigned long long)__builtin_bit_cast(long long, 1.0e50)};
  if (hipMemcpy(D.tsd, tsd0, sizeof tsd0, hipMemcpyHostToDevice) != hipSuccess) {
    *err = KZ_E_HIP;
    kz_sim_destroy(s);
    return nullptr;
  }
  return s;
}

void kz_sim_destroy(kz_sim *s) {
  if (!s) return;
  (void)hipStreamSynchronize(s->stream);
  for (void *p : s->allocs) (void)hipFree(p);
  for (hipEvent_t e : s->ev_pool) (void)hipEventDestroy(e);
  (void)hipStreamDestroy(s->stream);
  delete s;
}

// Wait for the sim's stream and return (and clear) the first error a
// launch or a kernel reported.
static int check_err(kz_sim *s) {
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(s->stream));
  int e = 0;
  HIPCHK(hipMemcpy(&e, s->d.err, sizeof(int), hipMemcpyDeviceToHost));
  if (e) {
    int z = 0;
    HIPCHK(hipMemcpy(s->d.err, &z, sizeof(int), hipMemcpyHostToDevice));
  }
  return e;
}

int kz_sim_set_state(kz_sim *s, const double *p_aos, const double *u_aos) {
  size_t fb = (size_t)s->g.ncell * NV * sizeof(double);
  long tot = s->g.ncell * NV;
  if (p_aos) {
    HIPCHK(hipMemcpyAsync(s->d.stage, p_aos, fb, hipMemcpyHostToDevice, s->stream));
    k_aos2soa<<<nblk(tot, 256), 256, 0, s->stream>>>(s->d.stage, s->d.p, s->g.ncell, NV);
  }
  if (u_aos) {
    HIPCHK(hipMemcpyAsync(s->d.stage, u_aos, fb, hipMemcpyHostToDevice, s->stream));
    k_aos2soa<<<nblk(tot, 256), 256, 0, s->stream>>>(s->d.stage, s->d.u, s->g.ncell, NV);
  }
  HIPCHK(hipStreamSynchronize(s->stream));
  return 0;
}

int kz_sim_get_state(kz_sim *s, double *p_aos, double *u_aos, int32_t *flags) {
  size_t fb = (size_t)s->g.ncell * NV * sizeof(double);
  long tot = s->g.ncell * NV;
  if (p_aos) {
    k_soa2aos<<<nblk(tot, 256), 256, 0, s->stream>>>(s->d.p, s->d.stage, s->g.ncell, NV);
    HIPCHK(hipMemcpyAsync(p_aos, s->d.stage, fb, hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
  }
  if (u_aos) {
    k_soa2aos<<<nblk(tot, 256), 256, 0, s->stream>>>(s->d.u, s->d.stage, s->g.ncell, NV);
    HIPCHK(hipMemcpyAsync(u_aos, s->d.stage, fb, hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
  }
  if (flags) {
    long nf = s->g.ncell * KZ_NFLAGS;
    k_flags2aos<<<nblk(nf, 256), 256, 0, s->stream>>>(s->d.flags, (int32_t *)s->d.stage, s->g.ncell);
    HIPCHK(hipMemcpyAsync(flags, s->d.stage, nf * sizeof(int32_t), hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
  }
  return 0;
}

static long ndom(const kz_sim *s) { return (long)s->g.nx * s->g.ny * s->g.nz; }

int kz_sim_p2u_domain(kz_sim *s) {
  k_p2u_domain<<<nblk(ndom(s), 256), 256, 0, s->stream>>>(s->g, s->d, s->ph.gam);
  return check_err(s);
}

static void launch_set_bc(kz_sim *s) {
  const GridD &g = s->g;
  TimedLaunch tl(s, KZ_K_BC);
  long nb = (long)g.nz * g.ny * g.ngx * 2 + (long)g.nz * g.nx * g.ngy * 2;
  k_bc_xy<<<nblk(nb, 64), 64, 0, s->stream>>>(g, s->d, s->ph);
  k_bc_corners2d<<<1, 64, 0, s->stream>>>(g, s->d, s->ph);
}

int kz_sim_set_bc(kz_sim *s, int ifinit) {
  (void)ifinit;
  launch_set_bc(s);
  return check_err(s);
}

static void launch_calc_u2p(kz_sim *s) {
  long n = ndom(s);
  {
    TimedLaunch tl(s, KZ_K_U2P);
    k_u2p<<<nblk(n, 128), 128, 0, s->stream>>>(s->g, s->d, s->ph);
  }
  if (s->ph.do_fixups) k_fixup<<<nblk(n, 256), 256, 0, s->stream>>>(s->g, s->d, s->ph, KZ_FLAG_HD_FIXUP);
  if (s->ph.do_fixups && s->ph.do_u2prad_fixups)
    k_fixup<<<nblk(n, 256), 256, 0, s->stream>>>(s->g, s->d, s->ph, KZ_FLAG_RAD_FIXUP);
  launch_set_bc(s);
}

static void launch_op_explicit(kz_sim *s, int dt_slot) {
  const GridD &g = s->g;
  long nw = (long)(g.nx + 2) * (g.ny + 2) * g.nz;
  {
    TimedLaunch tl(s, KZ_K_WAVESPEEDS);
    k_wavespeeds<<<nblk(nw, 128), 128, 0, s->stream>>>(g, s->d, s->ph);
  }
  for (int dim = 0; dim < 2; ++dim) {
    long nd = dim == 0 ? g.nx : g.ny;
    long w0 = dim == 0 ? g.ny + 2 : g.nx + 2;
    long nf = (nd + 1) * w0 * g.nz;
    TimedLaunch tl(s, KZ_K_FACE_FLUX);
    k_face_flux<<<nblk(nf, 128), 128, 0, s->stream>>>(g, s->d, s->ph, dim);
  }
  long nc = (long)(g.nx + 1) * (g.ny + 1) * (g.nz + 1);
  {
    TimedLaunch tl(s, KZ_K_CT);
    k_emf<<<nblk(nc, 256), 256, 0, s->stream>>>(g, s->d);
    k_ct_rebuild<<<nblk(nc, 256), 256, 0, s->stream>>>(g, s->d);
  }
  {
    TimedLaunch tl(s, KZ_K_UPDATE);
    k_update<<<nblk(ndom(s), 128), 128, 0, s->stream>>>(g, s->d, s->ph, dt_slot);
  }
  launch_calc_u2p(s);
}

static void launch_op_implicit(kz_sim *s, int dt_slot) {
  if (!s->ph.has_opac) return;
  {
    TimedLaunch tl(s, KZ_K_IMPLICIT);
    unsigned nb = nblk(ndom(s), 64);
    k_imp_prep<<<nb, 64, 0, s->stream>>>(s->g, s->d, s->ph, dt_slot);
    k_imp_rung0<WP_MHD><<<nb, 64, 0, s->stream>>>(s->g, s->d, s->ph, dt_slot);
    k_imp_rung0<WP_RAD><<<nb, 64, 0, s->stream>>>(s->g, s->d, s->ph, dt_slot);
    k_imp_slow<<<nb, 64, 0, s->stream>>>(s->g, s->d, s->ph, dt_slot);
  }
  if (s->ph.do_fixups && s->ph.do_radimp_fixups)
    k_fixup<<<nblk(ndom(s), 256), 256, 0, s->stream>>>(s->g, s->d, s->ph, KZ_FLAG_RADIMP_FIXUP);
}

static void launch_do_correct(kz_sim *s) {
  if (!s->ph.correct_polaraxis) return;
  long n = (long)s->g.nx * s->ph.nccorrectpolar * 2 * s->g.nz;
  k_polar<<<nblk(n, 128), 128, 0, s->stream>>>(s->g, s->d, s->ph);
}

int kz_calc_u2p(kz_sim *s) {
  launch_calc_u2p(s);
  return check_err(s);
}

// a caller-supplied dt goes into a spare step-scalar slot
static int put_dt(kz_sim *s, double dt) {
  HIPCHK(hipMemcpyAsync(s->d.ss + SS_USER, &dt, sizeof(double), hipMemcpyHostToDevice, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  return 0;
}

int kz_op_explicit(kz_sim *s, double dt) {
  if (int e = put_dt(s, dt)) return e;
  launch_op_explicit(s, SS_USER);
  return check_err(s);
}

int kz_op_implicit(kz_sim *s, double dt) {
  if (int e = put_dt(s, dt)) return e;
  launch_op_implicit(s, SS_USER);
  return check_err(s);
}

int kz_do_correct(kz_sim *s) {
  launch_do_correct(s);
  return check_err(s);
}

// Enqueue one RK2IMEX step (sim.zig:1792-1899) on the sim's stream; no host
// synchronisation inside.
static void launch_step(kz_sim *s, double forced, int use_forced) {
  const GridD &g = s->g;
  Dev &d = s->d;
  hipStream_t st = s->stream;
  long n = ndom(s);
  long nn = n * NV;
  size_t fb = (size_t)g.ncell * NV * sizeof(double);
  k_begin_step<<<1, 1, 0, st>>>(d, forced, use_forced);
  // saveTimesteps writes cell_dt, which nothing on this path reads.
  (void)hipMemcpyAsync(d.ut0, d.u, fb, hipMemcpyDeviceToDevice, st);
  launch_op_implicit(s, SS_DTG);
  k_stage_deriv<<<nblk(nn, 256), 256, 0, st>>>(g, d.drt1, d.u, d.ut0, d.ss, SS_INV_DTG, SS_NINV_DTG);
  (void)hipMemcpyAsync(d.ut1, d.u, fb, hipMemcpyDeviceToDevice, st);
  launch_calc_u2p(s);
  launch_do_correct(s);
  launch_set_bc(s);
  launch_op_explicit(s, SS_DT);
  k_copy_entropy_flag<<<nblk(n, 256), 256, 0, st>>>(g, d);
  k_stage_deriv<<<nblk(nn, 256), 256, 0, st>>>(g, d.dut1, d.u, d.ut1, d.ss, SS_INV_DT, SS_NINV_DT);
  // u = ut0 + dt*dut1 + dt(1-2g)*drt1
  k_stage_combine<<<nblk(nn, 256), 256, 0, st>>>(g, d.u, d.ut0, d.dut1, d.drt1, d.ss, SS_DT, SS_C12);
  (void)hipMemcpyAsync(d.uforget, d.u, fb, hipMemcpyDeviceToDevice, st);
  launch_calc_u2p(s);
  launch_do_correct(s);
  launch_op_implicit(s, SS_DTG);
  k_stage_deriv<<<nblk(nn, 256), 256, 0, st>>>(g, d.drt2, d.u, d.uforget, d.ss, SS_INV_DTG, SS_NINV_DTG);
  (void)hipMemcpyAsync(d.ut2, d.u, fb, hipMemcpyDeviceToDevice, st);
  launch_calc_u2p(s);
  launch_do_correct(s);
  launch_set_bc(s);
  launch_op_explicit(s, SS_DT);
  k_stage_deriv<<<nblk(nn, 256), 256, 0, st>>>(g, d.dut2, d.u, d.ut2, d.ss, SS_INV_DT, SS_NINV_DT);
  k_stage_combine<<<nblk(nn, 256), 256, 0, st>>>(g, d.u, d.ut0, d.dut1, d.dut2, d.ss, SS_HALF, SS_HALF);
  k_stage_combine<<<nblk(nn, 256), 256, 0, st>>>(g, d.u, d.u, d.drt1, d.drt2, d.ss, SS_HALF, SS_HALF);
  launch_calc_u2p(s);
  launch_do_correct(s);
  launch_set_bc(s);
  k_end_step<<<1, 1, 0, st>>>(d);
  k_update_entropy<<<nblk(n, 256), 256, 0, st>>>(g, d, s->ph);
}

int kz_step(kz_sim *s, const double *forced_dt) {
  launch_step(s, forced_dt ? *forced_dt : 0.0, forced_dt != nullptr);
  s->nstep += 1;
  return check_err(s);
}

// enqueue `nsteps` CFL steps without host synchronisation (benchmarks)
int kz_step_many(kz_sim *s, int nsteps) {
  for (int i = 0; i < nsteps; ++i) {
    launch_step(s, 0.0, 0);
    s->nstep += 1;
  }
  return 0;
}

int kz_sim_sync(kz_sim *s) { return check_err(s); }

void *kz_sim_stream(kz_sim *s) { return (void *)s->stream; }

int kz_sim_postinit(kz_sim *s, double maxbeta, double *fac_out) {
  unsigned long long *dmax = s->d.cnt + 7;
  HIPCHK(hipMemsetAsync(dmax, 0, sizeof(unsigned long long), s->stream));
  k_beta_max<<<nblk(ndom(s), 256), 256, 0, s->stream>>>(s->g, s->d, s->ph, dmax);
  unsigned long long bits = 0;
  HIPCHK(hipMemcpyAsync(&bits, dmax, sizeof bits, hipMemcpyDeviceToHost, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  double maxb = __builtin_bit_cast(double, (long long)bits);  // >= 0 (starts at +0)
  double fac = __builtin_sqrt(maxbeta / maxb);
  k_beta_scale<<<nblk(ndom(s), 256), 256, 0, s->stream>>>(s->g, s->d, s->ph, fac);
  if (fac_out) *fac_out = fac;
  return check_err(s);
}

int kz_sim_init_timestep_guess(kz_sim *s) {
  const double ws = 10000.0;
  double tsd;
  if (s->g.nz > 1) tsd = ws / s->min_dx + ws / s->min_dy + ws / s->min_dz;
  else if (s->g.ny > 1) tsd = ws / s->min_dx + ws / s->min_dy;
  else tsd = ws / s->min_dx;
  tsd /= s->ph.tsteplim;
  k_init_tsd<<<nblk(ndom(s), 256), 256, 0, s->stream>>>(s->g, s->d, tsd);
  return check_err(s);
}

int kz_sim_set_time(kz_sim *s, double t, double tstepdenmax, double tstepdenmin) {
  double tt = t;
  HIPCHK(hipMemcpy(s->d.ss + SS_T, &tt, sizeof tt, hipMemcpyHostToDevice));
  unsigned long long b[2] = {(unsigned long long)__builtin_bit_cast(long long, tstepdenmax),
                             (unsigned long long)__builtin_bit_cast(long long, tstepdenmin)};
  HIPCHK(hipMemcpy(s->d.tsd, b, sizeof b, hipMemcpyHostToDevice));
  return 0;
}

int kz_sim_counters(kz_sim *s, kz_counters *c) {
  HIPCHK(hipStreamSynchronize(s->stream));
  unsigned long long cnt[8], tsd[2];
  double ss[SS_N];
  HIPCHK(hipMemcpy(cnt, s->d.cnt, sizeof cnt, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(tsd, s->d.tsd, sizeof tsd, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(ss, s->d.ss, sizeof ss, hipMemcpyDeviceToHost));
  c->t = ss[SS_T];
  c->dt = ss[SS_DT];
  c->own_dt = ss[SS_OWN_DT];
  c->tstepdenmax = __builtin_bit_cast(double, (long long)tsd[0]);
  c->tstepdenmin = __builtin_bit_cast(double, (long long)tsd[1]);
  c->nstep = s->nstep;
  c->n_radimp_failures = cnt[0];
  c->n_radimp_iters = cnt[1];
  c->n_radimp_solves = cnt[2];
  c->n_guard_recoveries = cnt[3];
  return 0;
}

int kz_sim_time_kernel(kz_sim *s, int kernel) {
  HIPCHK(hipStreamSynchronize(s->stream));
  if (kernel < -1 || kernel >= KZ_K_COUNT) return KZ_E_INVALID_CONFIG;
  s->timed = kernel;
  s->ev_used = 0;
  return 0;
}

int kz_sim_kernel_time(kz_sim *s, double *ms_total, uint64_t *launches) {
  HIPCHK(hipStreamSynchronize(s->stream));
  double tot = 0;
  uint64_t n = 0;
  for (size_t i = 0; i + 1 < s->ev_used; i += 2) {
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, s->ev_pool[i], s->ev_pool[i + 1]));
    tot += ms;
    ++n;
  }
  s->ev_used = 0;
  *ms_total = tot;
  *launches = n;
  return 0;
}

}  // extern "C"
