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
#include "kz_sim_dev.h"
#include "kz_sim_kernels.h"

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
// puffy.zig:1089-1166 (xhi) / :1167-1171 (xlo): the primitives of x ghost
// (ix, iy, iz) from the domain edge cell of its row; geom is the ghost's
// centre geometry.  Returns an error code, E_OK on success.
__device__ int bc_x_prims(const GridD &g, const Dev &d, long ix, long iy, long iz, const Geo &geom, double *pp) {
  const long nx = g.nx;
  if (ix < 0) {
    ld(d.p, g.ncell, cidx(g, 0, iy, iz), pp);
    return E_OK;
  }
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
    if (e) return e;
    pp[VX] = uc[1]; pp[VY] = uc[2]; pp[VZ] = uc[3];
  }
  double ur[4] = {0.0, pp[FX], pp[FY], pp[FZ]};
  velr_to_vel4(ur, geom, ur);
  tvec(J12, ur, ur);
  if (ur[1] < 0.0) {
    ur[1] = 0.0;
    tvec(J21, ur, ur);
    int e = vel4_to_velr(ur, geom, false, ur);
    if (e) return e;
    pp[FX] = ur[1]; pp[FY] = ur[2]; pp[FZ] = ur[3];
  }
  return E_OK;
}
// puffy.zig:1172-1187 polar reflection: the primitives of y ghost (ix, iy, iz)
__device__ void bc_y_prims(const GridD &g, const Dev &d, long ix, long iy, long iz, double *pp) {
  const long ny = g.ny;
  const long iiy = iy >= ny ? ny - (iy - ny) - 1 : -iy - 1;
  ld(d.p, g.ncell, cidx(g, ix, iiy, iz), pp);
  pp[VY] = -pp[VY];
  pp[B2] = -pp[B2];
  pp[FY] = -pp[FY];
}
__device__ void store_p2u(const GridD &g, const Dev &d, double gam, long ix, long iy, long iz, const Geo &geom,
                          double *pp) {
  double uu[NV];
  p2u(pp, uu, geom, gam);
  const long c = cidx(g, ix, iy, iz);
  st(d.p, g.ncell, c, pp);
  st(d.u, g.ncell, c, uu);
}
// x ghost t < nz*ny*ngx*2; a ghost whose transform fails keeps its old state
__device__ void bc_x_cell(const GridD &g, const Dev &d, const kz_physics &ph, long t) {
  int side = t % 2;
  long i = (t / 2) % g.ngx + 1;
  long iy = (t / (2 * g.ngx)) % g.ny;
  long iz = t / (2L * g.ngx * g.ny);
  long ix = side ? g.nx - 1 + i : -i;
  double pp[NV];
  Geo geom;
  geo_c(g, d, ix, iy, iz, geom);
  int e = bc_x_prims(g, d, ix, iy, iz, geom, pp);
  if (e) {
    set_err(d.err, e);
    return;
  }
  store_p2u(g, d, ph.gam, ix, iy, iz, geom, pp);
}
// y ghost t < nz*nx*ngy*2
__device__ void bc_y_cell(const GridD &g, const Dev &d, const kz_physics &ph, long t) {
  int side = t % 2;
  long i = (t / 2) % g.ngy + 1;
  long ix = (t / (2 * g.ngy)) % g.nx;
  long iz = t / (2L * g.ngy * g.nx);
  long iy = side ? g.ny - 1 + i : -i;
  double pp[NV];
  bc_y_prims(g, d, ix, iy, iz, pp);
  Geo geom;
  geo_c(g, d, ix, iy, iz, geom);
  store_p2u(g, d, ph.gam, ix, iy, iz, geom, pp);
}
// bc.zig:286-377 fillCorners2d.  Per corner quadrant the reference copies
// 2(ng-1) cells from the x and y ghosts next to the corner (the x ghosts of
// the edge row into the first ghost row, the y ghosts of the edge column
// into the first ghost column), averages the inner diagonal cell from the x
// ghost and the y ghost beside it, then the outer diagonal cell from two of
// the copies.  Every value involved is a ghost value of this setBc, which is
// a function of domain cells only, so each corner cell is formed by its own
// thread from those ghost functions, in the same launch as the ghosts: the
// same primitives, p2u on the corner cell's geometry.  A ghost whose
// transform fails is not written by this setBc; its old state is used, as
// the reference's copy reads it.  t < 4 (2 (ng-1) + 2).
__device__ void corner_x_src(const GridD &g, const Dev &d, long x, long yin, double *pp) {
  Geo gs;
  geo_c(g, d, x, yin, 0, gs);
  if (bc_x_prims(g, d, x, yin, 0, gs, pp)) ld(d.p, g.ncell, cidx(g, x, yin, 0), pp);
}
__device__ void bc_corner2d_cell(const GridD &g, const Dev &d, const kz_physics &ph, long t) {
  const long nx = g.nx, ny = g.ny, ng = g.ng;
  const int per = 2 * (int)(ng - 1) + 2;  // copies, inner and outer diagonal
  const int q = (int)(t / per), k = (int)(t % per);
  // quadrant q: bit 0 = high y side, bit 1 = high x side
  const bool hx = q & 2, hy = q & 1;
  const long xin = hx ? nx - 1 : 0, yin = hy ? ny - 1 : 0;  // edge domain index
  const long xg1 = hx ? nx : -1, yg1 = hy ? ny : -1;        // first ghost index
  const long xg2 = hx ? nx + 1 : -2, yg2 = hy ? ny + 1 : -2;
  // the corner cell: copies first (k < 2 (ng - 1)), then the inner and the
  // outer diagonal
  const bool copy = k < 2 * (ng - 1), inner = k == 2 * (ng - 1);
  const long i = k / 2;
  const long dx_ = copy ? (k % 2 == 0 ? (hx ? nx + i + 1 : -ng + i) : xg1) : (inner ? xg1 : xg2);
  const long dy_ = copy ? (k % 2 == 0 ? yg1 : (hy ? ny + i + 1 : -ng + i)) : (inner ? yg1 : yg2);
  Geo gm;  // the corner cell's geometry, loaded before the ghost values are formed
  geo_c(g, d, dx_, dy_, 0, gm);
  double pp[NV];
  if (copy && k % 2 == 0) {  // (x, yg1) <- x ghost (x, yin)
    corner_x_src(g, d, dx_, yin, pp);
  } else if (copy) {  // (xg1, y) <- y ghost (xin, y)
    bc_y_prims(g, d, xin, dy_, 0, pp);
  } else {
    // inner: x ghost (xg1, yin) and y ghost (xin, yg1); outer: the copies
    // (xg2, yg1) of x ghost (xg2, yin) and (xg1, yg2) of y ghost (xin, yg2)
    double pa[NV], pb[NV];
    corner_x_src(g, d, dx_, yin, pa);
    bc_y_prims(g, d, xin, dy_, 0, pb);
#pragma unroll
    for (int iv = 0; iv < NV; ++iv) pp[iv] = 0.5 * (pa[iv] + pb[iv]);
  }
  store_p2u(g, d, ph.gam, dx_, dy_, 0, gm, pp);
}

// z faces, periodic in phi (puffy.zig:237, bc.zig:187-252 .periodic, with its
// nz < ng pin); t < nx*ny*ngz*2
__device__ void bc_z_cell(const GridD &g, const Dev &d, const kz_physics &ph, long t) {
  const int side = t % 2;
  const long i = (t / 2) % g.ngz + 1;
  const long ix = (t / (2 * g.ngz)) % g.nx;
  const long iy = t / (2L * g.ngz * g.nx);
  const long nz = g.nz;
  const long iz = side ? nz - 1 + i : -i;
  long iiz = iz;
  if (iz < 0) iiz = iz + nz;
  if (iz > nz - 1) iiz = iz - nz;
  if (nz < g.ng) iiz = 0;
  double pp[NV], uu[NV];
  ld(d.p, g.ncell, cidx(g, ix, iy, iiz), pp);
  Geo geom;
  geo_c(g, d, ix, iy, iz, geom);
  p2u(pp, uu, geom, ph.gam);
  const long c = cidx(g, ix, iy, iz);
  st(d.p, g.ncell, c, pp);
  st(d.u, g.ncell, c, uu);
}
// bc.zig:378-520 fillCorners3d.  The reference fills the 12 edge tubes in
// turn (along z, y, x) over their full through-axis span and then the 8
// corner cubes, which overwrite every tube cell that lies in a z (or y, x)
// ghost slice of another tube; those cells are the only ones whose tube
// value reads a not-yet-refreshed ghost.  The surviving tube cells are the
// ones at domain slices of the through-axis, which read only face ghosts and
// their own tube's copies, so the device fills those in two phases (copies
// and inner diagonals, then outer diagonals) and the cubes alongside the
// second phase; the result equals the reference's sequence.
__device__ __forceinline__ void c3set(long o[3], int ax, long av, int bx, long bv, int cx, long cv) {
  o[ax] = av;
  o[bx] = bv;
  o[cx] = cv;
}
__device__ void cp_p2u3(const GridD &g, const Dev &d, double gam, const long dst[3], const long src[3]) {
  double pp[NV], uu[NV];
  ld(d.p, g.ncell, cidx(g, src[0], src[1], src[2]), pp);
  Geo gm;
  geo_c(g, d, dst[0], dst[1], dst[2], gm);
  p2u(pp, uu, gm, gam);
  const long c = cidx(g, dst[0], dst[1], dst[2]);
  st(d.p, g.ncell, c, pp);
  st(d.u, g.ncell, c, uu);
}
__device__ void avg_p2u3(const GridD &g, const Dev &d, double gam, const long dst[3], const long a[3],
                         const long b[3]) {
  double pa[NV], pb[NV], pp[NV], uu[NV];
  ld(d.p, g.ncell, cidx(g, a[0], a[1], a[2]), pa);
  ld(d.p, g.ncell, cidx(g, b[0], b[1], b[2]), pb);
#pragma unroll
  for (int iv = 0; iv < NV; ++iv) pp[iv] = 0.5 * (pa[iv] + pb[iv]);
  Geo gm;
  geo_c(g, d, dst[0], dst[1], dst[2], gm);
  p2u(pp, uu, gm, gam);
  const long c = cidx(g, dst[0], dst[1], dst[2]);
  st(d.p, g.ncell, c, pp);
  st(d.u, g.ncell, c, uu);
}
// One (tube, corner, domain slice) of the 12 edge tubes: type 0 = x-y corners
// along z, 1 = x-z along y, 2 = y-z along x; q = (a_hi, b_hi).
struct EdgeSlice {
  int pa, pb, pc;
  long na, nb, c;
  bool a_hi, b_hi;
};
__device__ __forceinline__ bool edge_slice(const GridD &g, long s, EdgeSlice &e) {
  const long dims[3] = {g.nx, g.ny, g.nz};
  const long n0 = 4L * g.nz, n1 = 4L * g.ny, n2 = 4L * g.nx;
  int type;
  if (s < n0) type = 0;
  else if (s < n0 + n1) { type = 1; s -= n0; }
  else if (s < n0 + n1 + n2) { type = 2; s -= n0 + n1; }
  else return false;
  const int q = (int)(s % 4);
  e.c = s / 4;
  e.pa = type == 2 ? 1 : 0;
  e.pb = type == 0 ? 1 : 2;
  e.pc = type == 0 ? 2 : (type == 1 ? 1 : 0);
  e.na = dims[e.pa];
  e.nb = dims[e.pb];
  e.a_hi = q >> 1;
  e.b_hi = q & 1;
  return true;
}
__global__ void __launch_bounds__(64) k_bc_edges1(GridD g, Dev d, kz_physics ph) {
  const long t = TID;
  const long ng = g.ng;
  const int per = 2 * (int)(ng - 1) + 1;
  EdgeSlice e;
  if (!edge_slice(g, t / per, e)) return;
  const int k = (int)(t % per);
  const long a_edge = e.a_hi ? e.na - 1 : 0, a_g1 = e.a_hi ? e.na : -1;
  const long b_edge = e.b_hi ? e.nb - 1 : 0, b_g1 = e.b_hi ? e.nb : -1;
  long dst[3], x[3], y[3];
  if (k < 2 * (ng - 1)) {
    const long i = k / 2;
    if (k % 2 == 0) {  // (a_out, b_g1) <- (a_out, b_edge)
      const long a_out = e.a_hi ? e.na + i + 1 : -ng + i;
      c3set(dst, e.pa, a_out, e.pb, b_g1, e.pc, e.c);
      c3set(x, e.pa, a_out, e.pb, b_edge, e.pc, e.c);
    } else {  // (a_g1, b_out) <- (a_edge, b_out)
      const long b_out = e.b_hi ? e.nb + i + 1 : -ng + i;
      c3set(dst, e.pa, a_g1, e.pb, b_out, e.pc, e.c);
      c3set(x, e.pa, a_edge, e.pb, b_out, e.pc, e.c);
    }
    cp_p2u3(g, d, ph.gam, dst, x);
  } else {  // inner diagonal from the two face ghosts
    c3set(dst, e.pa, a_g1, e.pb, b_g1, e.pc, e.c);
    c3set(x, e.pa, a_g1, e.pb, b_edge, e.pc, e.c);
    c3set(y, e.pa, a_edge, e.pb, b_g1, e.pc, e.c);
    avg_p2u3(g, d, ph.gam, dst, x, y);
  }
}
__global__ void __launch_bounds__(64) k_bc_edges2(GridD g, Dev d, kz_physics ph) {
  const long t = TID;
  const long ng = g.ng;
  const long nslice = 4L * (g.nx + g.ny + g.nz);
  if (t < nslice) {  // outer diagonal from two of the slice's copies
    EdgeSlice e;
    edge_slice(g, t, e);
    const long a_g1 = e.a_hi ? e.na : -1, a_g2 = e.a_hi ? e.na + 1 : -2;
    const long b_g1 = e.b_hi ? e.nb : -1, b_g2 = e.b_hi ? e.nb + 1 : -2;
    long dst[3], x[3], y[3];
    c3set(dst, e.pa, a_g2, e.pb, b_g2, e.pc, e.c);
    c3set(x, e.pa, a_g2, e.pb, b_g1, e.pc, e.c);
    c3set(y, e.pa, a_g1, e.pb, b_g2, e.pc, e.c);
    avg_p2u3(g, d, ph.gam, dst, x, y);
    return;
  }
  const long u = t - nslice, ng3 = ng * ng * ng;
  if (u >= 8 * ng3) return;
  const int cube = (int)(u / ng3);  // bit 0 x_hi, bit 1 y_hi, bit 2 z_hi
  const long r = u % ng3;
  const bool x_hi = cube & 1, y_hi = (cube >> 1) & 1, z_hi = (cube >> 2) & 1;
  const long dst[3] = {(x_hi ? g.nx : -ng) + r % ng, (y_hi ? g.ny : -ng) + (r / ng) % ng,
                       (z_hi ? g.nz : -ng) + r / (ng * ng)};
  const long src[3] = {x_hi ? g.nx - 1 : 0, y_hi ? g.ny - 1 : 0, z_hi ? g.nz - 1 : 0};
  cp_p2u3(g, d, ph.gam, dst, src);
}

// setBc (bc.zig:40-73): x ghosts (reading only domain cells of rows
// 0..ny-1), y ghosts and z ghosts (reading only domain cells) are
// independent, so one launch does all three, and in 2D the corner fill as
// well (bc_corner2d_cell); in 3D the corner fill, which reads them, follows
// (k_bc_edges1/2).
__device__ __forceinline__ long bc_corner_count(const GridD &g) { return g.nz == 1 ? 4L * (2 * (g.ng - 1) + 2) : 0; }
__global__ void __launch_bounds__(64) KZ_WPE(KZ_WPE_BC) k_bc(GridD g, Dev d, kz_physics ph) {
  long t = TID;
  long nxg = (long)g.nz * g.ny * g.ngx * 2;
  long nyg = (long)g.nz * g.nx * g.ngy * 2;
  long nzg = (long)g.ny * g.nx * g.ngz * 2;
  if (t < nxg) bc_x_cell(g, d, ph, t);
  else if (t < nxg + nyg) bc_y_cell(g, d, ph, t - nxg);
  else if (t < nxg + nyg + nzg) bc_z_cell(g, d, ph, t - nxg - nyg);
  else if (t < nxg + nyg + nzg + bc_corner_count(g)) bc_corner2d_cell(g, d, ph, t - nxg - nyg - nzg);
}

// ---------------------------------------------------------------------------
// cellFixup (sim.zig:1027-1133) for one flagged cell: average the fixed slots
// from the unflagged in-domain neighbours in the order x-, x+, y-, y+, z-, z+
// (at least 1/2/3 of them in 1/2/3-D), then p2u.  `which` is one flag; the
// u2p path applies the hydro then the radiation fixup to a cell in turn,
// which equals the reference's two passes: they change disjoint slots, and
// each reads only the slots it changes from cells it does not change.
__device__ bool fixup_slots(const GridD &g, const Dev &d, const kz_physics &ph, long ix, long iy, long iz, int which,
                            double *pp) {
  long n = g.ncell;
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
  if (!enough) return false;
  for (int iv = 0; iv < NV; ++iv) {
    bool fix;
    if (which == KZ_FLAG_HD_FIXUP) fix = iv != RHO && iv < B1;
    else if (which == KZ_FLAG_RAD_FIXUP) fix = iv >= EE && iv <= FZ;
    else fix = iv != RHO && (iv < B1 || (iv >= EE && iv <= FZ));
    if (fix) {
      double sum = 0;
      for (int k = 0; k < in_n; ++k) sum += d.p[iv * n + ci[k]];
      pp[iv] = sum / (double)in_n;
    }
  }
  return true;
}
__device__ void fixup_cell(const GridD &g, const Dev &d, const kz_physics &ph, long t, bool hd, bool rad) {
  long ix = t % g.nx, iy = (t / g.nx) % g.ny, iz = t / ((long)g.nx * g.ny);
  if (polar_row(ph, g, iy)) return;
  long n = g.ncell;
  long c = cidx(g, ix, iy, iz);
  double pp[NV], uu[NV];
  ld(d.p, n, c, pp);
  bool changed = false;
  if (hd && d.flags[KZ_FLAG_HD_FIXUP * n + c] != 0) changed |= fixup_slots(g, d, ph, ix, iy, iz, KZ_FLAG_HD_FIXUP, pp);
  if (rad && d.flags[KZ_FLAG_RAD_FIXUP * n + c] != 0)
    changed |= fixup_slots(g, d, ph, ix, iy, iz, KZ_FLAG_RAD_FIXUP, pp);
  if (!changed) return;
  Geo gm;
  geo_c(g, d, ix, iy, iz, gm);
  p2u(pp, uu, gm, ph.gam);
  st(d.p, n, c, pp);
  st(d.u, n, c, uu);
}

// u2p pass, drift-frame floors (u2p_pass; the ZAMO-frame variant is in
// kz_options.hip)
__global__ void __launch_bounds__(128) KZ_WPE(KZ_WPE_U2P) k_u2p(GridD g, Dev d, kz_physics ph) {
  u2p_pass<false>(g, d, ph);
}
// the u2p fixups of the listed cells; one block, which then clears the list
__global__ void __launch_bounds__(256) k_u2p_fixups(GridD g, Dev d, kz_physics ph) {
  const bool fix_hd = ph.do_fixups, fix_rad = ph.do_fixups && ph.do_u2prad_fixups;
  unsigned nfix = d.sync[SY_NFIX];
  for (unsigned i = threadIdx.x; i < nfix; i += blockDim.x) fixup_cell(g, d, ph, d.fix_list[i], fix_hd, fix_rad);
  __syncthreads();
  if (threadIdx.x == 0) d.sync[SY_NFIX] = 0;
}

// the implicit-failure fixup pass (sim.zig:1771-1774), one thread per cell
__global__ void __launch_bounds__(256) k_fixup(GridD g, Dev d, kz_physics ph, int which) {
  long t = TID;
  if (t >= (long)g.nx * g.ny * g.nz) return;
  long ix = t % g.nx, iy = (t / g.nx) % g.ny, iz = t / ((long)g.nx * g.ny);
  long n = g.ncell;
  long c = cidx(g, ix, iy, iz);
  if (d.flags[which * n + c] == 0) return;
  if (polar_row(ph, g, iy)) return;
  double pp[NV], uu[NV];
  ld(d.p, n, c, pp);
  if (!fixup_slots(g, d, ph, ix, iy, iz, which, pp)) return;
  Geo gm;
  geo_c(g, d, ix, iy, iz, gm);
  p2u(pp, uu, gm, ph.gam);
  st(d.p, n, c, pp);
  st(d.u, n, c, uu);
}

// ---------------------------------------------------------------------------
// wavespeeds + CFL denominators (wavespeeds_pass; the synchrotron-bridge
// variant is in kz_options.hip)
__global__ void __launch_bounds__(128) KZ_WPE(KZ_WPE_WS) k_wavespeeds(GridD g, Dev d, kz_physics ph) {
  wavespeeds_pass<false>(g, d, ph);
}

// ---------------------------------------------------------------------------
// calcDivB (ct.zig:336-381): corner divergence of gdet*B from the cell-centred
// values of the cell and its lower neighbours (one ghost layer below each
// active dimension), divided by the cell's gdet; one thread per domain cell,
// out[ix + nx*(iy + ny*iz)].  The reference's 2D form divides by twice the
// centre spacing above the cell, its 3D form by four times the spacing below.
__device__ __forceinline__ double gdet_pb(const GridD &g, const Dev &d, int iv, long ix, long iy, long iz) {
  long c = cidx(g, ix, iy, iz);
  return d.gc[26 * g.ncell + c] * d.p[iv * g.ncell + c];
}
__global__ void __launch_bounds__(256) k_divb(GridD g, Dev d, double *out) {
  long t = TID;
  if (t >= (long)g.nx * g.ny * g.nz) return;
  long ix = t % g.nx, iy = (t / g.nx) % g.ny, iz = t / ((long)g.nx * g.ny);
  double divb;
  if (g.nz == 1) {
    divb = (gdet_pb(g, d, B1, ix, iy, iz) + gdet_pb(g, d, B1, ix, iy - 1, iz) - gdet_pb(g, d, B1, ix - 1, iy, iz) -
            gdet_pb(g, d, B1, ix - 1, iy - 1, iz)) /
               (2.0 * (xc_(g, ix + 1) - xc_(g, ix))) +
           (gdet_pb(g, d, B2, ix, iy, iz) + gdet_pb(g, d, B2, ix - 1, iy, iz) - gdet_pb(g, d, B2, ix, iy - 1, iz) -
            gdet_pb(g, d, B2, ix - 1, iy - 1, iz)) /
               (2.0 * (yc_(g, iy + 1) - yc_(g, iy)));
  } else {
    double sx = gdet_pb(g, d, B1, ix, iy, iz) + gdet_pb(g, d, B1, ix, iy - 1, iz) - gdet_pb(g, d, B1, ix - 1, iy, iz) -
                gdet_pb(g, d, B1, ix - 1, iy - 1, iz) + gdet_pb(g, d, B1, ix, iy, iz - 1) +
                gdet_pb(g, d, B1, ix, iy - 1, iz - 1) - gdet_pb(g, d, B1, ix - 1, iy, iz - 1) -
                gdet_pb(g, d, B1, ix - 1, iy - 1, iz - 1);
    double sy = gdet_pb(g, d, B2, ix, iy, iz) + gdet_pb(g, d, B2, ix - 1, iy, iz) - gdet_pb(g, d, B2, ix, iy - 1, iz) -
                gdet_pb(g, d, B2, ix - 1, iy - 1, iz) + gdet_pb(g, d, B2, ix, iy, iz - 1) +
                gdet_pb(g, d, B2, ix - 1, iy, iz - 1) - gdet_pb(g, d, B2, ix, iy - 1, iz - 1) -
                gdet_pb(g, d, B2, ix - 1, iy - 1, iz - 1);
    double sz = gdet_pb(g, d, B3, ix, iy, iz) + gdet_pb(g, d, B3, ix - 1, iy, iz) - gdet_pb(g, d, B3, ix, iy, iz - 1) -
                gdet_pb(g, d, B3, ix - 1, iy, iz - 1) + gdet_pb(g, d, B3, ix, iy - 1, iz) +
                gdet_pb(g, d, B3, ix - 1, iy - 1, iz) - gdet_pb(g, d, B3, ix, iy - 1, iz - 1) -
                gdet_pb(g, d, B3, ix - 1, iy - 1, iz - 1);
    divb = sx / (4.0 * (xc_(g, ix) - xc_(g, ix - 1))) + sy / (4.0 * (yc_(g, iy) - yc_(g, iy - 1))) +
           sz / (4.0 * (zc_(g, iz) - zc_(g, iz - 1)));
  }
  out[t] = divb / d.gc[26 * g.ncell + cidx(g, ix, iy, iz)];
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

// ---------------------------------------------------------------------------
// conserved update: flux divergence + metric source (sim.zig:1140-1175,
// 1542-1582); dt from the step scalars (slot `sslot`).
//
// The B rows of the face fluxes are the flux-CT rebuild of ct.zig:143-177,
// evaluated here from the corner EMFs with the same expressions instead of
// being written back to the face arrays first; the hydro/radiation rows come
// from flb.  The metric source visits (k, l, nu) in the reference order with
// T^k_l and R^k_l built one row k at a time.
__device__ __forceinline__ double emf_at(const GridD &g, const Dev &d, int c, long ix, long iy, long iz) {
  return d.emf[c][eidx(g, ix, iy, iz)];
}
// Two threads per cell, in different waves: for the block's 64 cells, wave 0
// forms the metric source terms and wave 1 the flux differences, so each
// wave runs one of the two paths with all its lanes; both leave their 13
// rows in LDS and the update u + du + ms dt is then written half by each
// wave (the same expression as one thread evaluating both).
__global__ void __launch_bounds__(128) KZ_WPE(KZ_WPE_UPDATE) k_update(GridD g, Dev d, kz_physics ph, int sslot) {
  __shared__ double sms[NV][64], sdu[NV][64];
  const int slot = threadIdx.x & 63, side = threadIdx.x >> 6;
  const long t = (long)blockIdx.x * 64 + slot;
  const bool live = t < (long)g.nx * g.ny * g.nz;
  long ix = 0, iy = 0, iz = 0, c = 0;
  const long n = g.ncell;
  const double dt = d.ss[sslot];
  if (live) {
    ix = t % g.nx; iy = (t / g.nx) % g.ny; iz = t / ((long)g.nx * g.ny);
    c = cidx(g, ix, iy, iz);
  }
  if (live && side == 0) {
    Geo gm;
    geo_c(g, d, ix, iy, iz, gm);
    double pp[NV];
    ld(d.p, n, c, pp);
    double ucon[4], ucov[4], bcon[4], bcov[4], bsq, urf[4];
    ucon_ucov(pp, gm, ucon, ucov);
    bcon_bcov_bsq(pp[B1], pp[B2], pp[B3], ucon, ucov, gm, bcon, bcov, &bsq);
    {
      double v[4] = {0, pp[FX], pp[FY], pp[FZ]};
      velr_to_vel4(v, gm, urf);
    }
    const double rho = pp[RHO], u = pp[UU], erf = pp[EE];
    const double pr = (ph.gam - 1.0) * u;
    const double eta = (rho + u + pr) + bsq;
    const double ptot = pr + 0.5 * bsq;
    double ms[NV];
#pragma unroll
    for (int i = 0; i < NV; ++i) ms[i] = 0;
    const double *kr = d.kris;
    const double gdetu = gm.gdet;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      double trow[4], rrow[4], tk[4], rk[4];
#pragma unroll
      for (int m = 0; m < 4; ++m) {
        trow[m] = eta * ucon[k] * ucon[m] + ptot * gm.G[k][m] - bcon[k] * bcon[m];  // hydro.zig:87-112
        rrow[m] = four_third * erf * urf[k] * urf[m] + one_third * erf * gm.G[k][m];  // radiation.zig:79-98
      }
#pragma unroll
      for (int l = 0; l < 4; ++l) {
        double st_ = 0, sr = 0;
#pragma unroll
        for (int m = 0; m < 4; ++m) {
          st_ += trow[m] * gm.g[m][l];
          sr += rrow[m] * gm.g[m][l];
        }
        tk[l] = st_;
        rk[l] = sr;
      }
#pragma unroll
      for (int l = 0; l < 4; ++l)
#pragma unroll
        for (int nu = 0; nu < 4; ++nu) {
          const int a = nu < k ? nu : k, b = nu < k ? k : nu;
          double K = kr[(l * 10 + tri(a, b)) * n + c];
          ms[UU + nu] += gdetu * tk[l] * K;
          ms[EE + nu] += gdetu * rk[l] * K;
        }
    }
#pragma unroll
    for (int i = 0; i < NV; ++i) sms[i][slot] = ms[i];
  } else if (live) {
    double dx = csize(g, ix, 0), dy = csize(g, iy, 1), dz = csize(g, iz, 2);
    long fx0 = fidx(g, 0, ix, iy, iz), fx1 = fidx(g, 0, ix + 1, iy, iz);
    long fy0 = fidx(g, 1, ix, iy, iz), fy1 = fidx(g, 1, ix, iy + 1, iz);
    long fz0 = fidx(g, 2, ix, iy, iz), fz1 = fidx(g, 2, ix, iy, iz + 1);
    long m0 = g.nface[0], m1 = g.nface[1], m2 = g.nface[2];
    const bool xa = g.nx > 1, ya = g.ny > 1, za = g.nz > 1;
    // CT-rebuilt B fluxes (ct.zig:160-175) at the six faces of the cell
    double bx0[3], bx1[3], by0[3], by1[3], bz0[3], bz1[3];
    for (int q = 0; q < 2; ++q) {
      long jx = ix + q, jy = iy + q, jz = iz + q;
      double *fxb = q ? bx1 : bx0, *fyb = q ? by1 : by0, *fzb = q ? bz1 : bz0;
      if (xa) {
        fxb[0] = 0;
        fxb[1] = 0.5 * (emf_at(g, d, 2, jx, iy, iz) + emf_at(g, d, 2, jx, iy + 1, iz));
        fxb[2] = -0.5 * (emf_at(g, d, 1, jx, iy, iz) + emf_at(g, d, 1, jx, iy, iz + 1));
      } else {
        long f = q ? fx1 : fx0;
        for (int k = 0; k < 3; ++k) fxb[k] = d.flb[0][(B1 + k) * m0 + f];
      }
      if (ya) {
        fyb[0] = -0.5 * (emf_at(g, d, 2, ix, jy, iz) + emf_at(g, d, 2, ix + 1, jy, iz));
        fyb[1] = 0;
        fyb[2] = 0.5 * (emf_at(g, d, 0, ix, jy, iz) + emf_at(g, d, 0, ix, jy, iz + 1));
      } else {
        long f = q ? fy1 : fy0;
        for (int k = 0; k < 3; ++k) fyb[k] = d.flb[1][(B1 + k) * m1 + f];
      }
      if (za) {
        fzb[0] = 0.5 * (emf_at(g, d, 1, ix, iy, jz) + emf_at(g, d, 1, ix + 1, iy, jz));
        fzb[1] = -0.5 * (emf_at(g, d, 0, ix, iy, jz) + emf_at(g, d, 0, ix, iy + 1, jz));
        fzb[2] = 0;
      } else {
        long f = q ? fz1 : fz0;
        for (int k = 0; k < 3; ++k) fzb[k] = d.flb[2][(B1 + k) * m2 + f];
      }
    }
#pragma unroll
    for (int iv = 0; iv < NV; ++iv) {
      double fxl, fxr, fyl, fyr, fzl, fzr;
      if (iv >= B1 && iv <= B3) {
        fxl = bx0[iv - B1]; fxr = bx1[iv - B1];
        fyl = by0[iv - B1]; fyr = by1[iv - B1];
        fzl = bz0[iv - B1]; fzr = bz1[iv - B1];
      } else {
        fxl = d.flb[0][iv * m0 + fx0]; fxr = d.flb[0][iv * m0 + fx1];
        fyl = d.flb[1][iv * m1 + fy0]; fyr = d.flb[1][iv * m1 + fy1];
        fzl = d.flb[2][iv * m2 + fz0]; fzr = d.flb[2][iv * m2 + fz1];
      }
      sdu[iv][slot] = -(fxr - fxl) * dt / dx - (fyr - fyl) * dt / dy - (fzr - fzl) * dt / dz;
    }
  }
  __syncthreads();
  if (!live) return;
  const int iv0 = side ? 7 : 0, iv1 = side ? NV : 7;
  for (int iv = iv0; iv < iv1; ++iv) {
    double uo = d.u[iv * n + c];
    d.u[iv * n + c] = uo + sdu[iv][slot] + sms[iv][slot] * dt;
  }
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
// radiative shear viscosity (sim/rijvisc.zig, physics/radvisc.zig)
__device__ __forceinline__ double krv(const Dev &d, long n, long c, int l, int a, int b) {
  return d.kris[(l * 10 + (a < b ? tri(a, b) : tri(b, a))) * n + c];
}
// math/misc.zig:9-12 stepFunction
__device__ __forceinline__ double step_function(double x, double x9) {
  double k = 1.47222 / x9;
  return 1.0 / (1.0 + kzm_exp(-2.0 * k * x));
}
// relele.convertBoth(VELR) of the radiation-frame velocity of cell c
__device__ __forceinline__ void rad_vel_both(const Dev &d, long n, long c, const Geo &gm, double con[4],
                                             double cov[4]) {
  double v[4] = {0.0, d.p[FX * n + c], d.p[FY * n + c], d.p[FZ * n + c]};
  velr_to_vel4(v, gm, con);
  lower4(con, gm, cov);
}
__device__ __forceinline__ bool is_corner(const GridD &g, long ix, long iy, long iz) {  // sim.zig:660-666
  int k = 0;
  if (ix < 0 || ix >= g.nx) k += 1;
  if (iy < 0 || iy >= g.ny) k += 1;
  if (iz < 0 || iz >= g.nz) k += 1;
  return k >= 2;
}
__device__ __forceinline__ double ccoord(const GridD &g, int idim, long i) {
  return idim == 1 ? xc_(g, i) : (idim == 2 ? yc_(g, i) : zc_(g, i));
}
// rijvisc.zig:46-172 calcShearLab then radvisc.zig:72-131
// shearFromGradients: sigma_{mu nu} of the radiation frame in the lab frame,
// centred differences except the one-sided ones next to the skipped corners.
__device__ void shear_lab(const GridD &g, const Dev &d, long ix, long iy, long iz, const Geo &geom,
                          double sig[4][4]) {
  const long n = g.ncell, c = cidx(g, ix, iy, iz);
  double du[4][4], du2[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) du[i][j] = du2[i][j] = 0.0;
  double ucon[4], ucov[4];
  rad_vel_both(d, n, c, geom, ucon, ucov);
  const long nd[3] = {g.nx, g.ny, g.nz};
  const long nx = g.nx, ny = g.ny, nz = g.nz;
#pragma unroll 1
  for (int idim = 1; idim < 4; ++idim) {
    const int dim = idim - 1;
    if (nd[dim] <= 1) continue;
    long cm[3] = {ix, iy, iz}, cp[3] = {ix, iy, iz};
    cm[dim] -= 1;
    cp[dim] += 1;
    double umc[4], umv[4], upc[4], upv[4];
    {
      Geo gx;
      geo_c(g, d, cm[0], cm[1], cm[2], gx);
      rad_vel_both(d, n, cidx(g, cm[0], cm[1], cm[2]), gx, umc, umv);
      geo_c(g, d, cp[0], cp[1], cp[2], gx);
      rad_vel_both(d, n, cidx(g, cp[0], cp[1], cp[2]), gx, upc, upv);
    }
    const double xm = ccoord(g, idim, cm[dim]), xc = ccoord(g, idim, dim == 0 ? ix : (dim == 1 ? iy : iz)),
                 xp = ccoord(g, idim, cp[dim]);
    // 0 centred, 1 right-sided, 2 left-sided
    int side = 0;
    if ((ix < 0 && iy == 0 && iz == 0 && idim != 1) || (iy < 0 && ix == 0 && iz == 0 && idim != 2) ||
        (iz < 0 && ix == 0 && iy == 0 && idim != 3))
      side = 1;
    else if ((ix < 0 && iy == ny - 1 && iz == nz - 1 && idim != 1) ||
             (iy < 0 && ix == nx - 1 && iz == nz - 1 && idim != 2) ||
             (iz < 0 && ix == nx - 1 && iy == ny - 1 && idim != 3))
      side = 2;
    else if ((ix >= nx && iy == 0 && iz == 0 && idim != 1) || (iy >= ny && ix == 0 && iz == 0 && idim != 2) ||
             (iz >= nz && ix == 0 && iy == 0 && idim != 3))
      side = 1;
    else if ((ix >= nx && iy == ny - 1 && iz == nz - 1 && idim != 1) ||
             (iy >= ny && ix == nx - 1 && iz == nz - 1 && idim != 2) ||
             (iz >= nz && ix == nx - 1 && iy == ny - 1 && idim != 3))
      side = 2;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      double a, b;
      if (side == 1) {
        a = (upv[i] - ucov[i]) / (xp - xc);
        b = (upc[i] - ucon[i]) / (xp - xc);
      } else if (side == 2) {
        a = (ucov[i] - umv[i]) / (xc - xm);
        b = (ucon[i] - umc[i]) / (xc - xm);
      } else {
        a = (upv[i] - umv[i]) / (xp - xm);
        b = (upc[i] - umc[i]) / (xp - xm);
      }
      // du[i][idim], du2[i][idim] with a register-friendly column select
#pragma unroll
      for (int q = 1; q < 4; ++q)
        if (q == idim) {
          du[i][q] = a;
          du2[i][q] = b;
        }
    }
  }
  double dcu[4][4], dcudiag[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      double krsum = 0;
#pragma unroll
      for (int k = 0; k < 4; ++k) krsum += krv(d, n, c, k, i, j) * ucov[k];
      dcu[i][j] = du[i][j] - krsum;
    }
    double krsum = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) krsum += krv(d, n, c, i, i, k) * ucon[k];
    dcudiag[i] = du2[i][i] + krsum;
  }
  double theta = 0;
#pragma unroll
  for (int i = 0; i < 4; ++i) theta += dcudiag[i];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) sig[i][j] = 0.0;
#pragma unroll
  for (int i = 1; i < 4; ++i)
#pragma unroll
    for (int j = 1; j < 4; ++j) {
      double sum1 = 0, sum2 = 0;
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        sum1 += dcu[i][k] * ((k == j ? 1.0 : 0.0) + ucon[k] * ucov[j]);
        sum2 += dcu[j][k] * ((k == i ? 1.0 : 0.0) + ucon[k] * ucov[i]);
      }
      sig[i][j] = 0.5 * (sum1 + sum2) - 1.0 / 3.0 * theta * (geom.g[i][j] + ucov[i] * ucov[j]);
    }
#pragma unroll
  for (int i = 1; i < 4; ++i) {
    double v = -1.0 / ucon[0] * (ucon[1] * sig[1][i] + ucon[2] * sig[2][i] + ucon[3] * sig[3][i]);
    sig[i][0] = v;
    sig[0][i] = v;
  }
  sig[0][0] = -1.0 / ucon[0] * (ucon[1] * sig[1][0] + ucon[2] * sig[2][0] + ucon[3] * sig[3][0]);
}
// radvisc.zig:154-175 viscCoeff
__device__ __forceinline__ double visc_coeff(double chi, double mindx, double r_bl, double rhor, double alpha,
                                             double global_dt) {
  double mfp = 1.0 / chi;
  const double rmin = 1.2 * rhor;
  const double mfplim = r_bl;
  if (mfp > mfplim || chi < 1.0e-80) mfp = mfplim;
  if (mfp < 0 || !kzm_isfinite(mfp)) mfp = 0;
  mfp *= step_function(r_bl - rmin, 0.2 * rmin);
  if (r_bl <= 1.0 * rhor) mfp = 0;
  double nu = alpha * mfp;
  const double nulimit = mindx * mindx / 2.0 / global_dt / 2.0;
  if (nu > nulimit) nu = nulimit;
  return nu;
}
// rijvisc.zig:266-311 calcRijViscTotal -> calcRijVisc (:175-264): one
// thread per cell of the domain and its one-cell ring, corners skipped (their
// R^i_j stays 0).  global_dt is the step-scalar slot `sslot`.
__global__ void __launch_bounds__(64) k_rijvisc(GridD g, Dev d, kz_physics ph, int sslot, double rhor) {
  long t = TID;
  const long lx = g.nx > 1, ly = g.ny > 1, lz = g.nz > 1;
  const long wx = g.nx + 2 * lx, wy = g.ny + 2 * ly, wz = g.nz + 2 * lz;
  if (t >= wx * wy * wz) return;
  const long ix = t % wx - lx, iy = (t / wx) % wy - ly, iz = t / (wx * wy) - lz;
  if (is_corner(g, ix, iy, iz)) return;
  const long n = g.ncell, c = cidx(g, ix, iy, iz);
  const double global_dt = d.ss[sslot];
  Geo geom;
  geo_c(g, d, ix, iy, iz, geom);
  double sl[4][4];
  shear_lab(g, d, ix, iy, iz, geom, sl);
  // relele.zig:115-128 raiseBoth
  double su[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      double acc = 0;
#pragma unroll
      for (int k = 0; k < 4; ++k)
#pragma unroll
        for (int l = 0; l < 4; ++l) acc += sl[k][l] * geom.G[i][k] * geom.G[j][l];
      su[i][j] = acc;
    }
  double pp[NV];
  ld(d.p, n, c, pp);
  const double chi = calc_chi_slim(pp, geom, ph.gam, ph.opac);
  const double dx0 = csize(g, ix, 0) * kzm_sqrt(geom.g[1][1]), dx1 = csize(g, iy, 1) * kzm_sqrt(geom.g[2][2]),
               dx2 = csize(g, iz, 2) * kzm_sqrt(geom.g[3][3]);
  const double mindx = (g.ny == 1 && g.nz == 1) ? dx0
                       : (g.nz == 1)            ? kzm_fmin(dx0, dx1)
                       : (g.ny == 1)            ? kzm_fmin(dx0, dx2)
                                                : kzm_fmin(dx0, kzm_fmin(dx1, dx2));
  const double nu = visc_coeff(chi, mindx, d.bl[BL_R * n + c], rhor, ph.radvisc.alpha, global_dt);
  double rv[4][4], lo[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) rv[i][j] = -2.0 * nu * pp[EE] * su[i][j];
  lower_second(rv, geom, lo);
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) d.rijvisc[(i * 4 + j) * n + c] = lo[i][j];
}

// ---------------------------------------------------------------------------
// mean-field dynamo (sim/dynamo.zig)
__device__ __forceinline__ void bl_geo(const Dev &d, long n, long c, Geo &o) {
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      o.g[i][j] = d.bl[(BL_G + i * 4 + j) * n + c];
      o.G[i][j] = d.bl[(BL_GC + i * 4 + j) * n + c];
    }
  o.alpha = d.bl[BL_ALPHA * n + c];
  o.gdet = d.bl[BL_GDET * n + c];
  o.gttpert = 0;
  o.x1 = d.bl[BL_R * n + c];
}
// dynamo.zig:226-256 calcScaleHeight (single rank): per radial column the
// density-weighted RMS of pi/2 - theta, summed in the reference's row order;
// column 0 keeps the raw sum (the reference's gix > 0 condition).
__global__ void __launch_bounds__(64) k_scale_height(GridD g, Dev d) {
  const long ix = TID;
  if (ix >= g.nx) return;
  const long n = g.ncell;
  double sigma = 0, sth = 0;
  for (long iy = 0; iy < g.ny; ++iy)
    for (long iz = 0; iz < g.nz; ++iz) {
      const long c = cidx(g, ix, iy, iz);
      const double gd = d.gc[26 * n + c];
      const double th = d.bl[BL_TH * n + c];
      const double rho = d.p[RHO * n + c];
      const double dth = KZ_PI / 2.0 - th;
      sigma += rho * gd;
      sth += rho * gd * dth * dth;
    }
  d.scaleth[ix] = ix > 0 ? kzm_sqrt(sth / sigma) : sth;
}
// dynamo.zig:67-73 dampBphi
__device__ __forceinline__ double damp_bphi(double alphabeta, double facradius, double faczh, double dt_over_pk,
                                            double beta, double betasat, double bphi) {
  double dbphi = -alphabeta * facradius * faczh * dt_over_pk * kzm_fmax(0.0, beta - betasat) / betasat * bphi;
  if ((dbphi + bphi) * bphi < 0.0) dbphi = -bphi;
  return dbphi;
}
// dynamo.zig:104-151 dynamoDeltaA (the caller skips r < 1.0001 r_horizon)
__device__ double dynamo_delta_a(const kz_dynamo_params &dp, double a, double dt, double risco, double r, double th,
                                 double bphi, double bsq, double angle, double prermhd, double gg33, double scaleth,
                                 double *bphi_out) {
  *bphi_out = bphi;
  const double omk = 1.0 / (a + kzm_sqrt(r * r * r));
  const double pk = 2.0 * KZ_PI / omk;
  double facangle = 0;
  if (kzm_isfinite(angle)) {
    if (angle < -1.0) angle = -1.0;
    facangle = kzm_fmax(0.0, (dp.thetaangle - angle) / dp.thetaangle);
  }
  const double facradius = step_function(r - 1.0 * risco, 0.1 * risco);
  const double beta = 0.5 * bsq / prermhd;
  double hrdtheta;
  if (dp.calchronthego) {
    hrdtheta = scaleth;
    if (hrdtheta > 0.9 * KZ_PI / 2.0) hrdtheta = 0.9 * KZ_PI / 2.0;
  } else {
    hrdtheta = dp.expectedhr * KZ_PI / 2.0;
  }
  const double zh = (KZ_PI / 2.0 - th) / hrdtheta;
  const double faczh = kzm_fmax(0.0, 1.0 - zh * zh);
  const double facmagnetization = faczh;
  double effalpha = dp.alphadynamo;
  if (dp.alphaflipssign) effalpha = -(KZ_PI / 2.0 - th) / (hrdtheta / 2.0) * dp.alphadynamo;
  const double aphi = effalpha * (hrdtheta / (KZ_PI / 2.0)) / 0.4 * dt / pk * r * gg33 * bphi * facradius *
                      facmagnetization * facangle;
  if (dp.dampbeta)
    *bphi_out = bphi + damp_bphi(dp.alphabeta, facradius, faczh, dt / pk, beta, dp.betasaturated, bphi);
  return aphi;
}
// dynamo.zig:370-430 mimicDynamo, first half: Delta A_phi on the domain and
// its one-cell ring (corners included) with the field angle of
// dynamo.zig:335-368 evaluated in BL; DAMPBETA rewrites p[B3] in place.
__global__ void __launch_bounds__(64) k_dynamo_a(GridD g, Dev d, kz_physics ph, int sslot, double rhor,
                                                 double risco) {
  long t = TID;
  const long lx = g.nx > 1, ly = g.ny > 1, lz = g.nz > 1;
  const long wx = g.nx + 2 * lx, wy = g.ny + 2 * ly, wz = g.nz + 2 * lz;
  if (t >= wx * wy * wz) return;
  const long ix = t % wx - lx, iy = (t / wx) % wy - ly, iz = t / (wx * wy) - lz;
  const long n = g.ncell, c = cidx(g, ix, iy, iz);
  const double dt = d.ss[sslot];
  Geo geom, gbl;
  geo_c(g, d, ix, iy, iz, geom);
  bl_geo(d, n, c, gbl);
  double pp[NV], ppbl[NV];
  ld(d.p, n, c, pp);
#pragma unroll
  for (int iv = 0; iv < NV; ++iv) ppbl[iv] = pp[iv];
  double J[16];
#pragma unroll
  for (int k = 0; k < 16; ++k) J[k] = d.bl[(BL_J + k) * n + c];
  const int e = trans_pmhd(ppbl, geom, gbl, J);
  if (e) {
    set_err(d.err, e);
    return;
  }
  double angle, bsq;
  {
    double ucon[4], ucov[4], bcon[4], bcov[4], b2;
    ucon_ucov(ppbl, gbl, ucon, ucov);
    bcon_bcov_bsq(ppbl[B1], ppbl[B2], ppbl[B3], ucon, ucov, gbl, bcon, bcov, &b2);
    const double brbphi = kzm_sqrt(gbl.g[1][1] * gbl.g[3][3]) * bcon[1] * bcon[3];
    angle = -brbphi / b2;
    bsq = b2;
  }
  const double r = d.bl[BL_R * n + c], th = d.bl[BL_TH * n + c];
  if (r < 1.0001 * rhor) {
    d.dyn_a[c] = 0.0;
    return;
  }
  double prermhd = (ph.gam - 1.0) * pp[UU];
  const double ehat = -calc_ff_rtt(pp, geom);
  prermhd += ehat / 3.0;
  const long gix = ix < 0 ? 0 : (ix > g.nx - 1 ? g.nx - 1 : ix);
  double bphi_new;
  const double aphi = dynamo_delta_a(ph.dynamo_params, ph.bh_spin, dt, risco, r, th, pp[B3], bsq, angle, prermhd,
                                     geom.g[3][3], d.scaleth[gix], &bphi_new);
  d.dyn_a[c] = aphi;
  if (ph.dynamo_params.dampbeta) d.p[B3 * n + c] = bphi_new;
}
// mimicDynamo, second half: the corner average of (0, 0, Delta A_phi)
// (ct.zig:228-264 cornerAverageA), its curl (ct.zig:279-331 calcBfromACore)
// and the superposition on the domain's B (dynamo.zig:432-460), with u[B]
// from p2u.  The corner values are formed per cell from the four (2D) or
// eight (3D) cells around each corner, in the reference's sum order.
__device__ __forceinline__ double dyn_a_at(const GridD &g, const Dev &d, long ix, long iy, long iz) {
  return d.dyn_a[cidx(g, ix, iy, iz)];
}
// corner A_phi (cornerAverageA): 4 cells around the corner in 2D, 8 in 3D
__device__ __forceinline__ double corner_a(const GridD &g, const Dev &d, long ix, long iy, long iz) {
  if (g.nz == 1)
    return 1.0 / 4.0 *
           (dyn_a_at(g, d, ix, iy, iz) + dyn_a_at(g, d, ix, iy - 1, iz) + dyn_a_at(g, d, ix - 1, iy, iz) +
            dyn_a_at(g, d, ix - 1, iy - 1, iz));
  return 1.0 / 8.0 *
         (dyn_a_at(g, d, ix, iy, iz) + dyn_a_at(g, d, ix, iy - 1, iz) + dyn_a_at(g, d, ix - 1, iy, iz) +
          dyn_a_at(g, d, ix - 1, iy - 1, iz) + dyn_a_at(g, d, ix, iy, iz - 1) + dyn_a_at(g, d, ix, iy - 1, iz - 1) +
          dyn_a_at(g, d, ix - 1, iy, iz - 1) + dyn_a_at(g, d, ix - 1, iy - 1, iz - 1));
}
__global__ void __launch_bounds__(128) k_dynamo_curl(GridD g, Dev d, kz_physics ph) {
  long t = TID;
  if (t >= (long)g.nx * g.ny * g.nz) return;
  const long ix = t % g.nx, iy = (t / g.nx) % g.ny, iz = t / ((long)g.nx * g.ny);
  const long n = g.ncell, c = cidx(g, ix, iy, iz);
  const bool d3 = g.nz > 1;
  // the A_r and A_theta corners are averages of zeros
  const double z = d3 ? 1.0 / 8.0 * (0.0 + 0.0 + 0.0 + 0.0 + 0.0 + 0.0 + 0.0 + 0.0)
                      : 1.0 / 4.0 * (0.0 + 0.0 + 0.0 + 0.0);
  // corners (jx, jy, jz) in {ix, ix+1} x {iy, iy+1} x {iz, iz+1}: index jx-ix + 2 (jy-iy) + 4 (jz-iz)
  double ac[8];
#pragma unroll
  for (int q = 0; q < 8; ++q) ac[q] = (q >= 4 && !d3) ? 0.0 : corner_a(g, d, ix + (q & 1), iy + ((q >> 1) & 1), iz + (q >> 2));
  double dA[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) dA[i][j] = 0.0;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
#define A_(q) (k == 2 ? ac[q] : z)
    if (!d3) {
      // A(ix,iy)=q0, A(ix,iy+1)=q2, A(ix+1,iy)=q1, A(ix+1,iy+1)=q3
      dA[k + 1][2] = -(A_(0) - A_(2) + A_(1) - A_(3)) / (2.0 * csize(g, iy, 1));
      dA[k + 1][1] = -(A_(0) + A_(2) - A_(1) - A_(3)) / (2.0 * csize(g, ix, 0));
      dA[k + 1][3] = 0;
    } else {
      dA[k + 1][1] = (A_(1) - A_(0) + A_(3) - A_(2) + A_(5) - A_(4) + A_(7) - A_(6)) / (4.0 * csize(g, ix, 0));
      dA[k + 1][2] = (A_(2) - A_(0) + A_(3) - A_(1) + A_(6) - A_(4) + A_(7) - A_(5)) / (4.0 * csize(g, iy, 1));
      dA[k + 1][3] = (A_(4) - A_(0) + A_(5) - A_(1) + A_(6) - A_(2) + A_(7) - A_(3)) / (4.0 * csize(g, iz, 2));
    }
#undef A_
  }
  const double gdet = d.gc[26 * n + c];
  const double v3 = (dA[2][3] - dA[3][2]) / gdet;
  const double v4 = (dA[3][1] - dA[1][3]) / gdet;
  const double v5 = (dA[1][2] - dA[2][1]) / gdet;
  double pp[NV], uu[NV];
  ld(d.p, n, c, pp);
  pp[B1] += v3;
  pp[B2] += v4;
  pp[B3] += v5;
  Geo geom;
  geo_c(g, d, ix, iy, iz, geom);
  p2u(pp, uu, geom, ph.gam);
  d.p[B1 * n + c] = pp[B1];
  d.p[B2 * n + c] = pp[B2];
  d.p[B3 * n + c] = pp[B3];
  d.u[B1 * n + c] = uu[B1];
  d.u[B2 * n + c] = uu[B2];
  d.u[B3 * n + c] = uu[B3];
}

// stage arithmetic over domain cells (sim.zig:1586-1709), fused per step
// phase.  Each kernel applies, per element, the reference's expressions in
// the reference's order (stageDeriv f1*a + f2*b, stageCombine a + f1*b +
// f2*c).  The stage copies ut0/ut1/uforget/ut2 are only ever read on domain
// cells, so only domain cells are copied.
__device__ __forceinline__ bool dom_cell(const GridD &g, long t, long &c) {
  if (t >= (long)g.nx * g.ny * g.nz) return false;
  long ix = t % g.nx, iy = (t / g.nx) % g.ny, iz = t / ((long)g.nx * g.ny);
  c = cidx(g, ix, iy, iz);
  return true;
}
// step start (sim.zig:1792-1830): own_dt = 1/tstepdenmax, dt, the stage
// coefficients, the CFL accumulators reset; ut0 = u.
__global__ void k_begin_step(GridD g, Dev d, double forced, int use_forced) {
  long t = TID, c;
  if (t == 0) {
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
  if (!dom_cell(g, t, c)) return;
  long n = g.ncell;
#pragma unroll
  for (int iv = 0; iv < NV; ++iv) d.ut0[iv * n + c] = d.u[iv * n + c];
}
// after an implicit stage: drt = u/(dt g) - base/(dt g); next stage copy = u
__global__ void k_stage_implicit(GridD g, Dev d, double *drt, const double *base, double *copy) {
  long t = TID, c;
  if (!dom_cell(g, t, c)) return;
  long n = g.ncell;
  const double f1 = d.ss[SS_INV_DTG], f2 = d.ss[SS_NINV_DTG];
#pragma unroll
  for (int iv = 0; iv < NV; ++iv) {
    long o = iv * n + c;
    double u = d.u[o];
    drt[o] = f1 * u + f2 * base[o];
    copy[o] = u;
  }
}
// after the first explicit stage: copyEntropyCount, dut1, the first
// together-combine u = ut0 + dt dut1 + dt(1-2g) drt1, uforget = u
__global__ void k_stage_explicit1(GridD g, Dev d) {
  long t = TID, c;
  if (!dom_cell(g, t, c)) return;
  long n = g.ncell;
  d.flags[KZ_FLAG_ENTROPY3 * n + c] = d.flags[KZ_FLAG_ENTROPY * n + c];
  const double f1 = d.ss[SS_INV_DT], f2 = d.ss[SS_NINV_DT], dt = d.ss[SS_DT], c12 = d.ss[SS_C12];
#pragma unroll
  for (int iv = 0; iv < NV; ++iv) {
    long o = iv * n + c;
    double du = f1 * d.u[o] + f2 * d.ut1[o];
    d.dut1[o] = du;
    double u = d.ut0[o] + dt * du + c12 * d.drt1[o];
    d.u[o] = u;
    d.uforget[o] = u;
  }
}
// after the second explicit stage: dut2, u = ut0 + dt/2 dut1 + dt/2 dut2,
// then u += dt/2 drt1 + dt/2 drt2.  dut2 is read only here, so it stays in
// a register (the reference keeps it as a field, sim.zig:1877-1880).
__global__ void k_stage_explicit2(GridD g, Dev d) {
  long t = TID, c;
  if (!dom_cell(g, t, c)) return;
  long n = g.ncell;
  const double f1 = d.ss[SS_INV_DT], f2 = d.ss[SS_NINV_DT], h = d.ss[SS_HALF];
#pragma unroll
  for (int iv = 0; iv < NV; ++iv) {
    long o = iv * n + c;
    double du2 = f1 * d.u[o] + f2 * d.ut2[o];
    double u = d.ut0[o] + h * d.dut1[o] + h * du2;
    d.u[o] = u + h * d.drt1[o] + h * d.drt2[o];
  }
}
// step end: t += dt, then updateEntropy (sim.zig:1683-1709)
__global__ void k_end_step(GridD g, Dev d, kz_physics ph) {
  long t = TID;
  if (t == 0) d.ss[SS_T] = d.ss[SS_T] + d.ss[SS_DT];
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
  // BL horizon and ISCO radii (metric.zig:20-32) and whether the BL sidecar
  // has been loaded (radiative viscosity and dynamo need it)
  double rhor = 0, risco = 0;
  bool has_bl = false;
  // event timing of one kernel class (kz_sim_time_kernel)
  int timed = -1;
  std::vector<hipEvent_t> ev_pool;  // created events, reused
  size_t ev_used = 0;               // events recorded since the last read
  // device bytes allocated for this run, and the captured multi-step graph
  size_t dev_bytes = 0;
  double *scal_buf = nullptr;  // per-cell terms of the scalar diagnostics (allocated on first use)
  double wedge = 0;            // maxz - minz of the caller's grid
  unsigned persist_blocks = 1024;  // one wave per SIMD for the persistent implicit launches
  int imp_launch = 0;              // 0 by size, 1 merged first rung, 2 persistent
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  int graph_steps = 0;
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

// keep the first error of a sequence of set-up calls (error codes are
// distinct negative values: OR-ing them would name another error)
static inline void first_err(int &e, int code) {
  if (!e) e = code;
}
static inline int hip_err(hipError_t r) { return r == hipSuccess ? 0 : KZ_E_HIP; }

static int dalloc(kz_sim *s, void **p, size_t bytes) {
  if (hipMalloc(p, bytes) != hipSuccess) return KZ_E_OOM;
  s->allocs.push_back(*p);
  s->dev_bytes += bytes;
  return hipMemsetAsync(*p, 0, bytes, s->stream) == hipSuccess ? 0 : KZ_E_HIP;  // ordered before any use
}


static inline uint64_t bits_of(double x) {
  uint64_t b;
  memcpy(&b, &x, sizeof b);
  return b;
}
// Repack the reference's 20-double g/G blocks into the GEO_NC SoA planes.
// Returns false if a g block is not exactly symmetric.
static bool pack_geo(const double *g20, const double *G20, long n, std::vector<double> &out) {
  out.assign((size_t)n * GEO_NC, 0.0);
  for (long c = 0; c < n; ++c) {
    const double *g = g20 + c * 20, *G = G20 + c * 20;
    for (int a = 0; a < 4; ++a)
      for (int b = a; b < 4; ++b) {
        if (bits_of(g[a * 5 + b]) != bits_of(g[b * 5 + a])) return false;
        out[(size_t)tri(a, b) * n + c] = g[a * 5 + b];
      }
    for (int a = 0; a < 4; ++a)
      for (int b = 0; b < 4; ++b) out[(size_t)(10 + a * 4 + b) * n + c] = G[a * 5 + b];
    out[26 * n + c] = g[19];
    out[27 * n + c] = G[19];
    out[28 * n + c] = __builtin_sqrt(-1.0 / G[0]);  // precompute.zig:494
  }
  return true;
}
// Christoffel blocks kr[l*16 + nu*4 + k] -> KR_NC planes (l, tri(nu, k));
// false if a block is not exactly symmetric in (nu, k).
static bool pack_kris(const double *kris, long n, std::vector<double> &out) {
  out.assign((size_t)n * KR_NC, 0.0);
  for (long c = 0; c < n; ++c) {
    const double *k = kris + c * 64;
    for (int l = 0; l < 4; ++l)
      for (int a = 0; a < 4; ++a)
        for (int b = a; b < 4; ++b) {
          if (bits_of(k[l * 16 + a * 4 + b]) != bits_of(k[l * 16 + b * 4 + a])) return false;
          out[(size_t)(l * 10 + tri(a, b)) * n + c] = k[l * 16 + a * 4 + b];
        }
  }
  return true;
}

extern "C" {

kz_sim *kz_sim_create(const kz_grid *gr, const kz_physics *ph, const double *g, const double *gcon,
                      const double *kris, const double *gb0, const double *gconb0, const double *gb1,
                      const double *gconb1, const double *gb2, const double *gconb2, const double *bcx, int *err) {
  *err = 0;
  if (gr->ny <= 1 || gr->nx <= 1 || gr->ng < 3) { *err = KZ_E_UNSUPPORTED; return nullptr; }
  if (ph->floors.b2rhofloorframe < 0 || ph->floors.b2rhofloorframe > 1) {
    *err = KZ_E_INVALID_CONFIG;
    return nullptr;
  }
  if (ph->correct_polaraxis && gr->ny <= 2 * ph->nccorrectpolar) { *err = KZ_E_INVALID_CONFIG; return nullptr; }
  if (ph->radviscosity && !ph->has_opac) { *err = KZ_E_INVALID_CONFIG; return nullptr; }
  kz_sim *s = new kz_sim();
  s->ph = *ph;
  s->wedge = gr->maxz - gr->minz;
  {
    int dev = 0, ncu = 0;
    if (hipGetDevice(&dev) == hipSuccess && hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && ncu > 0)
      s->persist_blocks = (unsigned)(KZ_PERSIST_PER_CU * ncu);
  }
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
  // repack and check the metric tables before touching the device
  std::vector<double> pk_gc, pk_gf[3], pk_kr;
  bool sym = pack_geo(g, gcon, G.ncell, pk_gc);
  const double *gbs[3] = {gb0, gb1, gb2}, *gcbs[3] = {gconb0, gconb1, gconb2};
  for (int d = 0; d < 3; ++d) sym = sym && pack_geo(gbs[d], gcbs[d], G.nface[d], pk_gf[d]);
  sym = sym && pack_kris(kris, G.ncell, pk_kr);
  if (!sym) {  // MetricCache builds g and the Christoffels symmetric bit for bit
    *err = KZ_E_INVALID_CONFIG;
    delete s;
    return nullptr;
  }
  if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess) { *err = KZ_E_HIP; delete s; return nullptr; }
  size_t fb = (size_t)G.ncell * NV * sizeof(double);
  Dev &D = s->d;
  double **fl[] = {&D.p, &D.u, &D.ut0, &D.ut1, &D.ut2, &D.dut1, &D.drt1, &D.drt2, &D.uforget, &D.stage};
  int e = 0;
  for (auto f : fl) first_err(e, dalloc(s, (void **)f, fb));
  for (int d = 0; d < 3; ++d) first_err(e, dalloc(s, (void **)&D.flb[d], (size_t)G.nface[d] * NV * sizeof(double)));
  first_err(e, dalloc(s, (void **)&D.scal, (size_t)G.ncell * SC_N * sizeof(double)));
  first_err(e, dalloc(s, (void **)&D.flags, (size_t)G.ncell * KZ_NFLAGS * sizeof(int32_t)));
  long ncorn = (long)(gr->nx + 1) * (gr->ny + 1) * (gr->nz + 1);
  for (int c = 0; c < 3; ++c) first_err(e, dalloc(s, (void **)&D.emf[c], ncorn * sizeof(double)));
  first_err(e, dalloc(s, (void **)&D.gc, (size_t)G.ncell * GEO_NC * sizeof(double)));
  for (int d = 0; d < 3; ++d) first_err(e, dalloc(s, (void **)&D.gf[d], (size_t)G.nface[d] * GEO_NC * sizeof(double)));
  first_err(e, dalloc(s, (void **)&D.kris, (size_t)G.ncell * KR_NC * sizeof(double)));
  size_t nbcx = (size_t)gr->nz * gr->ny * gr->ng * BCX_REC;
  first_err(e, dalloc(s, (void **)&D.bcx, nbcx * sizeof(double)));
  first_err(e, dalloc(s, (void **)&D.cnt, 8 * sizeof(unsigned long long)));
  first_err(e, dalloc(s, (void **)&D.tsd, 2 * sizeof(unsigned long long)));
  first_err(e, dalloc(s, (void **)&D.err, sizeof(int)));
  first_err(e, dalloc(s, (void **)&D.ss, SS_N * sizeof(double)));
  long nd = (long)gr->nx * gr->ny * gr->nz;
  first_err(e, dalloc(s, (void **)&D.imp_uu00, (size_t)nd * NV * sizeof(double)));
  first_err(e, dalloc(s, (void **)&D.imp_pp0, (size_t)nd * NV * sizeof(double)));
  first_err(e, dalloc(s, (void **)&D.imp_aux, (size_t)nd * IMP_NAUX * sizeof(double)));
  first_err(e, dalloc(s, (void **)&D.imp_slow, (size_t)nd * sizeof(int32_t)));
  first_err(e, dalloc(s, (void **)&D.imp_nslow, sizeof(unsigned long long)));
  first_err(e, dalloc(s, (void **)&D.imp_list, (size_t)2 * nd * sizeof(int32_t)));
  first_err(e, dalloc(s, (void **)&D.imp_ncnt, 2 * sizeof(unsigned long long)));
  first_err(e, dalloc(s, (void **)&D.imp_rl, (size_t)2 * nd * sizeof(int32_t)));
  first_err(e, dalloc(s, (void **)&D.imp_rc, 2 * sizeof(unsigned long long)));
  first_err(e, dalloc(s, (void **)&D.imp_qpos, 3 * sizeof(unsigned long long)));
  first_err(e, dalloc(s, (void **)&D.sync, SY_N * sizeof(unsigned)));
  first_err(e, dalloc(s, (void **)&D.fix_list, (size_t)nd * sizeof(int32_t)));
  if (ph->radviscosity || ph->dynamo) first_err(e, dalloc(s, (void **)&D.bl, (size_t)G.ncell * BL_NC * sizeof(double)));
  if (ph->radviscosity) first_err(e, dalloc(s, (void **)&D.rijvisc, (size_t)G.ncell * 16 * sizeof(double)));
  if (ph->dynamo) {
    first_err(e, dalloc(s, (void **)&D.dyn_a, (size_t)G.ncell * sizeof(double)));
    first_err(e, dalloc(s, (void **)&D.scaleth, (size_t)gr->nx * sizeof(double)));
  }
  if (e) { *err = e < 0 ? e : KZ_E_HIP; kz_sim_destroy(s); return nullptr; }
  // metric tables
  first_err(e, hip_err(hipMemcpyAsync(D.gc, pk_gc.data(), pk_gc.size() * sizeof(double), hipMemcpyHostToDevice, s->stream)));
  for (int d = 0; d < 3; ++d)
    first_err(e, hip_err(hipMemcpyAsync(D.gf[d], pk_gf[d].data(), pk_gf[d].size() * sizeof(double), hipMemcpyHostToDevice, s->stream)));
  first_err(e, hip_err(hipMemcpyAsync(D.kris, pk_kr.data(), pk_kr.size() * sizeof(double), hipMemcpyHostToDevice, s->stream)));
  first_err(e, hip_err(hipMemcpyAsync(D.bcx, bcx, nbcx * sizeof(double), hipMemcpyHostToDevice, s->stream)));
  // the host staging vectors must outlive the copies
  first_err(e, hip_err(hipStreamSynchronize(s->stream)));
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
  {  // metric.zig:20-32 rHorizonBL / rIscoBL (pow(., 1/2) taken as a square root)
    const double a = ph->bh_spin, c3 = 1.0 / 3.0;
    s->rhor = 1.0 + kzm_sqrt(1.0 - a * a);
    double z1 = 1.0 + kzm_pow(1.0 - a * a, c3) * (kzm_pow(1.0 + a, c3) + kzm_pow(1.0 - a, c3));
    double z2 = kzm_sqrt(3.0 * a * a + z1 * z1);
    s->risco = 3.0 + z2 - kzm_sqrt((3.0 - z1) * (3.0 + z1 + 2.0 * z2));
  }
  s->t_host = 0;
  s->nstep = 0;
  unsigned long long tsd0[2] = {(unsigned long long)__builtin_bit_cast(long long, 0.0),
                                (unsigned long long)__builtin_bit_cast(long long, 1.0e50)};
  if (hipMemcpyAsync(D.tsd, tsd0, sizeof tsd0, hipMemcpyHostToDevice, s->stream) != hipSuccess ||
      hipStreamSynchronize(s->stream) != hipSuccess) {
    *err = KZ_E_HIP;
    kz_sim_destroy(s);
    return nullptr;
  }
  return s;
}

void kz_sim_destroy(kz_sim *s) {
  if (!s) return;
  (void)hipStreamSynchronize(s->stream);
  if (s->graph_exec) (void)hipGraphExecDestroy(s->graph_exec);
  if (s->graph) (void)hipGraphDestroy(s->graph);
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
  // stream-ordered copies: no legacy-stream work, which HIP refuses while
  // another thread captures a graph
  HIPCHK(hipMemcpyAsync(&e, s->d.err, sizeof(int), hipMemcpyDeviceToHost, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  if (e) {
    int z = 0;
    HIPCHK(hipMemcpyAsync(s->d.err, &z, sizeof(int), hipMemcpyHostToDevice, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
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

int kz_sim_calc_divb(kz_sim *s, double *out) {
  if (!s || !out) return KZ_E_INVALID_CONFIG;
  const long nd = ndom(s);
  k_divb<<<nblk(nd, 256), 256, 0, s->stream>>>(s->g, s->d, s->d.stage);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(out, s->d.stage, nd * sizeof(double), hipMemcpyDeviceToHost, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  return 0;
}

// cells of the domain plus one ring in every active dimension
static long ring_count(const GridD &g) {
  return (long)(g.nx > 1 ? g.nx + 2 : g.nx) * (g.ny > 1 ? g.ny + 2 : g.ny) * (g.nz > 1 ? g.nz + 2 : g.nz);
}

static void launch_set_bc(kz_sim *s) {
  const GridD &g = s->g;
  TimedLaunch tl(s, KZ_K_BC);
  long nb = (long)g.nz * g.ny * g.ngx * 2 + (long)g.nz * g.nx * g.ngy * 2 + (long)g.ny * g.nx * g.ngz * 2 +
            (g.nz == 1 ? 4L * (2 * (g.ng - 1) + 2) : 0);  // + the 2D corner cells
  k_bc<<<nblk(nb, 64), 64, 0, s->stream>>>(g, s->d, s->ph);
  if (g.nz > 1) {
    const long nslice = 4L * (g.nx + g.ny + g.nz), per = 2L * (g.ng - 1) + 1;
    k_bc_edges1<<<nblk(nslice * per, 64), 64, 0, s->stream>>>(g, s->d, s->ph);
    k_bc_edges2<<<nblk(nslice + 8L * g.ng * g.ng * g.ng, 64), 64, 0, s->stream>>>(g, s->d, s->ph);
  }
}

int kz_sim_set_bc(kz_sim *s, int ifinit) {
  (void)ifinit;
  launch_set_bc(s);
  return check_err(s);
}

// calcU2p (sim.zig:924-1026): u2p + fixups, then setBc when `with_bc`.
// Inside a step every calcU2p is followed either by doCorrect + setBc or by
// passes that read no ghost cell before the next setBc (the implicit solve,
// fixups, polar overwrite and stage arithmetic read domain cells only), so
// the step omits that setBc: it would write ghosts that are overwritten
// before they are read.
static void launch_calc_u2p(kz_sim *s, bool with_bc) {
  {
    TimedLaunch tl(s, KZ_K_U2P);
    if (s->ph.floors.b2rhofloorframe == 1) opt_launch_u2p_zamo(s->g, s->d, s->ph, nblk(ndom(s), 128), s->stream);
    else k_u2p<<<nblk(ndom(s), 128), 128, 0, s->stream>>>(s->g, s->d, s->ph);
  }
  if (s->ph.do_fixups) k_u2p_fixups<<<1, 256, 0, s->stream>>>(s->g, s->d, s->ph);
  if (with_bc) launch_set_bc(s);
}

// The face-flux blocks of each dimension (k_face_flux): x faces in runs of
// fx whose adjacent cells, over at most ceil((fx - 1) / (nx + 1)) row
// changes, number at most 64; y and z faces in FT_T x FT_F tiles over the
// cross extent (ghost ring included) and the nd + 1 faces along the dimension.
static FaceTiles face_tiles(const GridD &g) {
  FaceTiles ft{};
  const long nd3[3] = {g.nx, g.ny, g.nz};
  const long wf0 = g.nx + 1;
  int fx = 62;
  while (fx > 1 && fx + 1 + (fx - 1 + wf0 - 1) / wf0 > 64) --fx;
  ft.fx = fx;
  for (int dim = 0; dim < 3; ++dim) {
    if (nd3[dim] <= 1) continue;
    const long n0 = dim == 0 ? g.ny : g.nx, n1 = dim == 2 ? g.ny : g.nz;
    const long w0 = n0 > 1 ? n0 + 2 : n0, w1 = n1 > 1 ? n1 + 2 : n1, wf = nd3[dim] + 1;
    if (dim == 0) ft.nb[0] = (wf * w0 * w1 + fx - 1) / fx;
    else ft.nb[dim] = ((w0 + FT_T - 1) / FT_T) * ((wf + FT_F - 1) / FT_F) * w1;
  }
  return ft;
}
static void launch_op_explicit(kz_sim *s, int dt_slot, bool with_bc) {
  const GridD &g = s->g;
  long nw = ring_count(g);
  {
    TimedLaunch tl(s, KZ_K_WAVESPEEDS);
    if (s->ph.opac.synchrotron_bridge) opt_launch_wavespeeds(g, s->d, s->ph, nblk(nw, 128), s->stream);
    else k_wavespeeds<<<nblk(nw, 128), 128, 0, s->stream>>>(g, s->d, s->ph);
  }
  {
    const FaceTiles ft = face_tiles(g);
    const unsigned nb = (unsigned)(ft.nb[0] + ft.nb[1] + ft.nb[2]);
    TimedLaunch tl(s, KZ_K_FACE_FLUX);
    const bool opts = s->ph.floors.b2rhofloorframe == 1 || s->ph.reduceorderatbh || s->ph.reduceorderafterfixup;
    if (opts) opt_launch_face_flux(g, s->d, s->ph, ft, nb, s->stream);
    else if (s->ph.radviscosity) k_face_flux<true, false><<<nb, 128, 0, s->stream>>>(g, s->d, s->ph, ft);
    else k_face_flux<false, false><<<nb, 128, 0, s->stream>>>(g, s->d, s->ph, ft);
  }
  long nc = (long)(g.nx + 1) * (g.ny + 1) * (g.nz + 1);
  {
    TimedLaunch tl(s, KZ_K_CT);
    k_emf<<<nblk(nc, 256), 256, 0, s->stream>>>(g, s->d);
  }
  {
    TimedLaunch tl(s, KZ_K_UPDATE);
    k_update<<<nblk(ndom(s), 64), 128, 0, s->stream>>>(g, s->d, s->ph, dt_slot);
  }
  launch_calc_u2p(s, with_bc);
}

// dynamo.zig:462-469 applyDynamo: scale height, setBc, mimicDynamo,
// calcU2p (its setBc only when `with_bc`, as for launch_calc_u2p).
static void launch_dynamo(kz_sim *s, int dt_slot, bool with_bc) {
  const GridD &g = s->g;
  {
    TimedLaunch tl(s, KZ_K_DYNAMO);
    k_scale_height<<<nblk(g.nx, 64), 64, 0, s->stream>>>(g, s->d);
  }
  launch_set_bc(s);
  {
    TimedLaunch tl(s, KZ_K_DYNAMO);
    long nw = ring_count(g);
    k_dynamo_a<<<nblk(nw, 64), 64, 0, s->stream>>>(g, s->d, s->ph, dt_slot, s->rhor, s->risco);
    k_dynamo_curl<<<nblk((long)g.nx * g.ny * g.nz, 128), 128, 0, s->stream>>>(g, s->d, s->ph);
  }
  launch_calc_u2p(s, with_bc);
}

// rijvisc.zig:266-311 calcRijViscTotal over the domain and its ring
static void launch_rijvisc(kz_sim *s, int dt_slot) {
  const GridD &g = s->g;
  TimedLaunch tl(s, KZ_K_VISC);
  long nw = ring_count(g);
  k_rijvisc<<<nblk(nw, 64), 64, 0, s->stream>>>(g, s->d, s->ph, dt_slot, s->rhor);
}

static ImpWork imp_work(const Dev &d, bool persist) {
  if (persist)
    return ImpWork{d.imp_uu00, d.imp_pp0, d.imp_aux, d.imp_slow, d.imp_list, d.imp_nslow, d.imp_ncnt, d.imp_rl,
                   d.imp_rc, d.imp_qpos};
  return ImpWork{d.imp_uu00, d.imp_pp0, d.imp_aux, d.imp_slow, d.imp_list, d.imp_nslow, d.imp_ncnt, nullptr, nullptr,
                 nullptr};
}

// persistent launches pay where the cells span many rounds of waves
// (192x192x64: 11.4 -> 10.6 ms of implicit work per step); at 384x360
// (two rounds) the merged launch is faster (0.85 against 0.91 ms)
static bool imp_persistent(long n, unsigned persist_blocks, int imp_launch) {
  return imp_launch == 2 || (imp_launch == 0 && n >= 8L * 64 * (long)persist_blocks);
}
// The implicit kernels over the n cells of a cell source (a grid, or the
// runs of a group); the bridge variants are in kz_options.hip.
extern "C++" template <class S>
static void launch_implicit(const S &cells, const ImpWork &w, const kz_physics &ph, long n, bool persist,
                            unsigned persist_blocks, hipStream_t st) {
  if (ph.opac.synchrotron_bridge) opt_launch_implicit_bridge(cells, w, ph, n, persist, persist_blocks, st);
  else launch_implicit_br<S, false>(cells, w, ph, n, persist, persist_blocks, st);
}
static void launch_radimp_fixup(kz_sim *s) {
  if (s->ph.has_opac && s->ph.do_fixups && s->ph.do_radimp_fixups)
    k_fixup<<<nblk(ndom(s), 256), 256, 0, s->stream>>>(s->g, s->d, s->ph, KZ_FLAG_RADIMP_FIXUP);
}
static void launch_op_implicit(kz_sim *s, int dt_slot) {
  if (!s->ph.has_opac) return;
  {
    TimedLaunch tl(s, KZ_K_IMPLICIT);
    const bool persist = imp_persistent(ndom(s), s->persist_blocks, s->imp_launch);
    launch_implicit(GridCells{s->g, s->d, dt_slot}, imp_work(s->d, persist), s->ph, ndom(s), persist,
                    s->persist_blocks, s->stream);
  }
  launch_radimp_fixup(s);
}

static void launch_do_correct(kz_sim *s) {
  if (!s->ph.correct_polaraxis) return;
  long n = (long)s->g.nx * s->ph.nccorrectpolar * 2 * s->g.nz;
  k_polar<<<nblk(n, 128), 128, 0, s->stream>>>(s->g, s->d, s->ph);
}

int kz_calc_u2p(kz_sim *s) {
  launch_calc_u2p(s, true);
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
  launch_op_explicit(s, SS_USER, true);
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

int kz_sim_set_bl_tables(kz_sim *s, const double *bl) {
  if (!bl) return KZ_E_INVALID_CONFIG;
  const long n = s->g.ncell;
  // runs without radiative viscosity or the dynamo allocate the sidecar here
  // (the scalar diagnostics read it)
  if (!s->d.bl)
    if (int e = dalloc(s, (void **)&s->d.bl, (size_t)n * BL_NC * sizeof(double))) return e;
  std::vector<double> planes((size_t)n * BL_NC);
  for (long c = 0; c < n; ++c)
    for (int k = 0; k < BL_NC; ++k) planes[(size_t)k * n + c] = bl[(size_t)c * BL_NC + k];
  HIPCHK(hipMemcpyAsync(s->d.bl, planes.data(), planes.size() * sizeof(double), hipMemcpyHostToDevice, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  s->has_bl = true;
  return 0;
}

int kz_calc_rij_visc_total(kz_sim *s, double global_dt) {
  if (!s->d.rijvisc || !s->has_bl) return KZ_E_INVALID_CONFIG;
  if (int e = put_dt(s, global_dt)) return e;
  launch_rijvisc(s, SS_USER);
  return check_err(s);
}

int kz_sim_get_rijvisc(kz_sim *s, double *out) {
  if (!s->d.rijvisc) return KZ_E_INVALID_CONFIG;
  const long n = s->g.ncell;
  std::vector<double> planes((size_t)n * 16);
  HIPCHK(hipMemcpyAsync(planes.data(), s->d.rijvisc, planes.size() * sizeof(double), hipMemcpyDeviceToHost,
                        s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  for (long c = 0; c < n; ++c)
    for (int k = 0; k < 16; ++k) out[(size_t)c * 16 + k] = planes[(size_t)k * n + c];
  return 0;
}

int kz_apply_dynamo(kz_sim *s, double dt) {
  if (!s->d.dyn_a || !s->has_bl) return KZ_E_INVALID_CONFIG;
  if (int e = put_dt(s, dt)) return e;
  launch_dynamo(s, SS_USER, true);
  return check_err(s);
}

// Enqueue one RK2IMEX step (sim.zig:1792-1899) on the sim's stream; no host
// synchronisation inside (launch_step).  saveTimesteps (cell_dt) is
// omitted: nothing on this path reads it.  The step is written in three parts
// around its two implicit passes, so that a group of runs can solve the
// implicit cells of all its runs in one launch between them.
static void launch_step_head(kz_sim *s, double forced, int use_forced) {
  k_begin_step<<<nblk(ndom(s), 256), 256, 0, s->stream>>>(s->g, s->d, forced, use_forced);  // + ut0 = u
  if (s->ph.radviscosity) launch_rijvisc(s, SS_DT);                                        // sim.zig:1819-1823
}
static void launch_step_mid(kz_sim *s) {
  const GridD &g = s->g;
  Dev &d = s->d;
  hipStream_t st = s->stream;
  const unsigned nb = nblk(ndom(s), 256);
  k_stage_implicit<<<nb, 256, 0, st>>>(g, d, d.drt1, d.ut0, d.ut1);
  launch_calc_u2p(s, false);
  launch_do_correct(s);
  launch_set_bc(s);
  launch_op_explicit(s, SS_DT, false);                          // F(U(1))
  if (s->ph.dynamo) launch_dynamo(s, SS_DT, false);             // sim.zig:1845-1849
  k_stage_explicit1<<<nb, 256, 0, st>>>(g, d);
  launch_calc_u2p(s, false);
  launch_do_correct(s);
}
static void launch_step_tail(kz_sim *s) {
  const GridD &g = s->g;
  Dev &d = s->d;
  hipStream_t st = s->stream;
  const unsigned nb = nblk(ndom(s), 256);
  k_stage_implicit<<<nb, 256, 0, st>>>(g, d, d.drt2, d.uforget, d.ut2);
  launch_calc_u2p(s, false);
  launch_do_correct(s);
  launch_set_bc(s);
  launch_op_explicit(s, SS_DT, false);                          // F(U(2))
  if (s->ph.dynamo) launch_dynamo(s, SS_DT, false);             // sim.zig:1872-1876
  k_stage_explicit2<<<nb, 256, 0, st>>>(g, d);
  launch_calc_u2p(s, false);
  launch_do_correct(s);
  launch_set_bc(s);
  k_end_step<<<nb, 256, 0, st>>>(g, d, s->ph);                  // t += dt, updateEntropy
}
static void launch_step(kz_sim *s, double forced, int use_forced) {
  launch_step_head(s, forced, use_forced);
  launch_op_implicit(s, SS_DTG);  // U(1)
  launch_step_mid(s);
  launch_op_implicit(s, SS_DTG);  // U(2)
  launch_step_tail(s);
}

int kz_step(kz_sim *s, const double *forced_dt) {
  if ((s->ph.radviscosity || s->ph.dynamo) && !s->has_bl) return KZ_E_INVALID_CONFIG;
  launch_step(s, forced_dt ? *forced_dt : 0.0, forced_dt != nullptr);
  s->nstep += 1;
  return check_err(s);
}

// enqueue `nsteps` CFL steps without host synchronisation (benchmarks)
int kz_step_many(kz_sim *s, int nsteps) {
  if ((s->ph.radviscosity || s->ph.dynamo) && !s->has_bl) return KZ_E_INVALID_CONFIG;
  for (int i = 0; i < nsteps; ++i) {
    launch_step(s, 0.0, 0);
    s->nstep += 1;
  }
  return 0;
}

int kz_sim_sync(kz_sim *s) { return check_err(s); }

// Capture nsteps CFL steps (exactly the launches kz_step_many enqueues) into
// a HIP graph; kz_sim_replay enqueues that graph.  Every step reads its dt
// and clears its counters and lists with kernels on the device, so a replay
// continues the run like kz_step_many(nsteps).  Event timing must be off.
int kz_sim_capture_steps(kz_sim *s, int nsteps) {
  if (nsteps <= 0 || s->timed != -1) return KZ_E_INVALID_CONFIG;
  if ((s->ph.radviscosity || s->ph.dynamo) && !s->has_bl) return KZ_E_INVALID_CONFIG;
  HIPCHK(hipStreamSynchronize(s->stream));
  if (s->graph_exec) HIPCHK(hipGraphExecDestroy(s->graph_exec));
  if (s->graph) HIPCHK(hipGraphDestroy(s->graph));
  s->graph_exec = nullptr;
  s->graph = nullptr;
  s->graph_steps = 0;
  HIPCHK(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
  for (int i = 0; i < nsteps; ++i) launch_step(s, 0.0, 0);
  HIPCHK(hipStreamEndCapture(s->stream, &s->graph));
  HIPCHK(hipGraphInstantiate(&s->graph_exec, s->graph, nullptr, nullptr, 0));
  s->graph_steps = nsteps;
  return 0;
}

int kz_sim_replay(kz_sim *s) {
  if (!s->graph_exec) return KZ_E_INVALID_CONFIG;
  HIPCHK(hipGraphLaunch(s->graph_exec, s->stream));
  s->nstep += (uint64_t)s->graph_steps;
  return 0;
}

int kz_sim_set_implicit_launch(kz_sim *s, int mode) {
  if (mode < 0 || mode > 2) return KZ_E_INVALID_CONFIG;
  s->imp_launch = mode;
  return 0;
}

int kz_sim_query(kz_sim *s) {
  hipError_t e = hipStreamQuery(s->stream);
  if (e == hipSuccess) return 0;
  if (e == hipErrorNotReady) return 1;
  return KZ_E_HIP;
}

int kz_sim_device_bytes(kz_sim *s, uint64_t *bytes) {
  *bytes = s->dev_bytes;
  return 0;
}

void *kz_sim_stream(kz_sim *s) { return (void *)s->stream; }

// ---------------------------------------------------------------------------
// Run groups: several resident runs stepped together, the implicit cells of
// all of them solved by one set of implicit launches per pass.  Each run's
// other kernels stay on its own stream; the group's stream joins the runs'
// streams before each implicit pass and forks them again after it.
struct kz_group {
  hipStream_t stream = nullptr;
  std::vector<kz_sim *> sims;
  std::vector<hipEvent_t> ev;  // one per run: its stream reached the join
  hipEvent_t ev_fork = nullptr;
  kz_physics ph{};
  RunCells *runs = nullptr;  // device table
  int cap_runs = 0;
  long n = 0, cap_n = 0;
  ImpWork w{};
  std::vector<void *> wallocs;
  unsigned persist_blocks = 1024;
  int imp_launch = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  int graph_steps = 0;
};

// the fields of kz_physics the implicit kernels read must agree across a group
static bool group_compatible(const kz_physics &a, const kz_physics &b) {
  return bits_of(a.gam) == bits_of(b.gam) && !memcmp(&a.rad, &b.rad, sizeof a.rad) &&
         !memcmp(&a.opac, &b.opac, sizeof a.opac) && !memcmp(&a.implicit, &b.implicit, sizeof a.implicit) &&
         a.has_opac == b.has_opac && a.correct_polaraxis == b.correct_polaraxis &&
         a.nccorrectpolar == b.nccorrectpolar;
}

static void group_free_graph(kz_group *G) {
  if (G->graph_exec) (void)hipGraphExecDestroy(G->graph_exec);
  if (G->graph) (void)hipGraphDestroy(G->graph);
  G->graph_exec = nullptr;
  G->graph = nullptr;
  G->graph_steps = 0;
}

static void group_free_work(kz_group *G) {
  for (void *p : G->wallocs) (void)hipFree(p);
  G->wallocs.clear();
  G->w = ImpWork{};
  G->cap_n = 0;
}

static int group_alloc(kz_group *G, void **p, size_t bytes) {
  if (hipMalloc(p, bytes) != hipSuccess) return KZ_E_HIP;
  G->wallocs.push_back(*p);
  if (hipMemsetAsync(*p, 0, bytes, G->stream) != hipSuccess) return KZ_E_HIP;
  return 0;
}

int kz_group_create(kz_group **out) {
  *out = nullptr;
  kz_group *G = new kz_group();
  int dev = 0, ncu = 0;
  if (hipStreamCreateWithFlags(&G->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&G->ev_fork, hipEventDisableTiming) != hipSuccess || hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) {
    kz_group_destroy(G);
    return KZ_E_HIP;
  }
  G->persist_blocks = (unsigned)(ncu > 0 ? KZ_PERSIST_PER_CU * ncu : 1024);
  *out = G;
  return 0;
}

void kz_group_destroy(kz_group *G) {
  if (!G) return;
  if (G->stream) (void)hipStreamSynchronize(G->stream);
  group_free_graph(G);
  group_free_work(G);
  if (G->runs) (void)hipFree(G->runs);
  for (hipEvent_t e : G->ev) (void)hipEventDestroy(e);
  if (G->ev_fork) (void)hipEventDestroy(G->ev_fork);
  if (G->stream) (void)hipStreamDestroy(G->stream);
  delete G;
}

int kz_group_set_runs(kz_group *G, kz_sim *const *sims, int nsims) {
  if (nsims < 0 || (nsims > 0 && !sims)) return KZ_E_INVALID_CONFIG;
  for (int i = 0; i < nsims; ++i) {
    if (!sims[i] || !group_compatible(sims[0]->ph, sims[i]->ph)) return KZ_E_INVALID_CONFIG;
    for (int j = 0; j < i; ++j)
      if (sims[j] == sims[i]) return KZ_E_INVALID_CONFIG;
  }
  long n = 0;
  std::vector<RunCells> tab((size_t)nsims);
  for (int i = 0; i < nsims; ++i) {
    tab[i] = RunCells{sims[i]->g, sims[i]->d, n};
    n += ndom(sims[i]);
  }
  if (n > INT32_MAX) return KZ_E_INVALID_CONFIG;
  // the previous membership's work (and any replay of its graph) ends first
  for (kz_sim *s : G->sims) HIPCHK(hipStreamSynchronize(s->stream));
  HIPCHK(hipStreamSynchronize(G->stream));
  group_free_graph(G);
  if (n > G->cap_n) {
    group_free_work(G);
    const long c = n + n / 4;
    int e = 0;
    first_err(e, group_alloc(G, (void **)&G->w.uu00, (size_t)c * NV * sizeof(double)));
    first_err(e, group_alloc(G, (void **)&G->w.pp0, (size_t)c * NV * sizeof(double)));
    first_err(e, group_alloc(G, (void **)&G->w.aux, (size_t)c * IMP_NAUX * sizeof(double)));
    first_err(e, group_alloc(G, (void **)&G->w.slow, (size_t)c * sizeof(int32_t)));
    first_err(e, group_alloc(G, (void **)&G->w.list, (size_t)2 * c * sizeof(int32_t)));
    first_err(e, group_alloc(G, (void **)&G->w.nslow, sizeof(unsigned long long)));
    first_err(e, group_alloc(G, (void **)&G->w.ncnt, 2 * sizeof(unsigned long long)));
    first_err(e, group_alloc(G, (void **)&G->w.rl, (size_t)2 * c * sizeof(int32_t)));
    first_err(e, group_alloc(G, (void **)&G->w.rc, 2 * sizeof(unsigned long long)));
    first_err(e, group_alloc(G, (void **)&G->w.qpos, 3 * sizeof(unsigned long long)));
    if (e) {
      group_free_work(G);
      return KZ_E_HIP;
    }
    G->cap_n = c;
  }
  if (nsims > G->cap_runs) {
    if (G->runs) HIPCHK(hipFree(G->runs));
    G->runs = nullptr;
    G->cap_runs = 0;
    HIPCHK(hipMalloc((void **)&G->runs, (size_t)nsims * sizeof(RunCells)));
    G->cap_runs = nsims;
  }
  while ((int)G->ev.size() < nsims) {
    hipEvent_t e;
    HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    G->ev.push_back(e);
  }
  if (nsims) {
    HIPCHK(hipMemcpyAsync(G->runs, tab.data(), (size_t)nsims * sizeof(RunCells), hipMemcpyHostToDevice, G->stream));
    G->ph = sims[0]->ph;
  }
  HIPCHK(hipStreamSynchronize(G->stream));
  G->sims.assign(sims, sims + nsims);
  G->n = n;
  return 0;
}

int kz_group_set_implicit_launch(kz_group *G, int mode) {
  if (mode < 0 || mode > 2) return KZ_E_INVALID_CONFIG;
  G->imp_launch = mode;
  group_free_graph(G);
  return 0;
}

// the group stream waits for every run's stream
static int group_join(kz_group *G) {
  for (size_t i = 0; i < G->sims.size(); ++i) {
    HIPCHK(hipEventRecord(G->ev[i], G->sims[i]->stream));
    HIPCHK(hipStreamWaitEvent(G->stream, G->ev[i], 0));
  }
  return 0;
}
// every run's stream waits for the group stream
static int group_fork(kz_group *G) {
  HIPCHK(hipEventRecord(G->ev_fork, G->stream));
  for (kz_sim *s : G->sims) HIPCHK(hipStreamWaitEvent(s->stream, G->ev_fork, 0));
  return 0;
}
static void group_implicit(kz_group *G, int dt_slot) {
  if (!G->ph.has_opac || G->n == 0) return;
  const bool persist = imp_persistent(G->n, G->persist_blocks, G->imp_launch);
  launch_implicit(MultiCells{G->runs, (int)G->sims.size(), G->n, dt_slot}, G->w, G->ph, G->n, persist,
                  G->persist_blocks, G->stream);
}
// nsteps CFL steps of every run: the kernel sequence of launch_step per run,
// with both implicit passes of each step taken by the group's launches
static int group_enqueue(kz_group *G, int nsteps) {
  if (int e = group_fork(G)) return e;
  for (int k = 0; k < nsteps; ++k) {
    for (kz_sim *s : G->sims) launch_step_head(s, 0.0, 0);
    for (int pass = 0; pass < 2; ++pass) {
      if (int e = group_join(G)) return e;
      group_implicit(G, SS_DTG);  // U(1), U(2)
      if (int e = group_fork(G)) return e;
      for (kz_sim *s : G->sims) {
        launch_radimp_fixup(s);
        if (pass == 0) launch_step_mid(s);
        else launch_step_tail(s);
      }
    }
  }
  if (int e = group_join(G)) return e;
  HIPCHK(hipGetLastError());
  return 0;
}
static int group_check(kz_group *G) {
  for (kz_sim *s : G->sims) {
    if ((s->ph.radviscosity || s->ph.dynamo) && !s->has_bl) return KZ_E_INVALID_CONFIG;
    if (s->timed != -1) return KZ_E_INVALID_CONFIG;  // event timing is per run, not per group
  }
  return 0;
}

int kz_group_step_many(kz_group *G, int nsteps) {
  if (nsteps < 0) return KZ_E_INVALID_CONFIG;
  if (int e = group_check(G)) return e;
  if (int e = group_enqueue(G, nsteps)) return e;
  for (kz_sim *s : G->sims) s->nstep += (uint64_t)nsteps;
  return 0;
}

int kz_group_capture_steps(kz_group *G, int nsteps) {
  if (nsteps <= 0) return KZ_E_INVALID_CONFIG;
  if (int e = group_check(G)) return e;
  for (kz_sim *s : G->sims) HIPCHK(hipStreamSynchronize(s->stream));
  HIPCHK(hipStreamSynchronize(G->stream));
  group_free_graph(G);
  HIPCHK(hipStreamBeginCapture(G->stream, hipStreamCaptureModeThreadLocal));
  const int e = group_enqueue(G, nsteps);
  hipGraph_t gr = nullptr;
  const hipError_t ee = hipStreamEndCapture(G->stream, &gr);
  if (e || ee != hipSuccess) {
    if (gr) (void)hipGraphDestroy(gr);
    return e ? e : KZ_E_HIP;
  }
  G->graph = gr;
  if (hipGraphInstantiate(&G->graph_exec, G->graph, nullptr, nullptr, 0) != hipSuccess) {
    group_free_graph(G);
    return KZ_E_HIP;
  }
  G->graph_steps = nsteps;
  return 0;
}

// the replay runs after the work already enqueued on the members' streams,
// and work enqueued on a member's stream afterwards runs after the replay
int kz_group_replay(kz_group *G) {
  if (!G->graph_exec) return KZ_E_INVALID_CONFIG;
  if (int e = group_join(G)) return e;
  HIPCHK(hipGraphLaunch(G->graph_exec, G->stream));
  if (int e = group_fork(G)) return e;
  for (kz_sim *s : G->sims) s->nstep += (uint64_t)G->graph_steps;
  return 0;
}

int kz_group_query(kz_group *G) {
  const hipError_t e = hipStreamQuery(G->stream);
  if (e == hipSuccess) return 0;
  if (e == hipErrorNotReady) return 1;
  return KZ_E_HIP;
}

int kz_group_sync(kz_group *G) {
  HIPCHK(hipStreamSynchronize(G->stream));
  for (kz_sim *s : G->sims)
    if (int e = check_err(s)) return e;
  return 0;
}

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
  HIPCHK(hipStreamSynchronize(s->stream));
  double tt = t;
  HIPCHK(hipMemcpyAsync(s->d.ss + SS_T, &tt, sizeof tt, hipMemcpyHostToDevice, s->stream));
  unsigned long long b[2] = {(unsigned long long)__builtin_bit_cast(long long, tstepdenmax),
                             (unsigned long long)__builtin_bit_cast(long long, tstepdenmin)};
  HIPCHK(hipMemcpyAsync(s->d.tsd, b, sizeof b, hipMemcpyHostToDevice, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  return 0;
}

int kz_sim_counters(kz_sim *s, kz_counters *c) {
  HIPCHK(hipStreamSynchronize(s->stream));
  unsigned long long cnt[8], tsd[2];
  double ss[SS_N];
  HIPCHK(hipMemcpyAsync(cnt, s->d.cnt, sizeof cnt, hipMemcpyDeviceToHost, s->stream));
  HIPCHK(hipMemcpyAsync(tsd, s->d.tsd, sizeof tsd, hipMemcpyDeviceToHost, s->stream));
  HIPCHK(hipMemcpyAsync(ss, s->d.ss, sizeof ss, hipMemcpyDeviceToHost, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
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

// ---- scalar diagnostics (koral/io/scalars.zig, io/dump.zig:358-411) -------
// The per-cell terms of the row's sums are formed on the device with the
// reference's expressions; the host adds them in the reference's loop order
// (each sum is a serial fold in the reference), so the row equals a serial
// evaluation bit for bit.
struct ScalArgs {
  long ix_h, ix_l, ix_s;
  double wedge;       // maxz - minz
  const double *dth;  // [nz][ny] BL theta widths of shell ix_l
  const double *dph;  // [nz][ny] BL phi widths of shell ix_l
  double *mass;       // [nd] rho dx dy dz sqrt(-g)
  double *ratio;      // [nd] pmag / ptot
  int32_t *diag;      // [nd] bit 0: a non-finite primitive, bit 1: HD fixup flag
  double *shell;      // [5][nz][ny] mdot, radlum, totallum, sigma, sigma dtheta^2 terms
};
// hydro.zig:87-112 calcTij (T^{mu nu} of the gas and field)
__device__ void calc_tij_full(const double *pp, const Geo &gm, double gam, double t[4][4]) {
  double ucon[4], ucov[4], bcon[4], bcov[4], bsq;
  ucon_ucov(pp, gm, ucon, ucov);
  bcon_bcov_bsq(pp[B1], pp[B2], pp[B3], ucon, ucov, gm, bcon, bcov, &bsq);
  const double rho = pp[RHO], u = pp[UU];
  const double p = (gam - 1.0) * u;
  const double w = rho + u + p;
  const double eta = w + bsq;
  const double ptot = p + 0.5 * bsq;
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) t[i][j] = eta * ucon[i] * ucon[j] + ptot * gm.G[i][j] - bcon[i] * bcon[j];
}
__global__ void __launch_bounds__(128) k_scalar_terms(GridD g, Dev d, kz_physics ph, ScalArgs a) {
  long t = TID;
  const long nd = (long)g.nx * g.ny * g.nz;
  if (t >= nd) return;
  const long ix = t % g.nx, iy = (t / g.nx) % g.ny, iz = t / ((long)g.nx * g.ny);
  const long n = g.ncell, c = cidx(g, ix, iy, iz);
  const long ns = (long)g.nz * g.ny, sh = iz * g.ny + iy;
  const bool twod = g.nz == 1;
  Geo gm;
  geo_c(g, d, ix, iy, iz, gm);
  double pp[NV];
  ld(d.p, n, c, pp);
  // totalMass (scalars.zig:79-101)
  const double phifac = g.nz > 1 ? 2.0 * KZ_PI / a.wedge : 1.0;
  const double dx = csize(g, ix, 0), dy = csize(g, iy, 1), dz = csize(g, iz, 2) * phifac;
  a.mass[t] = pp[RHO] * dx * dy * dz * gm.gdet;
  // maxPmagPtot (scalars.zig:275-313)
  {
    double ucon[4], ucov[4], bcon[4], bcov[4], bsq;
    ucon_ucov(pp, gm, ucon, ucov);
    bcon_bcov_bsq(pp[B1], pp[B2], pp[B3], ucon, ucov, gm, bcon, bcov, &bsq);
    const double pmag = bsq / 2.0;
    double ptot = (ph.gam - 1.0) * pp[UU];
    const double rtt = calc_ff_rtt(pp, gm);
    ptot += (-rtt) / 3.0;
    a.ratio[t] = pmag / ptot;
  }
  // collectDiag (dump.zig:358-380)
  int dg = 0;
#pragma unroll
  for (int iv = 0; iv < NV; ++iv)
    if (!kzm_isfinite(pp[iv])) dg |= 1;
  if (d.flags[KZ_FLAG_HD_FIXUP * n + c] != 0) dg |= 2;
  a.diag[t] = dg;
  // mdot (scalars.zig:108-133)
  if (ix == a.ix_h) {
    double v[4] = {0.0, pp[VX], pp[VY], pp[VZ]}, ucon[4];
    velr_to_vel4(v, gm, ucon);
    const double rhouconr = pp[RHO] * ucon[1];
    const double dzm = twod ? 2.0 * KZ_PI : csize(g, iz, 2) * (2.0 * KZ_PI / a.wedge);
    a.shell[0 * ns + sh] = gm.gdet * rhouconr * dy * dzm;
  }
  // lum, type 1, in BL (scalars.zig:141-185; frames.zig:300-310 transPallCoco)
  if (ix == a.ix_l) {
    Geo gbl;
    bl_geo(d, n, c, gbl);
    double J[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) J[k] = d.bl[(BL_J + k) * n + c];
    double q[NV];
#pragma unroll
    for (int iv = 0; iv < NV; ++iv) q[iv] = pp[iv];
    int e = trans_pmhd(q, gm, gbl, J);
    double ur[4], vr[4] = {0.0, q[FX], q[FY], q[FZ]};
    velr_to_vel4(vr, gm, ur);
    tvec(J, ur, ur);
    if (!e) e = vel4_to_velr(ur, gbl, true, ur);
    if (e) {
      atomicCAS(d.err, 0, e);
      return;
    }
    q[FX] = ur[1];
    q[FY] = ur[2];
    q[FZ] = ur[3];
    const double dthbl = a.dth[sh];
    const double dphbl = twod ? 2.0 * KZ_PI : a.dph[sh] * (2.0 * KZ_PI / a.wedge);
    double v[4] = {0.0, q[VX], q[VY], q[VZ]}, ucon[4];
    velr_to_vel4(v, gbl, ucon);
    const double rhour = q[RHO] * ucon[1];
    double tu[4][4], tl[4][4];
    calc_tij_full(q, gbl, ph.gam, tu);
    lower_second(tu, gbl, tl);
    const double trt = tl[1][0];
    double ru[4][4], rl[4][4];
    calc_rij(q, gbl, ru);
    lower_second(ru, gbl, rl);
    double rrt = -rl[1][0];
    if (rrt < 0) rrt = 0;
    a.shell[1 * ns + sh] = gbl.gdet * rrt * dthbl * dphbl;
    a.shell[2 * ns + sh] = gbl.gdet * (rhour + trt + rrt) * dthbl * dphbl;
  }
  // scaleHeightPartsAtIx (dynamo.zig:239-262)
  if (ix == a.ix_s) {
    const double th = d.bl[BL_TH * n + c];
    const double dthe = KZ_PI / 2.0 - th;
    a.shell[3 * ns + sh] = pp[RHO] * gm.gdet;
    a.shell[4 * ns + sh] = pp[RHO] * gm.gdet * dthe * dthe;
  }
}

int kz_sim_scalars(kz_sim *s, double dt, long ix_h, long ix_l, long ix_s, const double *dth_bl, const double *dph_bl,
                   kz_scalar_row *row) {
  const GridD &g = s->g;
  if (!s->has_bl || !row || !dth_bl || !dph_bl) return KZ_E_INVALID_CONFIG;
  if (ix_h < 0 || ix_h >= g.nx || ix_l < 0 || ix_l >= g.nx || ix_s < 0 || ix_s >= g.nx) return KZ_E_INVALID_CONFIG;
  const long nd = (long)g.nx * g.ny * g.nz, ns = (long)g.nz * g.ny;
  // layout: mass[nd] ratio[nd] shell[5 ns] dth[ns] dph[ns] diag[nd] (int32)
  const size_t nbytes = (size_t)(2 * nd + 7 * ns) * sizeof(double) + (size_t)nd * sizeof(int32_t);
  if (!s->scal_buf) {
    void *p = nullptr;
    if (int e = dalloc(s, &p, nbytes)) return e;
    s->scal_buf = (double *)p;
  }
  ScalArgs a;
  a.ix_h = ix_h;
  a.ix_l = ix_l;
  a.ix_s = ix_s;
  a.wedge = s->wedge;
  a.mass = s->scal_buf;
  a.ratio = a.mass + nd;
  a.shell = a.ratio + nd;
  double *ddth = a.shell + 5 * ns, *ddph = ddth + ns;
  a.dth = ddth;
  a.dph = ddph;
  a.diag = (int32_t *)(ddph + ns);
  HIPCHK(hipMemcpyAsync(ddth, dth_bl, ns * sizeof(double), hipMemcpyHostToDevice, s->stream));
  HIPCHK(hipMemcpyAsync(ddph, dph_bl, ns * sizeof(double), hipMemcpyHostToDevice, s->stream));
  k_scalar_terms<<<nblk(nd, 128), 128, 0, s->stream>>>(g, s->d, s->ph, a);
  if (int e = check_err(s)) return e;
  std::vector<double> mass(nd), ratio(nd), shell(5 * ns);
  std::vector<int32_t> diag(nd);
  HIPCHK(hipMemcpyAsync(mass.data(), a.mass, nd * sizeof(double), hipMemcpyDeviceToHost, s->stream));
  HIPCHK(hipMemcpyAsync(ratio.data(), a.ratio, nd * sizeof(double), hipMemcpyDeviceToHost, s->stream));
  HIPCHK(hipMemcpyAsync(shell.data(), a.shell, 5 * ns * sizeof(double), hipMemcpyDeviceToHost, s->stream));
  HIPCHK(hipMemcpyAsync(diag.data(), a.diag, nd * sizeof(int32_t), hipMemcpyDeviceToHost, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  // the reference's serial folds: mass and the maximum over (iz, iy, ix),
  // the shell sums over (iz, iy), the scale-height parts over (iy, iz)
  double m = 0, maxb = 0;
  uint64_t nhd = 0, nnan = 0;
  for (long t = 0; t < nd; ++t) {
    m += mass[t];
    if (ratio[t] > maxb) maxb = ratio[t];
    if (diag[t] & 1) nnan += 1;
    if (diag[t] & 2) nhd += 1;
  }
  double md = 0, radlum = 0, totallum = 0;
  for (long sh = 0; sh < ns; ++sh) {
    md += shell[sh];
    radlum += shell[ns + sh];
    totallum += shell[2 * ns + sh];
  }
  double sig = 0, sth = 0;
  for (long iy = 0; iy < g.ny; ++iy)
    for (long iz = 0; iz < g.nz; ++iz) {
      sig += shell[3 * ns + iz * g.ny + iy];
      sth += shell[4 * ns + iz * g.ny + iy];
    }
  kz_counters c;
  if (int e = kz_sim_counters(s, &c)) return e;
  row->t = c.t;
  row->dt = dt;
  row->nstep = c.nstep;
  row->mass = m;
  row->mdot = -md;
  row->radlum = radlum;
  row->totallum = totallum;
  row->scaleheight = ix_s > 0 ? std::sqrt(sth / sig) : sth;
  row->max_pmag_ptot = maxb;
  row->n_hd_fixup = nhd;
  row->n_radimp_fail = c.n_radimp_failures;
  row->n_nan = nnan;
  row->n_floorguard = c.n_guard_recoveries;
  return 0;
}

// ---- KDMP v2 checkpoints (koral/io/dump.zig:19-160) ----------------------
// The header and body are little-endian; both hosts this library targets
// (x86-64, aarch64) are, so fields are copied as they lie in memory.
int kz_kdmp_parse_header(const uint8_t *buf, uint64_t len, kz_dump_header *h) {
  if (len < KZ_KDMP_HEADER_BYTES) return KZ_E_TRUNCATED;
  if (memcmp(buf, "KDMP", 4) != 0) return KZ_E_BAD_MAGIC;
  uint32_t ver;
  memcpy(&ver, buf + 4, 4);
  if (ver != 2) return KZ_E_BAD_VERSION;
  memcpy(&h->nx, buf + 8, 4);
  memcpy(&h->ny, buf + 12, 4);
  memcpy(&h->nz, buf + 16, 4);
  memcpy(&h->nv, buf + 20, 4);
  memcpy(&h->t, buf + 24, 8);
  memcpy(&h->nstep, buf + 32, 8);
  memcpy(&h->out_idx, buf + 40, 4);
  h->_pad = 0;
  return 0;
}

int kz_sim_kdmp_size(kz_sim *s, uint64_t *bytes) {
  *bytes = KZ_KDMP_HEADER_BYTES + (uint64_t)s->g.nx * s->g.ny * s->g.nz * NV * 8;
  return 0;
}

// serializePrimDump (dump.zig:101-135): header, then the domain primitives
// p[iz][iy][ix][iv]; t and nstep are the run's current clock.
int kz_sim_write_kdmp(kz_sim *s, uint32_t out_idx, uint8_t *buf, uint64_t cap, uint64_t *written) {
  uint64_t need;
  kz_sim_kdmp_size(s, &need);
  if (cap < need) return KZ_E_TRUNCATED;
  kz_counters c;
  if (int e = kz_sim_counters(s, &c)) return e;
  std::vector<double> p((size_t)s->g.ncell * NV);
  if (int e = kz_sim_get_state(s, p.data(), nullptr, nullptr)) return e;
  uint8_t *w = buf;
  const uint32_t ver = 2, nv = NV, dims[3] = {(uint32_t)s->g.nx, (uint32_t)s->g.ny, (uint32_t)s->g.nz};
  memcpy(w, "KDMP", 4);
  memcpy(w + 4, &ver, 4);
  memcpy(w + 8, dims, 12);
  memcpy(w + 20, &nv, 4);
  memcpy(w + 24, &c.t, 8);
  memcpy(w + 32, &c.nstep, 8);
  memcpy(w + 40, &out_idx, 4);
  w += KZ_KDMP_HEADER_BYTES;
  for (long iz = 0; iz < s->g.nz; ++iz)
    for (long iy = 0; iy < s->g.ny; ++iy)
      for (long ix = 0; ix < s->g.nx; ++ix) {
        const long ci = (ix + s->g.ngx) + (iy + s->g.ngy) * (long)s->g.sx + (iz + s->g.ngz) * (long)s->g.sy * s->g.sx;
        memcpy(w, p.data() + ci * NV, NV * 8);
        w += NV * 8;
      }
  *written = (uint64_t)(w - buf);
  return 0;
}

// loadPrimDump (dump.zig:137-160) + the driver's clock restore
// (problems/puffy/main.zig:283-298): every domain cell gets its primitives
// and u = p2u(p) (Sim.initCell); ghosts are left for kz_sim_finish_init.
int kz_sim_load_kdmp(kz_sim *s, const uint8_t *buf, uint64_t len, kz_dump_header *h) {
  kz_dump_header hh;
  if (int e = kz_kdmp_parse_header(buf, len, &hh)) return e;
  if (hh.nx != (uint32_t)s->g.nx || hh.ny != (uint32_t)s->g.ny || hh.nz != (uint32_t)s->g.nz || hh.nv != NV)
    return KZ_E_DIM_MISMATCH;
  uint64_t need;
  kz_sim_kdmp_size(s, &need);
  if (len < need) return KZ_E_TRUNCATED;
  std::vector<double> p((size_t)s->g.ncell * NV);
  if (int e = kz_sim_get_state(s, p.data(), nullptr, nullptr)) return e;
  const uint8_t *r = buf + KZ_KDMP_HEADER_BYTES;
  for (long iz = 0; iz < s->g.nz; ++iz)
    for (long iy = 0; iy < s->g.ny; ++iy)
      for (long ix = 0; ix < s->g.nx; ++ix) {
        const long ci = (ix + s->g.ngx) + (iy + s->g.ngy) * (long)s->g.sx + (iz + s->g.ngz) * (long)s->g.sy * s->g.sx;
        memcpy(p.data() + ci * NV, r, NV * 8);
        r += NV * 8;
      }
  if (int e = kz_sim_set_state(s, p.data(), nullptr)) return e;
  if (int e = kz_sim_p2u_domain(s)) return e;
  double t = hh.t;
  HIPCHK(hipMemcpyAsync(s->d.ss + SS_T, &t, sizeof t, hipMemcpyHostToDevice, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  s->nstep = hh.nstep;
  if (h) *h = hh;
  return 0;
}

// Sim.finishInit (koral/sim.zig:686-690): ghosts, then the startup dt guess.
int kz_sim_finish_init(kz_sim *s) {
  if (int e = kz_sim_set_bc(s, 1)) return e;
  return kz_sim_init_timestep_guess(s);
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

// ---------------------------------------------------------------------------
// Independent-cell implicit batches (solveImplicitLab over n records, device
// resident): the split, compacted solve of opImplicit over BatchCells.
struct kz_implicit_batch {
  long n = 0;
  kz_physics ph;
  hipStream_t stream = nullptr;
  double *gp = nullptr, *dt = nullptr, *uu = nullptr, *pp = nullptr, *pp_in = nullptr, *stage = nullptr;
  kz_implicit_result *res = nullptr;
  unsigned long long *cnt = nullptr;
  ImpWork w{};
  std::vector<void *> allocs;
  hipEvent_t ev[2] = {nullptr, nullptr};
  bool has_input = false;
  unsigned persist_blocks = 1024;
};

static int balloc(kz_implicit_batch *b, void **p, size_t bytes) {
  if (hipMalloc(p, bytes) != hipSuccess) return KZ_E_OOM;
  b->allocs.push_back(*p);
  return hipMemset(*p, 0, bytes) == hipSuccess ? 0 : KZ_E_HIP;
}

__global__ void k_batch_reset(long n, const double *pp_in, double *pp, double *uu, kz_implicit_result *res,
                              unsigned long long *cnt) {
  long t = TID;
  if (t < 3) cnt[t] = 0;
  if (t >= n) return;
#pragma unroll
  for (int iv = 0; iv < NV; ++iv) {
    pp[iv * n + t] = pp_in[iv * n + t];
    uu[iv * n + t] = 0.0;
  }
  res[t] = kz_implicit_result{0, -1, 0, WP_RAD, WE_ENERGY, WF_LAB};
}

extern "C" {

kz_implicit_batch *kz_implicit_batch_create(int64_t n, const kz_physics *ph, int *err) {
  *err = 0;
  if (n <= 0 || n > INT32_MAX) { *err = KZ_E_INVALID_CONFIG; return nullptr; }
  if (!ph->has_opac) { *err = KZ_E_INVALID_CONFIG; return nullptr; }
  kz_implicit_batch *b = new kz_implicit_batch();
  b->n = n;
  b->ph = *ph;
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) {
    *err = KZ_E_HIP;
    delete b;
    return nullptr;
  }
  int e = 0;
  first_err(e, balloc(b, (void **)&b->gp, (size_t)n * (GEO_NC + 1) * sizeof(double)));
  first_err(e, balloc(b, (void **)&b->dt, (size_t)n * sizeof(double)));
  for (double **f : {&b->uu, &b->pp, &b->pp_in, &b->stage}) first_err(e, balloc(b, (void **)f, (size_t)n * NV * sizeof(double)));
  first_err(e, balloc(b, (void **)&b->res, (size_t)n * sizeof(kz_implicit_result)));
  first_err(e, balloc(b, (void **)&b->cnt, 4 * sizeof(unsigned long long)));
  first_err(e, balloc(b, (void **)&b->w.uu00, (size_t)n * NV * sizeof(double)));
  first_err(e, balloc(b, (void **)&b->w.pp0, (size_t)n * NV * sizeof(double)));
  first_err(e, balloc(b, (void **)&b->w.aux, (size_t)n * IMP_NAUX * sizeof(double)));
  first_err(e, balloc(b, (void **)&b->w.slow, (size_t)n * sizeof(int32_t)));
  first_err(e, balloc(b, (void **)&b->w.list, (size_t)2 * n * sizeof(int32_t)));
  first_err(e, balloc(b, (void **)&b->w.nslow, sizeof(unsigned long long)));
  first_err(e, balloc(b, (void **)&b->w.ncnt, 2 * sizeof(unsigned long long)));
  first_err(e, balloc(b, (void **)&b->w.rl, (size_t)2 * n * sizeof(int32_t)));
  first_err(e, balloc(b, (void **)&b->w.rc, 2 * sizeof(unsigned long long)));
  first_err(e, balloc(b, (void **)&b->w.qpos, 3 * sizeof(unsigned long long)));
  {
    int dev = 0, ncu = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
      first_err(e, KZ_E_HIP);
    b->persist_blocks = (unsigned)(ncu > 0 ? KZ_PERSIST_PER_CU * ncu : 1024);  // one wave per SIMD
  }
  for (auto &x : b->ev) first_err(e, hip_err(hipEventCreate(&x)));
  if (e) {
    *err = e < 0 ? e : KZ_E_HIP;
    kz_implicit_batch_destroy(b);
    return nullptr;
  }
  return b;
}

void kz_implicit_batch_destroy(kz_implicit_batch *b) {
  if (!b) return;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  for (void *p : b->allocs) (void)hipFree(p);
  for (auto x : b->ev)
    if (x) (void)hipEventDestroy(x);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
}

int kz_implicit_batch_set(kz_implicit_batch *b, const double *pp, const kz_geom *g, const double *dt) {
  const long n = b->n;
  std::vector<double> planes((size_t)n * (GEO_NC + 1));
  for (long i = 0; i < n; ++i) {
    const kz_geom &k = g[i];
    for (int a = 0; a < 4; ++a)
      for (int c = a; c < 4; ++c) {
        if (bits_of(k.gg[a][c]) != bits_of(k.gg[c][a])) return KZ_E_INVALID_CONFIG;
        planes[(size_t)tri(a, c) * n + i] = k.gg[a][c];
      }
    for (int a = 0; a < 4; ++a)
      for (int c = 0; c < 4; ++c) planes[(size_t)(10 + a * 4 + c) * n + i] = k.GG[a][c];
    planes[(size_t)26 * n + i] = k.gdet;
    planes[(size_t)27 * n + i] = k.gttpert;
    planes[(size_t)28 * n + i] = k.alpha;
    planes[(size_t)GEO_NC * n + i] = k.xx[1];
  }
  HIPCHK(hipMemcpyAsync(b->gp, planes.data(), planes.size() * sizeof(double), hipMemcpyHostToDevice, b->stream));
  HIPCHK(hipMemcpyAsync(b->dt, dt, (size_t)n * sizeof(double), hipMemcpyHostToDevice, b->stream));
  HIPCHK(hipMemcpyAsync(b->stage, pp, (size_t)n * NV * sizeof(double), hipMemcpyHostToDevice, b->stream));
  k_aos2soa<<<nblk(n * NV, 256), 256, 0, b->stream>>>(b->stage, b->pp_in, n, NV);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(b->stream));
  b->has_input = true;
  return 0;
}

int kz_implicit_batch_solve(kz_implicit_batch *b, double *ms) {
  if (!b->has_input) return KZ_E_INVALID_CONFIG;
  const long n = b->n;
  k_batch_reset<<<nblk(n < 3 ? 3 : n, 256), 256, 0, b->stream>>>(n, b->pp_in, b->pp, b->uu, b->res, b->cnt);
  const BatchCells cells{n, b->gp, b->dt, b->uu, b->pp, b->res, b->cnt};
  HIPCHK(hipEventRecord(b->ev[0], b->stream));
  launch_implicit(cells, b->w, b->ph, n, true, b->persist_blocks, b->stream);
  HIPCHK(hipEventRecord(b->ev[1], b->stream));
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(b->stream));
  if (ms) {
    float f = 0;
    HIPCHK(hipEventElapsedTime(&f, b->ev[0], b->ev[1]));
    *ms = f;
  }
  return 0;
}

int kz_implicit_batch_get(kz_implicit_batch *b, double *uu, double *pp, kz_implicit_result *r, uint64_t *counts) {
  const long n = b->n;
  const size_t fb = (size_t)n * NV * sizeof(double);
  if (uu) {
    k_soa2aos<<<nblk(n * NV, 256), 256, 0, b->stream>>>(b->uu, b->stage, n, NV);
    HIPCHK(hipMemcpyAsync(uu, b->stage, fb, hipMemcpyDeviceToHost, b->stream));
    HIPCHK(hipStreamSynchronize(b->stream));
  }
  if (pp) {
    k_soa2aos<<<nblk(n * NV, 256), 256, 0, b->stream>>>(b->pp, b->stage, n, NV);
    HIPCHK(hipMemcpyAsync(pp, b->stage, fb, hipMemcpyDeviceToHost, b->stream));
    HIPCHK(hipStreamSynchronize(b->stream));
  }
  if (r) HIPCHK(hipMemcpyAsync(r, b->res, (size_t)n * sizeof(kz_implicit_result), hipMemcpyDeviceToHost, b->stream));
  if (counts) HIPCHK(hipMemcpyAsync(counts, b->cnt, 3 * sizeof(uint64_t), hipMemcpyDeviceToHost, b->stream));
  HIPCHK(hipStreamSynchronize(b->stream));
  return 0;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// The implicit batch solve without device allocation (kz_implicit_solve_noalloc):
// the same kernels as kz_implicit_batch_solve over the caller's device
// arrays, with every working buffer carved out of one static device array.
// Per cell: the geometry planes (GEO_NC + 1), the SoA uu and pp (2 x 13),
// the Newton reference state and start (2 x 13), the reference-state
// scalars (IMP_NAUX) as doubles, and five int32 list entries.
#define NA_DBL_PER_CELL (GEO_NC + 1 + 4 * NV + IMP_NAUX)
#define NA_INT_PER_CELL 5
#define NA_NCNT 12  // cnt[4] nslow ncnt[2] rc[2] qpos[3]
static constexpr size_t NA_BYTES =
    ((size_t)KZ_NOALLOC_MAX_CELLS * (NA_DBL_PER_CELL * 8 + NA_INT_PER_CELL * 4) + NA_NCNT * 8 + 255) / 256 * 256;
static __device__ __align__(256) unsigned char scratch[NA_BYTES];

// AoS caller inputs -> the SoA planes the solve reads; outputs cleared
__global__ void k_na_pack(long n, const double *pp_aos, const kz_geom *g, double *gp, double *pp, double *uu,
                          kz_implicit_result *res, unsigned long long *ctr) {
  long t = TID;
  if (t < NA_NCNT) ctr[t] = 0;
  if (t >= n) return;
  const kz_geom &k = g[t];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int c = a; c < 4; ++c) gp[tri(a, c) * n + t] = k.gg[a][c];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int c = 0; c < 4; ++c) gp[(10 + a * 4 + c) * n + t] = k.GG[a][c];
  gp[26 * n + t] = k.gdet;
  gp[27 * n + t] = k.gttpert;
  gp[28 * n + t] = k.alpha;
  gp[GEO_NC * n + t] = k.xx[1];
#pragma unroll
  for (int iv = 0; iv < NV; ++iv) {
    pp[iv * n + t] = pp_aos[t * NV + iv];
    uu[iv * n + t] = 0.0;
  }
  res[t] = kz_implicit_result{0, -1, 0, WP_RAD, WE_ENERGY, WF_LAB};
}
__global__ void k_na_unpack(long n, const double *uu, const double *pp, double *uu_aos, double *pp_aos,
                            const unsigned long long *cnt, uint64_t *counts) {
  long t = TID;
  if (counts && t < 3) counts[t] = cnt[t];
  if (t >= n * NV) return;
  long c = t / NV;
  int iv = t % NV;
  if (uu_aos) uu_aos[t] = uu[iv * n + c];
  if (pp_aos) pp_aos[t] = pp[iv * n + c];
}

namespace {
struct NaDevice {
  unsigned char *base = nullptr;  // this device's copy of `scratch`
  unsigned blocks = 0;            // persistent launch width: one wave per SIMD
};
std::mutex na_mu;
NaDevice na_dev[64];
}  // namespace

extern "C" int kz_implicit_solve_noalloc(int64_t n, const double *pp_in, const kz_geom *g, const double *dt,
                                         double *uu, double *pp, kz_implicit_result *res, uint64_t *counts,
                                         const kz_physics *phys, void *stream) {
  if (n < 0 || n > KZ_NOALLOC_MAX_CELLS) {
    fprintf(stderr, "kz_implicit_solve_noalloc: %lld cells, at most %d fit the static scratch\n", (long long)n,
            KZ_NOALLOC_MAX_CELLS);
    return KZ_E_INVALID_CONFIG;
  }
  if (!phys || !phys->has_opac) return KZ_E_INVALID_CONFIG;
  if (n == 0) return KZ_OK;
  if (!pp_in || !g || !dt || !res) return KZ_E_INVALID_CONFIG;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return KZ_E_HIP;
  NaDevice nd;
  {
    std::lock_guard<std::mutex> lk(na_mu);
    if (!na_dev[dev].base) {
      void *ptr = nullptr;
      if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) return KZ_E_HIP;
      int ncu = 0;
      if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return KZ_E_HIP;
      na_dev[dev].blocks = (unsigned)(ncu > 0 ? KZ_PERSIST_PER_CU * ncu : 1024);
      na_dev[dev].base = (unsigned char *)ptr;
    }
    nd = na_dev[dev];
  }
  hipStream_t st = (hipStream_t)stream;
  // carve the scratch: doubles first, then the int32 lists, then counters
  double *dp = (double *)nd.base;
  double *gp = dp, *suu = gp + (size_t)(GEO_NC + 1) * n, *spp = suu + (size_t)NV * n;
  ImpWork w{};
  w.uu00 = spp + (size_t)NV * n;
  w.pp0 = w.uu00 + (size_t)NV * n;
  w.aux = w.pp0 + (size_t)NV * n;
  int32_t *ip = (int32_t *)(w.aux + (size_t)IMP_NAUX * n);
  w.slow = ip;
  w.list = ip + n;
  w.rl = ip + 3 * n;
  unsigned long long *ctr = (unsigned long long *)(((uintptr_t)(ip + 5 * n) + 7) & ~(uintptr_t)7);
  unsigned long long *cnt = ctr;
  w.nslow = ctr + 4;
  w.ncnt = ctr + 5;
  w.rc = ctr + 7;
  w.qpos = ctr + 9;
  kz_physics ph = *phys;
  k_na_pack<<<nblk(n < NA_NCNT ? NA_NCNT : n, 256), 256, 0, st>>>(n, pp_in, g, gp, spp, suu, res, ctr);
  const BatchCells cells{n, gp, dt, suu, spp, res, cnt};
  launch_implicit(cells, w, ph, n, true, nd.blocks, st);
  k_na_unpack<<<nblk(n * NV, 256), 256, 0, st>>>(n, suu, spp, uu, pp, cnt, counts);
  return hipGetLastError() == hipSuccess ? KZ_OK : KZ_E_HIP;
}
