// kz_sim_kernels.h — the step kernels that the reference's non-default
// options change, as templates over the option: the wavespeeds (the
// synchrotron bridge in the optical depth), the u2p pass (the b^2/rho floor
// frame), the face flux (reconstruction order per cell, floor frame) and the
// implicit ladder (the synchrotron bridge).  kz_step.hip
// instantiates the default of each, kz_options.hip the option variants, so
// that the default code object carries none of the variants' code.
#ifndef KZ_SIM_KERNELS_H
#define KZ_SIM_KERNELS_H

#include "kz_sim_dev.h"

// wavespeeds + CFL denominators (sim.zig:747-886)
__device__ __forceinline__ long long dbits(double x) { return __double_as_longlong(x); }
// one thread per cell of the domain and its one-cell ring; BR fixes the
// synchrotron bridge of the opacities on or off
template <bool BR>
__device__ __forceinline__ void wavespeeds_pass(const GridD &g, const Dev &d, const kz_physics &ph) {
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
    kz_opac_params op = ph.opac;
    op.synchrotron_bridge = BR;
    double chi = calc_chi_slim(pp, gm, ph.gam, op);
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

// u2p pass (sim.zig:940-1000), one thread per domain cell; cells flagged for
// the u2p fixups (sim.zig:1001-1003) are listed for k_u2p_fixups.  ZAMO
// floors in the ZAMO frame (invert.zig:649-684).
template <bool ZAMO>
__device__ __forceinline__ void u2p_pass(const GridD &g, const Dev &d, const kz_physics &ph) {
  long t = TID;
  long nd = (long)g.nx * g.ny * g.nz;
  const bool fix_hd = ph.do_fixups, fix_rad = ph.do_fixups && ph.do_u2prad_fixups;
  if (t < nd) {
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
    } else {
      int ent;
      int ret = u2p_mhd(uu, pp, gm, ph.gam, ph.floors, &ent);
      if (ent) d.flags[KZ_FLAG_ENTROPY * n + c] = 1;
      int corr;
      u2p_rad(uu, pp, gm, ph.rad, &corr);
      int guard = 0;
      check_floors_mhd<ZAMO>(pp, gm, ph.gam, ph.floors, &guard);
      check_floors_rad(pp, gm, ph.rad);
      st(d.p, n, c, pp);
      int fh = ret < -1 ? 1 : 0, fr = corr ? -1 : 0;
      d.flags[KZ_FLAG_HD_FIXUP * n + c] = fh;
      d.flags[KZ_FLAG_RAD_FIXUP * n + c] = fr;
      if (guard) atomicAdd(&d.cnt[3], (unsigned long long)guard);
      if ((fix_hd && fh) || (fix_rad && fr)) {
        unsigned k = atomicAdd(&d.sync[SY_NFIX], 1u);
        d.fix_list[k] = (int32_t)t;
      }
    }
  }
}
// ---------------------------------------------------------------------------
// fused face flux: PPM of the two adjacent cells, face floors, fFluxPrime,
// p2u, LAXF (sim.zig:1284-1503) for faces i in [0, n] of each active
// dimension, cross indices over [-1, n_c + 1) in the other active ones.

// rijvisc.zig:313-353 faceAvg + radvisc.zig:201-244 addRadViscFlux: row
// dim+1 of the face-averaged viscous R^i_j (rv), velocity-capped, added to
// the radiation rows of one face state's flux; uu is that state's p2u.  Of
// the reference's three capped velocities only the face-normal one is used.
template <int DIM>
__device__ __forceinline__ void add_visc_flux(const double rv[4], const double *uu, const Geo &gm, const GridD &g,
                                              const kz_physics &ph, double *ff) {
  const bool y_active = g.ny > 1, z_active = g.nz > 1;
  const double gdetu = gm.gdet;
  double vel = 0;
#pragma unroll
  for (int ii = 1; ii < 4; ++ii) {
    if (ii == 2 && !y_active) continue;
    if (ii == 3 && !z_active) continue;
    if (kzm_fabs(uu[EE + ii]) < 1.0e-10 * kzm_fabs(uu[EE])) continue;
    vel = rv[ii] / (uu[EE + ii] / gdetu) * kzm_sqrt(gm.g[DIM + 1][DIM + 1]);
  }
  const double maxv = kzm_fabs(vel);
  double dampfac = 1.0;
  if (maxv > ph.radvisc.maxvel) dampfac = ph.radvisc.maxvel / maxv;
#pragma unroll
  for (int nu = 0; nu < 4; ++nu) ff[EE + nu] += gdetu * dampfac * rv[nu];
}
// Face fluxes by tiles.  A block of 128 threads owns up to 64 faces of one
// dimension whose adjacent cells number at most 64:
//   x faces: a run of FaceTiles::fx consecutive faces of the x-fastest face
//   order (a run may continue on the next row);
//   y and z faces: FT_T consecutive x columns by FT_F consecutive faces
//   along the dimension.
// Phase 1: the block reconstructs each of its cells once, both face values
// of every slot, into LDS (the left state of face i is cell i-1's right
// value, the right state cell i's left one); the two waves share the slots.
// Phase 2: one thread per face side takes its state from LDS, applies the
// face floors and forms its flux fFluxPrime and conserved vector p2u (with
// the viscous flux when enabled).  Phase 3: the two sides of a face,
// neighbouring lanes, exchange flux and conserved vector and each writes the
// LAXF flux of half the slots.  Each side is exactly the reference's
// computation for that face state, so the fluxes and the floor guard count
// are the reference's.
#ifndef KZ_FT_T
#define KZ_FT_T 8
#endif
constexpr int FT_T = KZ_FT_T, FT_F = 64 / FT_T - 1;
struct FaceTiles {
  long nb[3];  // blocks of each dimension, launched in the order x, y, z
  int fx;      // x faces per block
};
struct FaceShared {
  double st[2][NV][64];  // [0] left, [1] right face value of each slot of each cell
};
// The faces and cells of block b of dimension DIM, numbered locally: face f
// lies between cells left(f) and right(f).  Coordinates are (ix, iy, iz)
// with the face index in component DIM.
template <int DIM>
struct FaceTile {
  static constexpr int c0d = DIM == 0 ? 1 : 0, c1d = DIM == 2 ? 1 : 2;
  long wf, w0, l0, l1;
  long t0, r0, e0;  // x runs: first face, its row, its left cell's place in rows of wf + 1 cells
  long x0, i0, c1;  // y/z tiles: first column, first face along DIM, the other cross index
  int nface, ncell;
  __device__ FaceTile(const GridD &g, const FaceTiles &ft, long b) {
    const long n0 = c0d == 1 ? g.ny : g.nx, n1 = c1d == 1 ? g.ny : g.nz;
    l0 = n0 > 1 ? -1 : 0;
    w0 = n0 > 1 ? n0 + 2 : n0;
    l1 = n1 > 1 ? -1 : 0;
    const long w1 = n1 > 1 ? n1 + 2 : n1;
    wf = (DIM == 0 ? g.nx : (DIM == 1 ? g.ny : g.nz)) + 1;
    if (DIM == 0) {
      t0 = b * ft.fx;
      const long left = wf * w0 * w1 - t0;
      nface = (int)(left < ft.fx ? left : ft.fx);
      r0 = t0 / wf;
      e0 = r0 * (wf + 1) + t0 % wf;
      ncell = right(nface - 1) + 1;
    } else {
      const long ntx = (w0 + FT_T - 1) / FT_T, nti = (wf + FT_F - 1) / FT_F;
      x0 = (b % ntx) * FT_T;
      i0 = ((b / ntx) % nti) * FT_F;
      c1 = b / (ntx * nti) + l1;
      nface = FT_T * FT_F;
      ncell = FT_T * (FT_F + 1);
    }
  }
  __device__ int left(int f) const { return DIM == 0 ? (int)(f + (t0 + f) / wf - r0) : f; }
  __device__ int right(int f) const { return DIM == 0 ? left(f) + 1 : f + FT_T; }
  __device__ bool face(int f, long cell[3]) const {
    if (f >= nface) return false;
    if (DIM == 0) {
      const long t = t0 + f;
      cell[0] = t % wf;
      cell[1] = (t / wf) % w0 + l0;
      cell[2] = t / (wf * w0) + l1;
      return true;
    }
    const long x = x0 + f % FT_T, i = i0 + f / FT_T;
    if (x >= w0 || i >= wf) return false;
    cell[DIM] = i;
    cell[c0d] = x + l0;
    cell[c1d] = c1;
    return true;
  }
  __device__ bool cellk(int k, long cell[3]) const {
    if (k >= ncell) return false;
    if (DIM == 0) {
      const long e = e0 + k, r = e / (wf + 1);
      cell[0] = e % (wf + 1) - 1;
      cell[1] = r % w0 + l0;
      cell[2] = r / w0 + l1;
      return true;
    }
    const long x = x0 + k % FT_T, i = i0 - 1 + k / FT_T;
    if (x >= w0 || i >= wf) return false;
    cell[DIM] = i;
    cell[c0d] = x + l0;
    cell[c1d] = c1;
    return true;
  }
};
// sim.zig:1316-1337: the reconstruction order of a cell: PPM, one lower with
// REDUCEORDERATBH inside the horizon (radial index <= reduce_order_ix_max)
// or REDUCEORDERAFTERFIXUP on a cell that the last u2p or implicit pass
// flagged for a fixup (ghost cells carry no flag)
__device__ __forceinline__ int recon_order(const GridD &g, const Dev &d, const kz_physics &ph, const long cell[3]) {
  if (ph.reduceorderatbh && cell[0] <= ph.reduce_order_ix_max) return 1;
  if (ph.reduceorderafterfixup) {
    const long c = cidx(g, cell[0], cell[1], cell[2]), n = g.ncell;
    if (d.flags[KZ_FLAG_HD_FIXUP * n + c] != 0 || d.flags[KZ_FLAG_RAD_FIXUP * n + c] != 0 ||
        d.flags[KZ_FLAG_RADIMP_FIXUP * n + c] != 0)
      return 1;
  }
  return 2;
}
template <int DIM, bool VISC, bool OPTS>
__device__ __forceinline__ void face_tile(const GridD &g, const Dev &d, const kz_physics &ph, const FaceTiles &ft,
                                          long b, FaceShared &sh) {
  constexpr int dim = DIM;
  const FaceTile<DIM> tl(g, ft, b);
  const int j = threadIdx.x;
  const long n = g.ncell;
  const long stride = dim == 0 ? 1 : (dim == 1 ? g.sx : (long)g.sx * g.sy);
  long cell[3];
  // phase 1: both face values of cell j & 63 over the stencil along DIM;
  // the block's first wave takes slots [0, 7), the second [7, 13)
  const int kc = j & 63;
  if (tl.cellk(kc, cell)) {
    const long i = cell[dim];
    const long c = cidx(g, cell[0], cell[1], cell[2]);
    double dxs[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) dxs[k] = csize(g, i - 2 + k, dim);
    const int iv0 = j < 64 ? 0 : 7, iv1 = j < 64 ? 7 : NV;
    const int order = OPTS ? recon_order(g, d, ph, cell) : 2;
    for (int iv = iv0; iv < iv1; ++iv) {
      const double *f = d.p + iv * n;
      double s5[5];
#pragma unroll
      for (int k = 0; k < 5; ++k) s5[k] = f[c + (k - 2) * stride];
      double lo, ro;
      if (order == 2) ppm(s5, dxs, &lo, &ro);
      else reconstruct(order, ph.minmod_theta, s5, dxs, &lo, &ro);
      sh.st[0][iv][kc] = lo;
      sh.st[1][iv][kc] = ro;
    }
  }
  __syncthreads();
  // phase 2: face side j & 1 of face j >> 1
  const int slot = j >> 1, side = j & 1;
  const bool act = tl.face(slot, cell);
  double fl[NV], ul[NV], ag = 0, agr = 0;
  long fi = 0;
  if (act) {
    const int k = side == 0 ? tl.left(slot) : tl.right(slot);
    double ps[NV];
#pragma unroll
    for (int iv = 0; iv < NV; ++iv) ps[iv] = sh.st[side == 0 ? 1 : 0][iv][k];
    const long cf = cidx(g, cell[0], cell[1], cell[2]);
    const long cm = cf - stride;
    Geo gm;
    geo_f(g, d, dim, cell[0], cell[1], cell[2], gm);
    int guard = 0;
    if (OPTS && ph.floors.b2rhofloorframe == 1) check_floors_mhd<true>(ps, gm, ph.gam, ph.floors, &guard);
    else check_floors_mhd<false>(ps, gm, ph.gam, ph.floors, &guard);
    if (flux_prime_d<DIM>(ps, gm, ph.gam, fl)) set_err(d.err, E_NANFLUX);
    p2u(ps, ul, gm, ph.gam);
    if (VISC) {
      double rv[4];
#pragma unroll
      for (int q = 0; q < 4; ++q)
        rv[q] = 0.5 * (d.rijvisc[((DIM + 1) * 4 + q) * n + cf] + d.rijvisc[((DIM + 1) * 4 + q) * n + cm]);
      add_visc_flux<DIM>(rv, ul, gm, g, ph, fl);
    }
    ag = kzm_fmax(d.scal[(SC_AHD + dim) * n + cf], d.scal[(SC_AHD + dim) * n + cm]);
    agr = kzm_fmax(d.scal[(SC_ARAD + dim) * n + cf], d.scal[(SC_ARAD + dim) * n + cm]);
    fi = fidx(g, dim, cell[0], cell[1], cell[2]);
    if (guard) atomicAdd(&d.cnt[3], (unsigned long long)guard);
  }
  // phase 3: the two sides of a face are neighbouring lanes; each takes the
  // other's flux and conserved vector and writes the LAXF flux of half the
  // slots (side 0: [0, 7), side 1: [7, 13))
  double fo[NV], uo[NV];
#pragma unroll
  for (int iv = 0; iv < NV; ++iv) {
    fo[iv] = __shfl_xor(fl[iv], 1);
    uo[iv] = __shfl_xor(ul[iv], 1);
  }
  if (!act) return;
  const long nfd = g.nface[dim];
  double *out = d.flb[dim];
#pragma unroll
  for (int iv = 0; iv < NV; ++iv) {
    if ((iv < 7) != (side == 0)) continue;
    const double fL = side == 0 ? fl[iv] : fo[iv], fR = side == 0 ? fo[iv] : fl[iv];
    const double uL = side == 0 ? ul[iv] : uo[iv], uR = side == 0 ? uo[iv] : ul[iv];
    out[iv * nfd + fi] = laxf(fL, fR, uL, uR, iv >= EE ? agr : ag);
  }
}
// One launch for the faces of every active dimension (faces of different
// dimensions are independent): blocks [0, nb[0]) are x tiles, then y, then
// z.  VISC adds the radiative-viscosity flux (sim.zig:1349-1364).  OPTS
// reads the reference options of this pass from the parameter block (the
// reconstruction order per cell, the b^2/rho floor frame); without it the
// kernel is compiled for PPM everywhere and the drift frame, so the default
// build carries none of their code.
template <bool VISC, bool OPTS>
__global__ void __launch_bounds__(128) KZ_WPE(KZ_WPE_FACE) k_face_flux(GridD g, Dev d, kz_physics ph, FaceTiles ft) {
  __shared__ FaceShared sh;
  long b = blockIdx.x;
  if (b < ft.nb[0]) face_tile<0, VISC, OPTS>(g, d, ph, ft, b, sh);
  else if (b < ft.nb[0] + ft.nb[1]) face_tile<1, VISC, OPTS>(g, d, ph, ft, b - ft.nb[0], sh);
  else face_tile<2, VISC, OPTS>(g, d, ph, ft, b - ft.nb[0] - ft.nb[1], sh);
}


// implicit four-force solve (sim.zig:1736-1776 -> implicit.zig:927-1004),
// split in three kernels so that each keeps its state in registers:
//   k_imp_prep     reference state, bisection warm start, starting block
//   k_imp_rung0    the first rung (energy equation, lab frame), one launch
//                  over both starting-block lists: the RAD-start cells (the
//                  longer solves) take the first blocks, so their waves are
//                  dispatched first and the MHD-start waves fill the SIMDs
//                  that remain and those freed as waves finish
//   k_imp_slow     rungs 1..5 (and damped-opacity levels) for the cells the
//                  first rung did not solve, gathered into a list
// Batches, whose cells need from one to forty iterations per rung, run the
// first rung and then each later rung of level 0 as persistent work-stealing
// launches (k_imp_rung0_persist, k_imp_rung_persist): every lane holds one
// cell's Newton state and takes the next cell of the queue when its cell is
// solved or has failed, so a wave's time is its lanes' total iterations, not
// 64 times its slowest cell's; k_imp_slow then takes the damped levels.
// The sequence of attempts per cell is the reference ladder's; inputs to
// every rung are the stored reference-state values, so results equal the
// single-pass ladder bit for bit.
// Scratch of the split solve, indexed by the solve's cell index t.
struct ImpWork {
  double *uu00;  // [13][n] conserved state of the incoming primitives
  double *pp0;   // [13][n] Newton start (after the bisection warm start)
  double *aux;   // [IMP_NAUX][n] reference-state scalars
  int32_t *slow;                  // cells that need rungs past the first
  int32_t *list;                  // [2][n] cells by starting block (MHD, RAD)
  unsigned long long *nslow;      // length of `slow`
  unsigned long long *ncnt;       // [2] lengths of the two lists
  // batches (null for grids): the cells the persistent first rung left,
  // rl = [2][n] by starting block (MHD, RAD) with lengths rc = [2], and the
  // queue claim positions qpos (first rung RAD, MHD; the later-rung ladder)
  int32_t *rl;
  unsigned long long *rc;
  unsigned long long *qpos;
};
// The cells of one run's domain (opImplicit, sim.zig:1736-1776): polar rows
// are skipped, dt is a step scalar, a solved cell stores u and p and clears
// its radimp flag (sim.zig:1762-1768), a failed one sets it to -1.
struct GridCells {
  GridD g;
  Dev d;
  int sslot;
  __device__ long count() const { return (long)g.nx * g.ny * g.nz; }
  __device__ void ijk(long t, long &ix, long &iy, long &iz) const {
    ix = t % g.nx;
    iy = (t / g.nx) % g.ny;
    iz = t / ((long)g.nx * g.ny);
  }
  __device__ bool skip(const kz_physics &ph, long t) const { return polar_row(ph, g, (t / g.nx) % g.ny); }
  __device__ void load_p(long t, double *pp) const {
    long ix, iy, iz;
    ijk(t, ix, iy, iz);
    ld(d.p, g.ncell, cidx(g, ix, iy, iz), pp);
  }
  __device__ void geo(long t, Geo &gm) const {
    long ix, iy, iz;
    ijk(t, ix, iy, iz);
    geo_c(g, d, ix, iy, iz, gm);
  }
  __device__ double dt(long) const { return d.ss[sslot]; }
  __device__ void solved(const kz_physics &ph, long t, const Geo &gm, const double *ppw, int iters, int, int, int,
                         int) const {
    long ix, iy, iz;
    ijk(t, ix, iy, iz);
    const long c = cidx(g, ix, iy, iz);
    double uu[NV];
    p2u(ppw, uu, gm, ph.gam);
    st(d.u, g.ncell, c, uu);
    st(d.p, g.ncell, c, ppw);
    d.flags[KZ_FLAG_RADIMP_FIXUP * g.ncell + c] = 0;
    atomicAdd(&d.cnt[1], (unsigned long long)iters);
    atomicAdd(&d.cnt[2], 1ULL);
  }
  __device__ void failed(long t) const {
    long ix, iy, iz;
    ijk(t, ix, iy, iz);
    d.flags[KZ_FLAG_RADIMP_FIXUP * g.ncell + cidx(g, ix, iy, iz)] = -1;
    atomicAdd(&d.cnt[0], 1ULL);
  }
};
// n independent cells (solveImplicitLab per record, implicit.zig:927-1009):
// geometry planes and dt per cell, the ladder's outcome recorded per cell.
struct BatchCells {
  long n;
  const double *gp;   // [GEO_NC + 1][n]: the geometry planes, then x1
  const double *dtv;  // [n]
  double *uu, *pp;    // [NV][n]
  kz_implicit_result *res;
  unsigned long long *cnt;  // [0] failures [1] iterations [2] solves
  __device__ long count() const { return n; }
  __device__ bool skip(const kz_physics &, long) const { return false; }
  __device__ void load_p(long t, double *p) const { ld(pp, n, t, p); }
  __device__ void geo(long t, Geo &gm) const { load_geo(gp, n, t, gp[GEO_NC * n + t], gm); }
  __device__ double dt(long t) const { return dtv[t]; }
  __device__ void solved(const kz_physics &ph, long t, const Geo &gm, const double *ppw, int iters, int ir, int wp,
                         int we, int wf) const {
    double u[NV];
    p2u(ppw, u, gm, ph.gam);
    st(uu, n, t, u);
    st(pp, n, t, ppw);
    res[t] = kz_implicit_result{1, ir, iters, wp, we, wf};
    atomicAdd(&cnt[1], (unsigned long long)iters);
    atomicAdd(&cnt[2], 1ULL);
  }
  __device__ void failed(long t) const {
    res[t] = kz_implicit_result{0, -1, 0, WP_RAD, WE_ENERGY, WF_LAB};
    atomicAdd(&cnt[0], 1ULL);
  }
};
// The cells of several runs as one sequence (a run group's implicit pass):
// run k owns t in [off_k, off_k + ndom_k), each with its own grid, fields,
// metric, dt slot value and counters.
struct RunCells {
  GridD g;
  Dev d;
  long off;
};
struct MultiCells {
  const RunCells *r;  // [nr], off ascending, r[0].off == 0
  int nr;
  long n;
  int sslot;
  __device__ int run(long t) const {
    int lo = 0, hi = nr - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (r[mid].off <= t) lo = mid;
      else hi = mid - 1;
    }
    return lo;
  }
  __device__ GridCells at(long t, long &tl) const {
    const int k = run(t);
    tl = t - r[k].off;
    return GridCells{r[k].g, r[k].d, sslot};
  }
  __device__ long count() const { return n; }
  __device__ bool skip(const kz_physics &ph, long t) const {
    long tl;
    const GridCells c = at(t, tl);
    return c.skip(ph, tl);
  }
  __device__ void load_p(long t, double *pp) const {
    long tl;
    const GridCells c = at(t, tl);
    c.load_p(tl, pp);
  }
  __device__ void geo(long t, Geo &gm) const {
    long tl;
    const GridCells c = at(t, tl);
    c.geo(tl, gm);
  }
  __device__ double dt(long t) const {
    long tl;
    const GridCells c = at(t, tl);
    return c.dt(tl);
  }
  __device__ void solved(const kz_physics &ph, long t, const Geo &gm, const double *ppw, int iters, int ir, int wp,
                         int we, int wf) const {
    long tl;
    const GridCells c = at(t, tl);
    c.solved(ph, tl, gm, ppw, iters, ir, wp, we, wf);
  }
  __device__ void failed(long t) const {
    long tl;
    const GridCells c = at(t, tl);
    c.failed(tl);
  }
};
template <class S, bool BR>
__global__ void __launch_bounds__(64) KZ_WPE(KZ_WPE_PREP) k_imp_prep(S s, ImpWork w, kz_physics ph) {
  ph.opac.synchrotron_bridge = BR;
  long t = TID;
  if (t == 0) {  // no appends run concurrently with this kernel
    *w.nslow = 0;
    if (w.qpos) {
      w.rc[0] = w.rc[1] = 0;
      for (int k = 0; k < 3; ++k) w.qpos[k] = 0;
    }
  }
  const long nd = s.count();
  if (t >= nd) return;
  if (s.skip(ph, t)) return;
  double dt = s.dt(t);
  double pp00[NV], pp0[NV];
  s.load_p(t, pp00);
  Geo gm;
  s.geo(t, gm);
  // the reference state and the starting block are stored before the
  // bisection, so that only its three inputs stay live through it
  double st_ehat, st_ut, st_tgas;
  int sw;
  {
    double uu00[NV];
    p2u(pp00, uu00, gm, ph.gam);
    st(w.uu00, nd, t, uu00);
    const double ehat = -calc_ff_rtt(pp00, gm);
    RadState st00;
    fill_radstate(pp00, gm, ph.gam, ph.opac, true, st00);
    sw = (ehat < ph.implicit.threshold * pp00[UU]) ? WP_RAD : WP_MHD;
    double *a = w.aux;
    a[IA_EHAT * nd + t] = st00.ehat;
    a[IA_UINT * nd + t] = st00.uint;
    a[IA_SGAS * nd + t] = st00.sgas;
    a[IA_KES * nd + t] = st00.kappaes;
    a[IA_KGA * nd + t] = st00.op.gas_abs;
    a[IA_KRA * nd + t] = st00.op.rad_abs;
    a[IA_KRR * nd + t] = st00.op.rad_ross;
    a[IA_SW * nd + t] = (double)sw;
    st_ehat = st00.ehat;
    st_ut = st00.ucon[0];
    st_tgas = st00.tgas;
  }
  // gather the cell into the list of its starting block, so that each
  // first-rung kernel runs full waves of one Newton variant
  if (sw == WP_MHD) {
    unsigned long long k = atomicAdd(&w.ncnt[0], 1ULL);
    w.list[k] = (int32_t)t;
  } else {
    unsigned long long k = atomicAdd(&w.ncnt[1], 1ULL);
    w.list[nd + k] = (int32_t)t;
  }
  ImpCtx cx{&gm, ph.gam, &ph.rad, &ph.opac, &ph.implicit, 1.0};
  if (ph.implicit.start_with_bisect) {
    solve_1d_prim(pp00, st_ehat, st_ut, st_tgas, dt, cx, pp0);
  } else {
#pragma unroll
    for (int i = 0; i < NV; ++i) pp0[i] = pp00[i];
  }
  st(w.pp0, nd, t, pp0);
}
__device__ __forceinline__ void imp_load(const ImpWork &w, long nd, long t, double *uu00, double *pp0, State0 &s0,
                                         LagOpac &lag) {
  ld(w.uu00, nd, t, uu00);
  ld(w.pp0, nd, t, pp0);
  const double *a = w.aux;
  s0.ehat = a[IA_EHAT * nd + t];
  s0.uint = a[IA_UINT * nd + t];
  s0.sgas = a[IA_SGAS * nd + t];
  lag.kappaes = a[IA_KES * nd + t];
  lag.op.gas_abs = a[IA_KGA * nd + t];
  lag.op.rad_abs = a[IA_KRA * nd + t];
  lag.op.rad_ross = a[IA_KRR * nd + t];
}
template <class S, int WP>
__device__ __forceinline__ void rung0_cell(const S &s, const ImpWork &w, const kz_physics &ph, long i) {
  constexpr int which = WP == WP_MHD ? 0 : 1;
  if (i >= (long)w.ncnt[which]) return;
  const long nd = s.count();
  long t = w.list[which * nd + i];
  double uu00[NV], pp0[NV], ppw[NV];
  State0 s0;
  LagOpac lag;
  imp_load(w, nd, t, uu00, pp0, s0, lag);
  Geo gm;
  s.geo(t, gm);
  double dt = s.dt(t);
  ImpCtx cx{&gm, ph.gam, &ph.rad, &ph.opac, &ph.implicit, 1.0};
  int iters = 0;
  bool ok = ph.implicit.opdamp_maxlevels >= 0 &&
            solve_4d_prim_core<WP>(uu00, lag, s0, pp0, dt, cx, WE_ENERGY, WF_LAB, ppw, &iters);
  if (ok) {
    s.solved(ph, t, gm, ppw, iters, 0, WP, WE_ENERGY, WF_LAB);
  } else {
    unsigned long long k = atomicAdd(w.nslow, 1ULL);
    w.slow[k] = (int32_t)t;
  }
}
// launched with nblk(n, 64) + 1 blocks: ceil(nrad/64) + ceil(nmhd/64) never
// exceeds that for nrad + nmhd <= n
template <class S, bool BR>
__global__ void __launch_bounds__(64) KZ_WPE(KZ_WPE_RUNG0) k_imp_rung0(S s, ImpWork w, kz_physics ph) {
  ph.opac.synchrotron_bridge = BR;
  const long brad = ((long)w.ncnt[1] + 63) / 64;
  const long b = blockIdx.x;
  if (b < brad) rung0_cell<S, WP_RAD>(s, w, ph, b * 64 + threadIdx.x);
  else rung0_cell<S, WP_MHD>(s, w, ph, (b - brad) * 64 + threadIdx.x);
}
// Where a cell of the persistent first-rung launch goes when its Newton
// ends: solved (recorded with the rung's attempt) or to the later rungs'
// list of its starting block.
template <class S>
struct RungDone {
  const S *s;
  const ImpWork *w;
  const kz_physics *ph;
  int ir, which, wp, we, wf;
  __device__ void operator()(long t, const Geo &gm, const NewtonState &n, bool ok) const {
    if (ok) {
      s->solved(*ph, t, gm, n.pp, n.iter, ir, wp, we, wf);
      return;
    }
    const long nd = s->count();
    unsigned long long k = atomicAdd(&w->rc[which], 1ULL);
    w->rl[which * nd + k] = (int32_t)t;
  }
};
// Work stealing over one queue (cells list[0 .. qlen)) with Newton variant
// WP at opdamp level 0: each lane holds one cell; the lanes without one claim
// the next queue entries together (one atomic per wave and refill) until the
// queue is exhausted; the loop ends when no lane holds a cell.  Every Newton
// ends within maxiter + 1 steps and the queue is finite, so every wave exits.
template <class S, int WP>
__device__ __forceinline__ void drain_queue(const S &s, const ImpWork &w, const kz_physics &ph, const int32_t *list,
                                            long qlen, unsigned long long *qpos, int we, int wf,
                                            const RungDone<S> &done) {
  const int lane = threadIdx.x & 63;
  const unsigned long long below = (1ULL << lane) - 1ULL;
  const long nd = s.count();
  bool has = false, exhausted = qlen <= 0;
  long t = 0;
  double dt = 0;
  NewtonState ns;
  double uu00[NV];
  State0 s0;
  LagOpac lag;
  Geo gm;
  for (;;) {
    const unsigned long long need = __ballot(!has && !exhausted);
    if (need) {
      const int leader = __ffsll((unsigned long long)need) - 1;
      const int cnt = __popcll(need);
      long base = 0;
      if (lane == leader) base = (long)atomicAdd(qpos, (unsigned long long)cnt);
      base = __shfl(base, leader);
      if (!has && !exhausted) {
        const long q = base + __popcll(need & below);
        if (q < qlen && ph.implicit.opdamp_maxlevels < 0) {
          done(list[q], gm, ns, false);  // the ladder has no level to run (implicit.zig:961)
        } else if (q < qlen) {
          t = list[q];
          double pp0[NV];
          imp_load(w, nd, t, uu00, pp0, s0, lag);
          s.geo(t, gm);
          dt = s.dt(t);
          const ImpCtx cx{&gm, ph.gam, &ph.rad, &ph.opac, &ph.implicit, 1.0};
          newton_start<WP>(ns, uu00, pp0, cx);
          has = true;
        }
      }
      if (base + cnt >= qlen) exhausted = true;  // the same for every lane of the wave
    }
    if (!__ballot(has)) break;
    if (has) {
      const ImpCtx cx{&gm, ph.gam, &ph.rad, &ph.opac, &ph.implicit, 1.0};
      const int r = newton_step<WP>(ns, uu00, lag, s0, dt, cx, we, wf);
      if (r != 0) {
        done(t, gm, ns, r > 0);
        has = false;
      }
    }
  }
}
// first rung over both starting-block lists, RAD-start queue first
template <class S, bool BR>
__global__ void __launch_bounds__(64) KZ_WPE(KZ_WPE_RUNG0) k_imp_rung0_persist(S s, ImpWork w, kz_physics ph) {
  ph.opac.synchrotron_bridge = BR;
  const long nd = s.count();
  const RungDone<S> dr{&s, &w, &ph, 0, 1, WP_RAD, WE_ENERGY, WF_LAB};
  drain_queue<S, WP_RAD>(s, w, ph, w.list + nd, (long)w.ncnt[1], &w.qpos[0], WE_ENERGY, WF_LAB, dr);
  const RungDone<S> dm{&s, &w, &ph, 0, 0, WP_MHD, WE_ENERGY, WF_LAB};
  drain_queue<S, WP_MHD>(s, w, ph, w.list, (long)w.ncnt[0], &w.qpos[1], WE_ENERGY, WF_LAB, dm);
}
// Rungs 1..5 of level 0 (implicit.zig:961-968) over the cells the first
// rung left (both starting-block lists, MHD start then RAD start), work
// stealing as drain_queue but with each lane walking its own cell up the
// ladder: when an attempt fails the lane starts the next rung's Newton (whose
// primitive block may differ from its neighbours', the two variants then run
// one after the other), so the long ladders of different rungs overlap
// instead of each rung waiting for its slowest cell.  Cells past rung 5 go
// to the slow list.
template <class S, bool BR>
__global__ void __launch_bounds__(64) KZ_WPE(KZ_WPE_RUNG0) k_imp_ladder_persist(S s, ImpWork w, kz_physics ph) {
  ph.opac.synchrotron_bridge = BR;
  const int lane = threadIdx.x & 63;
  const unsigned long long below = (1ULL << lane) - 1ULL;
  const long nd = s.count();
  const long n0 = (long)w.rc[0], qlen = n0 + (long)w.rc[1];
  bool has = false, exhausted = qlen <= 0;
  long t = 0;
  int sw = WP_MHD, ir = 1, wp = WP_MHD, we = WE_ENERGY, wf = WF_FF;
  double dt = 0;
  NewtonState ns;
  double uu00[NV], pp0[NV];
  State0 s0;
  LagOpac lag;
  Geo gm;
  for (;;) {
    const unsigned long long need = __ballot(!has && !exhausted);
    if (need) {
      const int leader = __ffsll((unsigned long long)need) - 1;
      const int cnt = __popcll(need);
      long base = 0;
      if (lane == leader) base = (long)atomicAdd(&w.qpos[2], (unsigned long long)cnt);
      base = __shfl(base, leader);
      if (!has && !exhausted) {
        const long q = base + __popcll(need & below);
        if (q < qlen) {
          sw = q < n0 ? WP_MHD : WP_RAD;
          t = q < n0 ? w.rl[q] : w.rl[nd + (q - n0)];
          if (ph.implicit.opdamp_maxlevels < 0) {  // no level to run (implicit.zig:961)
            unsigned long long k = atomicAdd(w.nslow, 1ULL);
            w.slow[k] = (int32_t)t;
          } else {
            imp_load(w, nd, t, uu00, pp0, s0, lag);
            s.geo(t, gm);
            dt = s.dt(t);
            ir = 0;  // the next attempt is rung 1
            has = true;
          }
        }
      }
      if (base + cnt >= qlen) exhausted = true;  // the same for every lane of the wave
    }
    if (!__ballot(has)) break;
    if (has) {
      const ImpCtx cx{&gm, ph.gam, &ph.rad, &ph.opac, &ph.implicit, 1.0};
      int r = -1;
      if (ir > 0) {
        r = (wp == WP_MHD) ? newton_step<WP_MHD>(ns, uu00, lag, s0, dt, cx, we, wf)
                           : newton_step<WP_RAD>(ns, uu00, lag, s0, dt, cx, we, wf);
      }
      if (r > 0) {
        s.solved(ph, t, gm, ns.pp, ns.iter, ir, wp, we, wf);
        has = false;
      } else if (r < 0) {
        if (ir == 5) {
          unsigned long long k = atomicAdd(w.nslow, 1ULL);
          w.slow[k] = (int32_t)t;
          has = false;
        } else {  // the next rung's attempt from the stored reference state
          ir += 1;
          const int ot = (sw == WP_RAD) ? WP_MHD : WP_RAD;
          wp = (ir == 1 || ir == 4) ? sw : ot;
          we = (ir >= 4) ? WE_ENTROPY : WE_ENERGY;
          wf = (ir == 2) ? WF_LAB : WF_FF;
          if (wp == WP_MHD) newton_start<WP_MHD>(ns, uu00, pp0, cx);
          else newton_start<WP_RAD>(ns, uu00, pp0, cx);
        }
      }
    }
  }
}

// The rest of the ladder over `slow` from opdamp level `first_level` (0:
// rungs 1..5 of level 0 first; 1: after the persistent rung launches).
template <class S, bool BR>
__global__ void __launch_bounds__(64) k_imp_slow(S s, ImpWork w, kz_physics ph, int first_level) {
  ph.opac.synchrotron_bridge = BR;
  long i = TID;
  // the first-rung lists have been consumed; clear them for the next pass
  if (i == 0) w.ncnt[0] = w.ncnt[1] = 0;
  if (i >= (long)*w.nslow) return;
  const long nd = s.count();
  long t = w.slow[i];
  double uu00[NV], pp0[NV];
  State0 s0;
  LagOpac lag;
  imp_load(w, nd, t, uu00, pp0, s0, lag);
  const int sw = (int)w.aux[IA_SW * nd + t];
  const int ot = (sw == WP_RAD) ? WP_MHD : WP_RAD;
  Geo gm;
  s.geo(t, gm);
  double dt = s.dt(t);
  const kz_implicit_params &ip = ph.implicit;
  ImpCtx cx{&gm, ph.gam, &ph.rad, &ph.opac, &ip, 1.0};
  for (int level = first_level; level <= ip.opdamp_maxlevels; ++level) {
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
        s.solved(ph, t, gm, ppw, iters, ir, wp, we, wf);
        return;
      }
    }
  }
  s.failed(t);
}

// The implicit kernels over the n cells of a cell source, with the
// synchrotron bridge of the opacities (BR) fixed on or off at kernel entry:
// left to a runtime test, the bridge's branch costs the ladder's register
// allocation about 64 spilled registers.  w carries the queue arrays when
// `persist`.
template <class S, bool BR>
static void launch_implicit_br(const S &cells, const ImpWork &w, const kz_physics &ph, long n, bool persist,
                               unsigned persist_blocks, hipStream_t st) {
  const unsigned nb = nblk(n, 64);
  k_imp_prep<S, BR><<<nb, 64, 0, st>>>(cells, w, ph);
  if (persist) {
    const unsigned np = persist_blocks < nb + 1 ? persist_blocks : nb + 1;
    k_imp_rung0_persist<S, BR><<<np, 64, 0, st>>>(cells, w, ph);
    k_imp_ladder_persist<S, BR><<<np, 64, 0, st>>>(cells, w, ph);
    k_imp_slow<S, BR><<<nb, 64, 0, st>>>(cells, w, ph, 1);
  } else {
    k_imp_rung0<S, BR><<<nb + 1, 64, 0, st>>>(cells, w, ph);
    k_imp_slow<S, BR><<<nb, 64, 0, st>>>(cells, w, ph, 0);
  }
}
// The option variants (kz_options.hip).
void opt_launch_wavespeeds(const GridD &g, const Dev &d, const kz_physics &ph, unsigned nb, hipStream_t st);
void opt_launch_u2p_zamo(const GridD &g, const Dev &d, const kz_physics &ph, unsigned nb, hipStream_t st);
void opt_launch_face_flux(const GridD &g, const Dev &d, const kz_physics &ph, const FaceTiles &ft, unsigned nb,
                          hipStream_t st);
template <class S>
void opt_launch_implicit_bridge(const S &cells, const ImpWork &w, const kz_physics &ph, long n, bool persist,
                                unsigned persist_blocks, hipStream_t st);

#endif
