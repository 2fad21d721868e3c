/* kzo_sim.c — oracle restatement of koral/sim.zig's Sim(config.puffy) passes
 * and RK2IMEX step, sim/ct.zig flux-CT, sim/bc.zig + the PUFFY Bc.calc
 * (koral/problems/puffy/puffy.zig:1080-1194), sim/polaraxis.zig.
 * TEST INFRASTRUCTURE (see kzo.h).
 *
 * Storage mirrors the reference exactly: AoS fields iv-fastest over the
 * padded grid (koral/field.zig), FaceStore (+1 slice in its own dim,
 * sim/storage.zig:62-135), MetricCache blocks (metric/precompute.zig:121-526)
 * passed in by the caller.  Also radiative viscosity (sim/rijvisc.zig,
 * physics/radvisc.zig) and the mimic dynamo (sim/dynamo.zig); 2D and 3D
 * grids (periodic z, fillCorners3d).
 */
#include <stdlib.h>
#include <string.h>
#include "kzo.h"

/* Threads of the cell-parallel passes (wavespeeds, sweeps, face fluxes,
 * update, u2p, implicit): rows are dealt out dynamically, every cell and face
 * is computed by one thread exactly as the serial loop computes it, and the
 * only cross-cell results are max/min and integer sums, so any thread count
 * gives the serial result bit for bit.  Default 1. */
static int kzo_nthreads = 1;
void kzo_set_threads(int n) { kzo_nthreads = n < 1 ? 1 : n; }
int kzo_get_threads(void) { return kzo_nthreads; }

#define NSCAL 20
enum {
  S_AHDXL, S_AHDXR, S_AHDX, S_ARADXL, S_ARADXR, S_ARADX,
  S_AHDYL, S_AHDYR, S_AHDY, S_ARADYL, S_ARADYR, S_ARADY,
  S_AHDZL, S_AHDZR, S_AHDZ, S_ARADZL, S_ARADZR, S_ARADZ,
  S_TSTEPDEN, S_CELL_DT
};
static const int ahd_l[3] = {S_AHDXL, S_AHDYL, S_AHDZL};
static const int ahd_r[3] = {S_AHDXR, S_AHDYR, S_AHDZR};
static const int ahd_m[3] = {S_AHDX, S_AHDY, S_AHDZ};
static const int arad_l[3] = {S_ARADXL, S_ARADYL, S_ARADZL};
static const int arad_r[3] = {S_ARADXR, S_ARADYR, S_ARADZR};
static const int arad_m[3] = {S_ARADX, S_ARADY, S_ARADZ};

#define BCX_REC 66 /* xhi ghost record: gBL[16] GBL[16] rghost rbound J12[16] J21[16] */

typedef struct kzo_sim {
  kz_grid grid;
  int ngx, ngy, ngz, sx, sy, sz;
  long ncell;
  double dx, dy, dz;
  kz_physics ph;
  const double *g, *gcon, *kris;
  const double *gb[3], *gconb[3];
  const double *bcx;
  double *p, *u, *ut0, *ut1, *ut2, *dut1, *dut2, *drt1, *drt2, *uforget, *u_bak, *p_bak;
  double *pb_l[3], *pb_r[3], *fl_l[3], *fl_r[3], *flb[3];
  long nface[3];
  int fsx[3], fsy[3], fsz[3];
  double *scal;
  int32_t *flags;
  double *emf[3];
  double t, dt, own_dt, tstepdenmax, tstepdenmin, min_dx, min_dy, min_dz;
  uint64_t nstep, n_fail, n_iters, n_solves, n_guard;
  /* radiative viscosity and dynamo state */
  const double *bl;   /* BL sidecar, BL_REC per padded cell (caller-owned) */
  double *rijvisc;    /* R^i_j viscous, 16 per padded cell */
  double *dyn_a;      /* Delta A_phi per padded cell */
  double *vecpot;     /* 6 per padded cell: corner A (0..2), curl B (3..5) */
  double *scaleth;    /* nx */
  double rhor, risco;
} kzo_sim;

#define BL_REC 52 /* gBL[16] GBL[16] alpha gdet r theta J_mks2->bl[16] */

/* ---- grid (koral/grid.zig:127-159) ---- */
static double xl_(const kzo_sim *s, long i) { return s->grid.minx + (double)i * s->dx; }
static double yl_(const kzo_sim *s, long i) { return s->grid.miny + (double)i * s->dy; }
static double zl_(const kzo_sim *s, long i) { return s->grid.minz + (double)i * s->dz; }
static double xc_(const kzo_sim *s, long i) { return 0.5 * (xl_(s, i) + xl_(s, i + 1)); }
static double yc_(const kzo_sim *s, long i) { return 0.5 * (yl_(s, i) + yl_(s, i + 1)); }
static double zc_(const kzo_sim *s, long i) { return 0.5 * (zl_(s, i) + zl_(s, i + 1)); }
static double csize(const kzo_sim *s, long i, int dim) {
  if (dim == 0) return xl_(s, i + 1) - xl_(s, i);
  if (dim == 1) return yl_(s, i + 1) - yl_(s, i);
  return zl_(s, i + 1) - zl_(s, i);
}
static long cidx(const kzo_sim *s, long ix, long iy, long iz) {
  return (ix + s->ngx) + (iy + s->ngy) * (long)s->sx + (iz + s->ngz) * (long)s->sy * s->sx;
}
static double *cell(double *f, const kzo_sim *s, long ix, long iy, long iz) { return f + cidx(s, ix, iy, iz) * NV; }
static long fidx(const kzo_sim *s, int dim, long ix, long iy, long iz) {
  long jx = ix + s->ngx, jy = iy + s->ngy, jz = iz + s->ngz;
  return ((jz * s->fsy[dim] + jy) * s->fsx[dim] + jx);
}
static double *face(double *f, const kzo_sim *s, int dim, long ix, long iy, long iz) {
  return f + fidx(s, dim, ix, iy, iz) * NV;
}
static long mfidx(const kzo_sim *s, int dim, long ix, long iy, long iz) {
  long jx = ix + s->ngx, jy = iy + s->ngy, jz = iz + s->ngz;
  if (dim == 0) return jx + jy * (s->sx + 1) + jz * (long)s->sy * (s->sx + 1);
  if (dim == 1) return jx + jy * s->sx + jz * (long)(s->sy + 1) * s->sx;
  return jx + jy * s->sx + jz * (long)s->sy * s->sx;
}
static void geo_c(const kzo_sim *s, long ix, long iy, long iz, kzo_geo *g) {
  long ci = cidx(s, ix, iy, iz);
  kzo_geo_from_blocks(s->g + ci * 20, s->gcon + ci * 20, xc_(s, ix), g);
}
static void geo_f(const kzo_sim *s, long ix, long iy, long iz, int dim, kzo_geo *g) {
  long fi = mfidx(s, dim, ix, iy, iz);
  double x1 = dim == 0 ? xl_(s, ix) : xc_(s, ix);
  kzo_geo_from_blocks(s->gb[dim] + fi * 20, s->gconb[dim] + fi * 20, x1, g);
}
static int32_t *flagp(kzo_sim *s, int f, long ix, long iy, long iz) { return s->flags + cidx(s, ix, iy, iz) * KZ_NFLAGS + f; }
static double *scp(kzo_sim *s, int k, long ix, long iy, long iz) { return s->scal + cidx(s, ix, iy, iz) * NSCAL + k; }
static long ndim(const kzo_sim *s, int d) { return d == 0 ? s->grid.nx : (d == 1 ? s->grid.ny : s->grid.nz); }

/* ---- create / destroy ---- */
kzo_sim *kzo_sim_create(const kz_grid *gr, const kz_physics *ph, const double *g, const double *gcon,
                        const double *kris, const double *gb0, const double *gconb0, const double *gb1,
                        const double *gconb1, const double *gb2, const double *gconb2, const double *bcx, int *err) {
  *err = 0;
  if (gr->ny <= 1 || gr->nx <= 1 || gr->ng < 3) { *err = KZ_E_UNSUPPORTED; return 0; }
  if (ph->floors.b2rhofloorframe < 0 || ph->floors.b2rhofloorframe > 1) { *err = KZ_E_INVALID_CONFIG; return 0; }
  if (ph->radviscosity && !ph->has_opac) { *err = KZ_E_INVALID_CONFIG; return 0; } /* sim.zig:335 */
  if (ph->correct_polaraxis && gr->ny <= 2 * ph->nccorrectpolar) { *err = KZ_E_INVALID_CONFIG; return 0; }
  kzo_sim *s = (kzo_sim *)calloc(1, sizeof(kzo_sim));
  s->grid = *gr;
  s->ph = *ph;
  s->ngx = gr->nx > 1 ? gr->ng : 0;
  s->ngy = gr->ny > 1 ? gr->ng : 0;
  s->ngz = gr->nz > 1 ? gr->ng : 0;
  s->sx = gr->nx + 2 * s->ngx;
  s->sy = gr->ny + 2 * s->ngy;
  s->sz = gr->nz + 2 * s->ngz;
  s->ncell = (long)s->sx * s->sy * s->sz;
  s->dx = (gr->maxx - gr->minx) / (double)gr->nx;
  s->dy = (gr->maxy - gr->miny) / (double)gr->ny;
  s->dz = (gr->maxz - gr->minz) / (double)gr->nz;
  s->g = g; s->gcon = gcon; s->kris = kris;
  s->gb[0] = gb0; s->gconb[0] = gconb0; s->gb[1] = gb1; s->gconb[1] = gconb1; s->gb[2] = gb2; s->gconb[2] = gconb2;
  s->bcx = bcx;
  size_t fb = (size_t)s->ncell * NV * sizeof(double);
  double **fl[] = {&s->p, &s->u, &s->ut0, &s->ut1, &s->ut2, &s->dut1, &s->dut2, &s->drt1, &s->drt2, &s->uforget, &s->u_bak, &s->p_bak};
  for (unsigned i = 0; i < sizeof(fl) / sizeof(fl[0]); ++i) *fl[i] = (double *)calloc(1, fb);
  for (int d = 0; d < 3; ++d) {
    s->fsx[d] = s->sx + (d == 0);
    s->fsy[d] = s->sy + (d == 1);
    s->fsz[d] = s->sz + (d == 2);
    s->nface[d] = (long)s->fsx[d] * s->fsy[d] * s->fsz[d];
    size_t b = (size_t)s->nface[d] * NV * sizeof(double);
    s->pb_l[d] = (double *)calloc(1, b); s->pb_r[d] = (double *)calloc(1, b);
    s->fl_l[d] = (double *)calloc(1, b); s->fl_r[d] = (double *)calloc(1, b);
    s->flb[d] = (double *)calloc(1, b);
  }
  s->scal = (double *)calloc((size_t)s->ncell * NSCAL, sizeof(double));
  s->flags = (int32_t *)calloc((size_t)s->ncell * KZ_NFLAGS, sizeof(int32_t));
  long ncorn = (long)(gr->nx + 1) * (gr->ny + 1) * (gr->nz + 1);
  for (int c = 0; c < 3; ++c) s->emf[c] = (double *)calloc(ncorn, sizeof(double));
  s->tstepdenmin = 1.0e50;
  s->rijvisc = (double *)calloc((size_t)s->ncell * 16, sizeof(double));
  s->dyn_a = (double *)calloc((size_t)s->ncell, sizeof(double));
  s->vecpot = (double *)calloc((size_t)s->ncell * 6, sizeof(double));
  s->scaleth = (double *)calloc((size_t)gr->nx, sizeof(double));
  s->rhor = kzo_r_horizon_bl(ph->bh_spin);
  s->risco = kzo_r_isco_bl(ph->bh_spin);
  /* set_grid min-size scan incl. its mdx quirk (sim.zig:447-466) */
  double mdx = -1, mdy = -1, mdz = -1;
  for (long ix = 0; ix < gr->nx; ++ix)
    for (long iy = 0; iy < gr->ny; ++iy)
      for (long iz = 0; iz < gr->nz; ++iz) {
        double dx = csize(s, ix, 0), dy = csize(s, iy, 1), dz = csize(s, iz, 2);
        if (dx < mdx || mdx < 0) mdx = dx;
        if (dy < mdx || mdy < 0) mdy = dy;
        if (dz < mdx || mdz < 0) mdz = dz;
      }
  s->min_dx = mdx; s->min_dy = mdy; s->min_dz = mdz;
  return s;
}

void kzo_sim_destroy(kzo_sim *s) {
  if (!s) return;
  double *fl[] = {s->p, s->u, s->ut0, s->ut1, s->ut2, s->dut1, s->dut2, s->drt1, s->drt2, s->uforget, s->u_bak, s->p_bak};
  for (unsigned i = 0; i < sizeof(fl) / sizeof(fl[0]); ++i) free(fl[i]);
  for (int d = 0; d < 3; ++d) { free(s->pb_l[d]); free(s->pb_r[d]); free(s->fl_l[d]); free(s->fl_r[d]); free(s->flb[d]); }
  free(s->scal); free(s->flags);
  free(s->rijvisc); free(s->dyn_a); free(s->vecpot); free(s->scaleth);
  for (int c = 0; c < 3; ++c) free(s->emf[c]);
  free(s);
}

void kzo_sim_set_state(kzo_sim *s, const double *p, const double *u) {
  if (p) memcpy(s->p, p, (size_t)s->ncell * NV * sizeof(double));
  if (u) memcpy(s->u, u, (size_t)s->ncell * NV * sizeof(double));
}
void kzo_sim_get_state(kzo_sim *s, double *p, double *u, int32_t *flags) {
  if (p) memcpy(p, s->p, (size_t)s->ncell * NV * sizeof(double));
  if (u) memcpy(u, s->u, (size_t)s->ncell * NV * sizeof(double));
  if (flags) memcpy(flags, s->flags, (size_t)s->ncell * KZ_NFLAGS * sizeof(int32_t));
}
void kzo_sim_get_scal(kzo_sim *s, double *scal) { memcpy(scal, s->scal, (size_t)s->ncell * NSCAL * sizeof(double)); }
void kzo_sim_get_flb(kzo_sim *s, int dim, double *out) { memcpy(out, s->flb[dim], (size_t)s->nface[dim] * NV * sizeof(double)); }
void kzo_sim_set_flb(kzo_sim *s, int dim, const double *in) { memcpy(s->flb[dim], in, (size_t)s->nface[dim] * NV * sizeof(double)); }
void kzo_sim_counters(kzo_sim *s, kz_counters *c) {
  c->t = s->t; c->dt = s->dt; c->own_dt = s->own_dt; c->tstepdenmax = s->tstepdenmax; c->tstepdenmin = s->tstepdenmin;
  c->nstep = s->nstep; c->n_radimp_failures = s->n_fail; c->n_radimp_iters = s->n_iters;
  c->n_radimp_solves = s->n_solves; c->n_guard_recoveries = s->n_guard;
}
void kzo_sim_set_time(kzo_sim *s, double t, double dtmax, double dtmin) {
  s->t = t; s->tstepdenmax = dtmax; s->tstepdenmin = dtmin;
}

/* ---- polar axis (sim/polaraxis.zig) ---- */
static int polar_row(const kzo_sim *s, long iy) {
  if (!s->ph.correct_polaraxis) return 0;
  long nc = s->ph.nccorrectpolar, ny = s->grid.ny;
  return iy < nc || iy > ny - nc - 1;
}

/* ---- boundary conditions (sim/bc.zig + puffy.zig:1080-1194) ---- */
static void tvec(const double *J, const double u[4], double out[4]) {
  double r[4] = {0, 0, 0, 0};
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) r[i] += J[i * 4 + j] * u[j];
  for (int i = 0; i < 4; ++i) out[i] = r[i];
}

/* frames.zig transPmhdCocoJ */
static int trans_pmhd(double *pp, const kzo_geo *g1, const kzo_geo *g2, const double *J) {
  double u1c[4], u1v[4];
  kzo_ucon_ucov_from_prims(pp, g1, u1c, u1v);
  double B[3] = {pp[B1], pp[B2], pp[B3]};
  double bcon[4], bcov[4], bsq;
  kzo_bcon_bcov_bsq(B, u1c, u1v, g1, bcon, bcov, &bsq);
  tvec(J, bcon, bcon);
  double ucon[4];
  tvec(J, u1c, ucon);
  int e = kzo_convert(ucon, VEL4, VELR, g2, 1, ucon);
  if (e) return e;
  pp[VX] = ucon[1]; pp[VY] = ucon[2]; pp[VZ] = ucon[3];
  double u2c[4], u2v[4];
  kzo_ucon_ucov_from_prims(pp, g2, u2c, u2v);
  for (int j = 1; j < 4; ++j) pp[B1 + j - 1] = bcon[j] * u2c[0] - bcon[0] * u2c[j];
  return KZO_OK;
}

static void geo_from16(const double *g16, const double *G16, kzo_geo *o) {
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) { o->g[i][j] = g16[i * 4 + j]; o->G[i][j] = G16[i * 4 + j]; }
  o->gdet = 0; o->alpha = kzm_sqrt(-1.0 / o->G[0][0]); o->gttpert = 0; o->x1 = 0;
}

/* puffy.zig:1080-1194 Bc.calc for one ghost cell: returns prims in pp */
static int puffy_bc(kzo_sim *s, long ix, long iy, long iz, int face, double *pp) {
  long nx = s->grid.nx, ny = s->grid.ny;
  if (face == 1) { /* xhi */
    memcpy(pp, cell(s->p, s, nx - 1, iy, iz), NV * sizeof(double));
    kzo_geo geom, gbl;
    geo_c(s, ix, iy, iz, &geom);
    const double *rec = s->bcx + (((iz * (long)ny) + iy) * s->grid.ng + (ix - nx)) * BCX_REC;
    geo_from16(rec, rec + 16, &gbl);
    double rghost = rec[32], rbound = rec[33];
    const double *J12 = rec + 34, *J21 = rec + 50;
    int e = trans_pmhd(pp, &geom, &gbl, J12);
    if (e) return e;
    double scale1 = rbound * rbound / rghost / rghost;
    double scale2 = rbound / rghost;
    pp[RHO] *= scale1; pp[UU] *= scale1;
    pp[B1] *= scale1; pp[B2] *= scale2; pp[B3] *= scale2;
    pp[VY] *= scale1; pp[VZ] *= scale1;
    pp[EE] *= scale1; pp[FY] *= scale1; pp[FZ] *= scale1;
    e = trans_pmhd(pp, &gbl, &geom, J21);
    if (e) return e;
    double ucon[4] = {0.0, pp[VX], pp[VY], pp[VZ]};
    e = kzo_convert(ucon, VELR, VEL4, &geom, 0, ucon);
    if (e) return e;
    tvec(J12, ucon, ucon);
    if (ucon[1] < 0.0) {
      ucon[1] = 0.0;
      tvec(J21, ucon, ucon);
      e = kzo_convert(ucon, VEL4, VELR, &geom, 0, ucon);
      if (e) return e;
      pp[VX] = ucon[1]; pp[VY] = ucon[2]; pp[VZ] = ucon[3];
    }
    double urf[4] = {0.0, pp[FX], pp[FY], pp[FZ]};
    e = kzo_convert(urf, VELR, VEL4, &geom, 0, urf);
    if (e) return e;
    tvec(J12, urf, urf);
    if (urf[1] < 0.0) {
      urf[1] = 0.0;
      tvec(J21, urf, urf);
      e = kzo_convert(urf, VEL4, VELR, &geom, 0, urf);
      if (e) return e;
      pp[FX] = urf[1]; pp[FY] = urf[2]; pp[FZ] = urf[3];
    }
  } else if (face == 0) { /* xlo */
    memcpy(pp, cell(s->p, s, 0, iy, iz), NV * sizeof(double));
  } else if (face == 2) { /* ylo */
    memcpy(pp, cell(s->p, s, ix, -iy - 1, iz), NV * sizeof(double));
    pp[VY] = -pp[VY]; pp[B2] = -pp[B2]; pp[FY] = -pp[FY];
  } else if (face == 3) { /* yhi */
    memcpy(pp, cell(s->p, s, ix, ny - (iy - ny) - 1, iz), NV * sizeof(double));
    pp[VY] = -pp[VY]; pp[B2] = -pp[B2]; pp[FY] = -pp[FY];
  } else { /* zlo / zhi: periodic (bc.zig:230-235, with its nz < ng pin) */
    long nz = s->grid.nz, iiz = iz;
    if (iz < 0) iiz = iz + nz;
    if (iz > nz - 1) iiz = iz - nz;
    if (nz < s->grid.ng) iiz = 0;
    memcpy(pp, cell(s->p, s, ix, iy, iiz), NV * sizeof(double));
  }
  return KZO_OK;
}

static int set_bc_cell(kzo_sim *s, long ix, long iy, long iz, int face) {
  double pp[NV], uu[NV];
  int e = puffy_bc(s, ix, iy, iz, face, pp);
  if (e) return e;
  kzo_geo g;
  geo_c(s, ix, iy, iz, &g);
  kzo_p2u(pp, uu, &g, s->ph.gam);
  memcpy(cell(s->p, s, ix, iy, iz), pp, sizeof pp);
  memcpy(cell(s->u, s, ix, iy, iz), uu, sizeof uu);
  return KZO_OK;
}

static void p2u_cell(kzo_sim *s, long ix, long iy, long iz) {
  kzo_geo g;
  geo_c(s, ix, iy, iz, &g);
  kzo_p2u(cell(s->p, s, ix, iy, iz), cell(s->u, s, ix, iy, iz), &g, s->ph.gam);
}
static void copy_p(kzo_sim *s, long dx, long dy, long sx_, long sy_) {
  memcpy(cell(s->p, s, dx, dy, 0), cell(s->p, s, sx_, sy_, 0), NV * sizeof(double));
}
static void avg_p(kzo_sim *s, long dx, long dy, long ax, long ay, long bx, long by) {
  double *pa = cell(s->p, s, ax, ay, 0), *pb = cell(s->p, s, bx, by, 0), pp[NV];
  for (int iv = 0; iv < NV; ++iv) pp[iv] = 0.5 * (pa[iv] + pb[iv]);
  memcpy(cell(s->p, s, dx, dy, 0), pp, sizeof pp);
}

/* bc.zig:286-377 fillCorners2d (bc_x, bc_y not periodic) */
static void fill_corners_2d(kzo_sim *s) {
  long nx = s->grid.nx, ny = s->grid.ny, ng = s->grid.ng;
  for (long i = 0; i < ng - 1; ++i) {
    copy_p(s, -ng + i, -1, -ng + i, 0); p2u_cell(s, -ng + i, -1, 0);
    copy_p(s, -1, -ng + i, 0, -ng + i); p2u_cell(s, -1, -ng + i, 0);
  }
  avg_p(s, -1, -1, -1, 0, 0, -1); p2u_cell(s, -1, -1, 0);
  avg_p(s, -2, -2, -2, -1, -1, -2); p2u_cell(s, -2, -2, 0);
  for (long i = 0; i < ng - 1; ++i) {
    copy_p(s, -ng + i, ny, -ng + i, ny - 1); p2u_cell(s, -ng + i, ny, 0);
    copy_p(s, -1, ny + i + 1, 0, ny + i + 1); p2u_cell(s, -1, ny + i + 1, 0);
  }
  avg_p(s, -1, ny, -1, ny - 1, 0, ny); p2u_cell(s, -1, ny, 0);
  avg_p(s, -2, ny + 1, -2, ny, -1, ny + 1); p2u_cell(s, -2, ny + 1, 0);
  for (long i = 0; i < ng - 1; ++i) {
    copy_p(s, nx + i + 1, -1, nx + i + 1, 0); p2u_cell(s, nx + i + 1, -1, 0);
    copy_p(s, nx, -ng + i, nx - 1, -ng + i); p2u_cell(s, nx, -ng + i, 0);
  }
  avg_p(s, nx, -1, nx - 1, -1, nx, 0); p2u_cell(s, nx, -1, 0);
  avg_p(s, nx + 1, -2, nx, -2, nx + 1, -1); p2u_cell(s, nx + 1, -2, 0);
  for (long i = 0; i < ng - 1; ++i) {
    copy_p(s, nx + i + 1, ny, nx + i + 1, ny - 1); p2u_cell(s, nx + i + 1, ny, 0);
    copy_p(s, nx, ny + i + 1, nx - 1, ny + i + 1); p2u_cell(s, nx, ny + i + 1, 0);
  }
  avg_p(s, nx, ny, nx - 1, ny, nx, ny - 1); p2u_cell(s, nx, ny, 0);
  avg_p(s, nx + 1, ny + 1, nx, ny + 1, nx + 1, ny); p2u_cell(s, nx + 1, ny + 1, 0);
}

/* bc.zig:378-520 fillCorners3d (serial: every z face is physical): the 12
 * edge tubes (along z, then y, then x), each filled at every slice of its
 * through-axis over the full ghost span, then the 8 corner cubes copied from
 * the nearest domain corner, which overwrite the shared cells last. */
static void copy3(kzo_sim *s, const long d[3], const long a[3]) {
  memcpy(cell(s->p, s, d[0], d[1], d[2]), cell(s->p, s, a[0], a[1], a[2]), NV * sizeof(double));
  p2u_cell(s, d[0], d[1], d[2]);
}
static void avg3(kzo_sim *s, const long d[3], const long a[3], const long b[3]) {
  double *pa = cell(s->p, s, a[0], a[1], a[2]), *pb = cell(s->p, s, b[0], b[1], b[2]), pp[NV];
  for (int iv = 0; iv < NV; ++iv) pp[iv] = 0.5 * (pa[iv] + pb[iv]);
  memcpy(cell(s->p, s, d[0], d[1], d[2]), pp, sizeof pp);
  p2u_cell(s, d[0], d[1], d[2]);
}
static void set3(long o[3], int ax, long av, int bx, long bv, int cx, long cv) { o[ax] = av; o[bx] = bv; o[cx] = cv; }
static void edge_tube(kzo_sim *s, int pa, int a_hi, int pb, int b_hi, int pc) {
  long ng = s->grid.ng;
  long dims[3] = {s->grid.nx, s->grid.ny, s->grid.nz};
  long na = dims[pa], nb = dims[pb], nc = dims[pc];
  long a_edge = a_hi ? na - 1 : 0, a_g1 = a_hi ? na : -1, a_g2 = a_hi ? na + 1 : -2;
  long b_edge = b_hi ? nb - 1 : 0, b_g1 = b_hi ? nb : -1, b_g2 = b_hi ? nb + 1 : -2;
  for (long c = -ng; c < nc + ng; ++c) {
    long d[3], x[3], y[3];
    for (long i = 0; i < ng - 1; ++i) {
      long a_out = a_hi ? na + i + 1 : -ng + i, b_out = b_hi ? nb + i + 1 : -ng + i;
      set3(d, pa, a_out, pb, b_g1, pc, c); set3(x, pa, a_out, pb, b_edge, pc, c); copy3(s, d, x);
      set3(d, pa, a_g1, pb, b_out, pc, c); set3(x, pa, a_edge, pb, b_out, pc, c); copy3(s, d, x);
    }
    set3(d, pa, a_g1, pb, b_g1, pc, c); set3(x, pa, a_g1, pb, b_edge, pc, c); set3(y, pa, a_edge, pb, b_g1, pc, c);
    avg3(s, d, x, y);
    set3(d, pa, a_g2, pb, b_g2, pc, c); set3(x, pa, a_g2, pb, b_g1, pc, c); set3(y, pa, a_g1, pb, b_g2, pc, c);
    avg3(s, d, x, y);
  }
}
static void corner_cube(kzo_sim *s, int x_hi, int y_hi, int z_hi) {
  long ng = s->grid.ng, nx = s->grid.nx, ny = s->grid.ny, nz = s->grid.nz;
  long src[3] = {x_hi ? nx - 1 : 0, y_hi ? ny - 1 : 0, z_hi ? nz - 1 : 0};
  for (long ix = x_hi ? nx : -ng; ix < (x_hi ? nx + ng : 0); ++ix)
    for (long iy = y_hi ? ny : -ng; iy < (y_hi ? ny + ng : 0); ++iy)
      for (long iz = z_hi ? nz : -ng; iz < (z_hi ? nz + ng : 0); ++iz) {
        long d[3] = {ix, iy, iz};
        copy3(s, d, src);
      }
}
static void fill_corners_3d(kzo_sim *s) {
  for (int k = 0; k < 4; ++k) edge_tube(s, 0, k >> 1, 1, k & 1, 2);
  for (int k = 0; k < 4; ++k) edge_tube(s, 0, k >> 1, 2, k & 1, 1);
  for (int k = 0; k < 4; ++k) edge_tube(s, 1, k >> 1, 2, k & 1, 0);
  for (int k = 0; k < 8; ++k) corner_cube(s, k & 1, (k >> 1) & 1, (k >> 2) & 1);
}

/* bc.zig:40-73 setBc */
int kzo_sim_set_bc(kzo_sim *s, int ifinit) {
  (void)ifinit;
  long nx = s->grid.nx, ny = s->grid.ny, nz = s->grid.nz;
  for (long iz = 0; iz < nz; ++iz)
    for (long iy = 0; iy < ny; ++iy)
      for (long i = 1; i <= s->ngx; ++i) {
        int e = set_bc_cell(s, -i, iy, iz, 0);
        if (e) return e;
        e = set_bc_cell(s, nx - 1 + i, iy, iz, 1);
        if (e) return e;
      }
  for (long iz = 0; iz < nz; ++iz)
    for (long ix = 0; ix < nx; ++ix)
      for (long i = 1; i <= s->ngy; ++i) {
        int e = set_bc_cell(s, ix, -i, iz, 2);
        if (e) return e;
        e = set_bc_cell(s, ix, ny - 1 + i, iz, 3);
        if (e) return e;
      }
  for (long iy = 0; iy < ny; ++iy)
    for (long ix = 0; ix < nx; ++ix)
      for (long i = 1; i <= s->ngz; ++i) {
        int e = set_bc_cell(s, ix, iy, -i, 4);
        if (e) return e;
        e = set_bc_cell(s, ix, iy, nz - 1 + i, 5);
        if (e) return e;
      }
  if (nz == 1) fill_corners_2d(s);
  else fill_corners_3d(s);
  return KZO_OK;
}

/* ---- p2u over the domain (Sim.initCell's p2u) ---- */
void kzo_sim_p2u_domain(kzo_sim *s) {
  for (long iz = 0; iz < s->grid.nz; ++iz)
    for (long iy = 0; iy < s->grid.ny; ++iy)
      for (long ix = 0; ix < s->grid.nx; ++ix) p2u_cell(s, ix, iy, iz);
}

/* ---- wavespeeds (sim.zig:747-886) ---- */
static void save_wavespeeds(kzo_sim *s, long ix, long iy, long iz, const double aaa[6], const double rad0[6],
                            const double radl[6], double *tmax, double *tmin) {
  for (int d = 0; d < 3; ++d) {
    *scp(s, ahd_l[d], ix, iy, iz) = aaa[2 * d];
    *scp(s, ahd_r[d], ix, iy, iz) = aaa[2 * d + 1];
  }
  double ax = kzm_fmax(kzm_fabs(aaa[0]), kzm_fabs(aaa[1]));
  double ay = kzm_fmax(kzm_fabs(aaa[2]), kzm_fabs(aaa[3]));
  double az = kzm_fmax(kzm_fabs(aaa[4]), kzm_fabs(aaa[5]));
  *scp(s, S_AHDX, ix, iy, iz) = ax;
  *scp(s, S_AHDY, ix, iy, iz) = ay;
  *scp(s, S_AHDZ, ix, iy, iz) = az;
  for (int d = 0; d < 3; ++d) {
    *scp(s, arad_l[d], ix, iy, iz) = radl[2 * d];
    *scp(s, arad_r[d], ix, iy, iz) = radl[2 * d + 1];
    *scp(s, arad_m[d], ix, iy, iz) = kzm_fmax(kzm_fabs(radl[2 * d]), kzm_fabs(radl[2 * d + 1]));
  }
  double wsx = kzm_fmax(ax, kzm_fmax(kzm_fabs(rad0[0]), kzm_fabs(rad0[1])));
  double wsy = kzm_fmax(ay, kzm_fmax(kzm_fabs(rad0[2]), kzm_fabs(rad0[3])));
  double wsz = kzm_fmax(az, kzm_fmax(kzm_fabs(rad0[4]), kzm_fabs(rad0[5])));
  if (!(ix >= 0 && ix < s->grid.nx && iy >= 0 && iy < s->grid.ny && iz >= 0 && iz < s->grid.nz)) return;
  double dx = csize(s, ix, 0), dy = csize(s, iy, 1), dz = csize(s, iz, 2);
  double tsd = s->grid.nz > 1 ? wsx / dx + wsy / dy + wsz / dz : wsx / dx + wsy / dy; /* ny > 1 */
  tsd /= s->ph.tsteplim;
  *scp(s, S_TSTEPDEN, ix, iy, iz) = tsd;
  if (tsd > *tmax) *tmax = tsd;
  if (tsd < *tmin) *tmin = tsd;
}

static int calc_wavespeeds(kzo_sim *s) {
  int active[3] = {s->grid.nx > 1, s->grid.ny > 1, s->grid.nz > 1};
  long lx = active[0] ? 1 : 0, ly = active[1] ? 1 : 0, lz = active[2] ? 1 : 0;
  double tmax = -1.0 / 0.0, tmin = 1.0 / 0.0;
  const long wy = s->grid.ny + 2 * ly, nrow = (s->grid.nz + 2 * lz) * wy;
#pragma omp parallel for num_threads(kzo_nthreads) schedule(dynamic, 1) reduction(max : tmax) reduction(min : tmin)
  for (long row = 0; row < nrow; ++row)
      for (long ix = -lx; ix < s->grid.nx + lx; ++ix) {
        long iz = row / wy - lz, iy = row % wy - ly;
        const double *pp = cell(s->p, s, ix, iy, iz);
        kzo_geo g;
        geo_c(s, ix, iy, iz, &g);
        double aaa[6], rad0[6] = {0}, radl[6] = {0}, aval[12];
        kzo_gas_wavespeeds_lr(pp, &g, s->ph.gam, active, aaa);
        double tautot[3] = {0, 0, 0};
        if (s->ph.has_opac) {
          double chi = kzo_calc_chi_slim(pp, &g, s->ph.gam, &s->ph.opac);
          long idx3[3] = {ix, iy, iz};
          for (int d = 0; d < 3; ++d) {
            double sg = kzm_sqrt(g.g[d + 1][d + 1]);
            double dxp = kzm_fmax(csize(s, idx3[d], d) * sg, csize(s, idx3[d] + 1, d) * sg);
            tautot[d] = chi * dxp;
          }
        }
        if (s->ph.dampradwavespeednearaxis > 0) {
          long nc = s->ph.dampradwavespeednearaxis;
          if (iy < nc || iy >= s->grid.ny - nc) tautot[0] = tautot[1] = tautot[2] = 0;
        }
        kzo_calc_rad_wavespeeds(pp, &g, tautot, active, aval);
        for (int i = 0; i < 6; ++i) { rad0[i] = aval[i]; radl[i] = aval[6 + i]; }
        for (int d = 0; d < 3; ++d) {
          if (radl[2 * d] > 0.0) radl[2 * d] = 0.0;
          if (radl[2 * d + 1] < 0.0) radl[2 * d + 1] = 0.0;
          if (rad0[2 * d] > 0.0) rad0[2 * d] = 0.0;
          if (rad0[2 * d + 1] < 0.0) rad0[2 * d + 1] = 0.0;
        }
        save_wavespeeds(s, ix, iy, iz, aaa, rad0, radl, &tmax, &tmin);
      }
  if (tmax > s->tstepdenmax) s->tstepdenmax = tmax;
  if (tmin < s->tstepdenmin) s->tstepdenmin = tmin;
  return KZO_OK;
}

/* sim.zig:693-718 initTimestepGuess (2D) */
void kzo_sim_init_timestep_guess(kzo_sim *s) {
  const double ws = 10000.0;
  double tsd = s->grid.nz > 1 ? ws / s->min_dx + ws / s->min_dy + ws / s->min_dz : ws / s->min_dx + ws / s->min_dy;
  tsd /= s->ph.tsteplim;
  s->tstepdenmax = tsd;
  s->tstepdenmin = tsd;
  for (long iz = 0; iz < s->grid.nz; ++iz)
    for (long iy = 0; iy < s->grid.ny; ++iy)
      for (long ix = 0; ix < s->grid.nx; ++ix) {
        *scp(s, S_TSTEPDEN, ix, iy, iz) = tsd;
        *scp(s, S_CELL_DT, ix, iy, iz) = 1.0 / tsd;
      }
}

static void save_timesteps(kzo_sim *s) {
  for (long iz = 0; iz < s->grid.nz; ++iz)
    for (long iy = 0; iy < s->grid.ny; ++iy)
      for (long ix = 0; ix < s->grid.nx; ++ix)
        *scp(s, S_CELL_DT, ix, iy, iz) = 1.0 / *scp(s, S_TSTEPDEN, ix, iy, iz);
}

/* ---- u2p pass + fixups (sim.zig:924-1133) ---- */
static void u2p_rows(kzo_sim *s) {
  const long ny = s->grid.ny, nrow = s->grid.nz * ny;
  uint64_t nguard = 0;
#pragma omp parallel for num_threads(kzo_nthreads) schedule(dynamic, 1) reduction(+ : nguard)
  for (long row = 0; row < nrow; ++row) {
    long iz = row / ny, iy = row % ny;
    int polar = polar_row(s, iy);
    for (long ix = 0; ix < s->grid.nx; ++ix) {
      double uu[NV], pp[NV];
      memcpy(uu, cell(s->u, s, ix, iy, iz), sizeof uu);
      memcpy(pp, cell(s->p, s, ix, iy, iz), sizeof pp);
      *flagp(s, KZ_FLAG_ENTROPY, ix, iy, iz) = 0;
      *flagp(s, KZ_FLAG_ENTROPY2, ix, iy, iz) = 0;
      kzo_geo g;
      geo_c(s, ix, iy, iz, &g);
      if (polar) {
        double gi = 1.0 / g.gdet;
        pp[B1] = uu[B1] * gi; pp[B2] = uu[B2] * gi; pp[B3] = uu[B3] * gi;
        memcpy(cell(s->p, s, ix, iy, iz), pp, sizeof pp);
        *flagp(s, KZ_FLAG_HD_FIXUP, ix, iy, iz) = 0;
        *flagp(s, KZ_FLAG_RAD_FIXUP, ix, iy, iz) = 0;
        continue;
      }
      kz_u2p_result r = kzo_u2p_mhd(uu, pp, &g, s->ph.gam, &s->ph.floors);
      if (r.entropy_used) *flagp(s, KZ_FLAG_ENTROPY, ix, iy, iz) = 1;
      kz_u2prad_result rr = kzo_u2p_rad(uu, pp, &g, &s->ph.rad);
      kzo_check_floors_mhd(pp, &g, s->ph.gam, &s->ph.floors, &nguard);
      kzo_check_floors_rad(pp, &g, &s->ph.rad);
      memcpy(cell(s->p, s, ix, iy, iz), pp, sizeof pp);
      *flagp(s, KZ_FLAG_HD_FIXUP, ix, iy, iz) = r.fixup ? 1 : 0;
      *flagp(s, KZ_FLAG_RAD_FIXUP, ix, iy, iz) = rr.corrected ? -1 : 0;
    }
  }
  s->n_guard += nguard;
}

static void cell_fixup(kzo_sim *s, int which) {
  int enabled = which == KZ_FLAG_HD_FIXUP ? s->ph.do_fixups
                : which == KZ_FLAG_RAD_FIXUP ? (s->ph.do_fixups && s->ph.do_u2prad_fixups)
                                             : (s->ph.do_fixups && s->ph.do_radimp_fixups);
  if (!enabled) return;
  long nx = s->grid.nx, ny = s->grid.ny, nz = s->grid.nz;
  int any = 0;
  for (long iz = 0; iz < nz && !any; ++iz)
    for (long iy = 0; iy < ny && !any; ++iy)
      for (long ix = 0; ix < nx; ++ix)
        if (*flagp(s, which, ix, iy, iz) != 0) { any = 1; break; }
  if (!any) return;
  size_t fb = (size_t)s->ncell * NV * sizeof(double);
  memcpy(s->u_bak, s->u, fb);
  memcpy(s->p_bak, s->p, fb);
  for (long iz = 0; iz < nz; ++iz)
  for (long iy = 0; iy < ny; ++iy) {
    if (polar_row(s, iy)) continue;
    for (long ix = 0; ix < nx; ++ix) {
      if (*flagp(s, which, ix, iy, iz) == 0) continue;
      double ppn[6][NV];
      int in_n = 0;
      long nb[6][3] = {{ix - 1, iy, iz}, {ix + 1, iy, iz}, {ix, iy - 1, iz}, {ix, iy + 1, iz}, {ix, iy, iz - 1}, {ix, iy, iz + 1}};
      for (int k = 0; k < 6; ++k) {
        if (nb[k][0] < 0 || nb[k][0] >= nx || nb[k][1] < 0 || nb[k][1] >= ny || nb[k][2] < 0 || nb[k][2] >= nz) continue;
        if (*flagp(s, which, nb[k][0], nb[k][1], nb[k][2]) != 0) continue;
        memcpy(ppn[in_n], cell(s->p, s, nb[k][0], nb[k][1], nb[k][2]), NV * sizeof(double));
        in_n += 1;
      }
      int enough = (nz == 1 && ny == 1 && in_n >= 1) || (nz == 1 && in_n >= 2) || (ny == 1 && in_n >= 2) || in_n >= 3;
      if (!enough) continue;
      double pp[NV], uu[NV];
      memcpy(pp, cell(s->p, s, ix, iy, iz), sizeof pp);
      for (int iv = 0; iv < NV; ++iv) {
        int fixthis;
        if (which == KZ_FLAG_HD_FIXUP) fixthis = iv != RHO && iv < B1;
        else if (which == KZ_FLAG_RAD_FIXUP) fixthis = iv >= EE && iv <= FZ;
        else fixthis = iv != RHO && (iv < B1 || (iv >= EE && iv <= FZ));
        if (fixthis) {
          double sum = 0;
          for (int k = 0; k < in_n; ++k) sum += ppn[k][iv];
          pp[iv] = sum / (double)in_n;
        }
      }
      kzo_geo g;
      geo_c(s, ix, iy, iz, &g);
      kzo_p2u(pp, uu, &g, s->ph.gam);
      memcpy(cell(s->u_bak, s, ix, iy, iz), uu, sizeof uu);
      memcpy(cell(s->p_bak, s, ix, iy, iz), pp, sizeof pp);
    }
  }
  memcpy(s->u, s->u_bak, fb);
  memcpy(s->p, s->p_bak, fb);
}

int kzo_sim_calc_u2p(kzo_sim *s) {
  u2p_rows(s);
  cell_fixup(s, KZ_FLAG_HD_FIXUP);
  cell_fixup(s, KZ_FLAG_RAD_FIXUP);
  return kzo_sim_set_bc(s, 0);
}

/* ---- metric source (sim.zig:1140-1175) ---- */
static void metric_source(kzo_sim *s, long ix, long iy, long iz, double ss[NV]) {
  kzo_geo g;
  geo_c(s, ix, iy, iz, &g);
  const double *pp = cell(s->p, s, ix, iy, iz);
  double gdetu = g.gdet;
  for (int i = 0; i < NV; ++i) ss[i] = 0;
  double tij[4][4], t[4][4], rup[4][4], rij[4][4];
  kzo_calc_tij(pp, &g, s->ph.gam, tij);
  kzo_lower_second(tij, &g, t);
  const double *kr = s->kris + cidx(s, ix, iy, iz) * 64;
  static const int rows[4] = {UU, VX, VY, VZ};
  static const int rrows[4] = {EE, FX, FY, FZ};
  kzo_calc_rij(pp, &g, rup);
  kzo_lower_second(rup, &g, rij);
  for (int k = 0; k < 4; ++k)
    for (int l = 0; l < 4; ++l)
      for (int nu = 0; nu < 4; ++nu) {
        ss[rows[nu]] += gdetu * t[k][l] * kr[l * 16 + nu * 4 + k];
        ss[rrows[nu]] += gdetu * rij[k][l] * kr[l * 16 + nu * 4 + k];
      }
}

/* ---- BL sidecar (metric/precompute.zig:361-380, :457-475) ---- */
void kzo_sim_set_bl_tables(kzo_sim *s, const double *bl) { s->bl = bl; }
static const double *bl_rec(const kzo_sim *s, long ix, long iy, long iz) { return s->bl + cidx(s, ix, iy, iz) * BL_REC; }
static void bl_geo(const kzo_sim *s, long ix, long iy, long iz, kzo_geo *g) {
  const double *r = bl_rec(s, ix, iy, iz);
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) { g->g[i][j] = r[i * 4 + j]; g->G[i][j] = r[16 + i * 4 + j]; }
  g->alpha = r[32];
  g->gdet = r[33];
  g->gttpert = 0;
  g->x1 = r[34];
}
static int is_corner(const kzo_sim *s, long ix, long iy, long iz) {  /* sim.zig:660-666 */
  int n = 0;
  if (ix < 0 || ix >= s->grid.nx) n += 1;
  if (iy < 0 || iy >= s->grid.ny) n += 1;
  if (iz < 0 || iz >= s->grid.nz) n += 1;
  return n >= 2;
}
static double ccoord(const kzo_sim *s, int idim, long i) {
  return idim == 1 ? xc_(s, i) : (idim == 2 ? yc_(s, i) : zc_(s, i));
}
/* math/misc.zig:9-12 stepFunction */
static double step_function(double x, double x9) {
  double k = 1.47222 / x9;
  return 1.0 / (1.0 + kzm_exp(-2.0 * k * x));
}
/* relele.convertBoth(VELR) for the velocity slots starting at `base` */
static void conv_both_velr(const double *pp, int base, const kzo_geo *g, double con[4], double cov[4]) {
  double v[4] = {0.0, pp[base], pp[base + 1], pp[base + 2]};
  kzo_velr_to_vel4(v, g, con);
  kzo_lower(con, g, cov);
}

/* ---- radiative shear viscosity (sim/rijvisc.zig, physics/radvisc.zig) ---- */
/* radvisc.zig:72-131 shearFromGradients: sigma_ij from the assembled gradients */
static void shear_from_gradients(const double du[4][4], const double du2[4][4], const double ucon[4],
                                 const double ucov[4], const kzo_geo *gm, const double *kr, double sig[4][4]) {
  double dcu[4][4], dcudiag[4];
  for (int i = 0; i < 4; ++i) {
    for (int j = 0; j < 4; ++j) {
      double krsum = 0;
      for (int k = 0; k < 4; ++k) krsum += kr[k * 16 + i * 4 + j] * ucov[k];
      dcu[i][j] = du[i][j] - krsum;
    }
    double krsum = 0;
    for (int k = 0; k < 4; ++k) krsum += kr[i * 16 + i * 4 + k] * ucon[k];
    dcudiag[i] = du2[i][i] + krsum;
  }
  double theta = 0;
  for (int i = 0; i < 4; ++i) theta += dcudiag[i];
  double p11[4][4], p21[4][4];
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) {
      p11[i][j] = gm->g[i][j] + ucov[i] * ucov[j];
      p21[i][j] = (i == j ? 1.0 : 0.0) + ucon[i] * ucov[j];
    }
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) sig[i][j] = 0.0;
  for (int i = 1; i < 4; ++i)
    for (int j = 1; j < 4; ++j) {
      double sum1 = 0, sum2 = 0;
      for (int k = 0; k < 4; ++k) {
        sum1 += dcu[i][k] * p21[k][j];
        sum2 += dcu[j][k] * p21[k][i];
      }
      sig[i][j] = 0.5 * (sum1 + sum2) - 1.0 / 3.0 * theta * p11[i][j];
    }
  for (int i = 1; i < 4; ++i) {
    double v = -1.0 / ucon[0] * (ucon[1] * sig[1][i] + ucon[2] * sig[2][i] + ucon[3] * sig[3][i]);
    sig[i][0] = v;
    sig[0][i] = v;
  }
  sig[0][0] = -1.0 / ucon[0] * (ucon[1] * sig[1][0] + ucon[2] * sig[2][0] + ucon[3] * sig[3][0]);
}

/* rijvisc.zig:46-172 calcShearLab (radiation-frame velocities FX..FZ) */
static void calc_shear_lab(const kzo_sim *s, long ix, long iy, long iz, double sig[4][4]) {
  const double *pp0 = s->p + cidx(s, ix, iy, iz) * NV;
  kzo_geo geom;
  geo_c(s, ix, iy, iz, &geom);
  double du[4][4] = {{0}}, du2[4][4] = {{0}};
  double ucon[4], ucov[4];
  conv_both_velr(pp0, FX, &geom, ucon, ucov);
  long nd[3] = {s->grid.nx, s->grid.ny, s->grid.nz};
  long nx = nd[0], ny = nd[1], nz = nd[2];
  for (int idim = 1; idim < 4; ++idim) {
    int dim = idim - 1;
    if (nd[dim] <= 1) continue;
    long cm[3] = {ix, iy, iz}, cp[3] = {ix, iy, iz};
    cm[dim] -= 1;
    cp[dim] += 1;
    const double *ppm1 = s->p + cidx(s, cm[0], cm[1], cm[2]) * NV;
    const double *ppp1 = s->p + cidx(s, cp[0], cp[1], cp[2]) * NV;
    kzo_geo gm, gp;
    geo_c(s, cm[0], cm[1], cm[2], &gm);
    geo_c(s, cp[0], cp[1], cp[2], &gp);
    double umc[4], umv[4], upc[4], upv[4];
    conv_both_velr(ppm1, FX, &gm, umc, umv);
    conv_both_velr(ppp1, FX, &gp, upc, upv);
    double xm = ccoord(s, idim, cm[dim]), xc = ccoord(s, idim, dim == 0 ? ix : (dim == 1 ? iy : iz)),
           xp = ccoord(s, idim, cp[dim]);
    for (int i = 0; i < 4; ++i) {
      double dc = (upv[i] - umv[i]) / (xp - xm);
      double dr = (upv[i] - ucov[i]) / (xp - xc);
      double dl = (ucov[i] - umv[i]) / (xc - xm);
      double dc2 = (upc[i] - umc[i]) / (xp - xm);
      double dr2 = (upc[i] - ucon[i]) / (xp - xc);
      double dl2 = (ucon[i] - umc[i]) / (xc - xm);
      if ((ix < 0 && iy == 0 && iz == 0 && idim != 1) || (iy < 0 && ix == 0 && iz == 0 && idim != 2) ||
          (iz < 0 && ix == 0 && iy == 0 && idim != 3)) {
        du[i][idim] = dr; du2[i][idim] = dr2;
      } else if ((ix < 0 && iy == ny - 1 && iz == nz - 1 && idim != 1) ||
                 (iy < 0 && ix == nx - 1 && iz == nz - 1 && idim != 2) ||
                 (iz < 0 && ix == nx - 1 && iy == ny - 1 && idim != 3)) {
        du[i][idim] = dl; du2[i][idim] = dl2;
      } else if ((ix >= nx && iy == 0 && iz == 0 && idim != 1) || (iy >= ny && ix == 0 && iz == 0 && idim != 2) ||
                 (iz >= nz && ix == 0 && iy == 0 && idim != 3)) {
        du[i][idim] = dr; du2[i][idim] = dr2;
      } else if ((ix >= nx && iy == ny - 1 && iz == nz - 1 && idim != 1) ||
                 (iy >= ny && ix == nx - 1 && iz == nz - 1 && idim != 2) ||
                 (iz >= nz && ix == nx - 1 && iy == ny - 1 && idim != 3)) {
        du[i][idim] = dl; du2[i][idim] = dl2;
      } else {
        du[i][idim] = dc; du2[i][idim] = dc2;
      }
    }
  }
  shear_from_gradients((const double(*)[4])du, (const double(*)[4])du2, ucon, ucov, &geom,
                       s->kris + cidx(s, ix, iy, iz) * 64, sig);
}

/* radvisc.zig:154-175 viscCoeff */
static double visc_coeff(double chi, double mindx, double r_bl, double rhor, double alpha, double global_dt) {
  double mfp = 1.0 / chi;
  double rmin = 1.2 * rhor;
  double mfplim = r_bl;
  if (mfp > mfplim || chi < 1.0e-80) mfp = mfplim;
  if (mfp < 0 || !kzm_isfinite(mfp)) mfp = 0;
  mfp *= step_function(r_bl - rmin, 0.2 * rmin);
  if (r_bl <= 1.0 * rhor) mfp = 0;
  double nu = alpha * mfp;
  double nulimit = mindx * mindx / 2.0 / global_dt / 2.0;
  if (nu > nulimit) nu = nulimit;
  return nu;
}

/* rijvisc.zig:175-268 calcRadViscCoeff + calcRijVisc: R^i_j viscous */
static void calc_rij_visc(const kzo_sim *s, long ix, long iy, long iz, double global_dt, double out[16]) {
  const double *pp = s->p + cidx(s, ix, iy, iz) * NV;
  kzo_geo geom;
  geo_c(s, ix, iy, iz, &geom);
  double sl[4][4], su[4][4] = {{0}};
  calc_shear_lab(s, ix, iy, iz, sl);
  /* relele.zig:115-128 raiseBoth */
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j)
      for (int k = 0; k < 4; ++k)
        for (int l = 0; l < 4; ++l) su[i][j] += sl[k][l] * geom.G[i][k] * geom.G[j][l];
  double chi = kzo_calc_chi_slim(pp, &geom, s->ph.gam, &s->ph.opac);
  double dx[3] = {csize(s, ix, 0) * kzm_sqrt(geom.g[1][1]), csize(s, iy, 1) * kzm_sqrt(geom.g[2][2]),
                  csize(s, iz, 2) * kzm_sqrt(geom.g[3][3])};
  long ny = s->grid.ny, nz = s->grid.nz;
  double mindx = (ny == 1 && nz == 1) ? dx[0]
                 : (nz == 1)          ? kzm_fmin(dx[0], dx[1])
                 : (ny == 1)          ? kzm_fmin(dx[0], dx[2])
                                      : kzm_fmin(dx[0], kzm_fmin(dx[1], dx[2]));
  double nu = visc_coeff(chi, mindx, bl_rec(s, ix, iy, iz)[34], s->rhor, s->ph.radvisc.alpha, global_dt);
  double rv[4][4], lo[4][4];
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) rv[i][j] = -2.0 * nu * pp[EE] * su[i][j];
  kzo_lower_second((const double(*)[4])rv, &geom, lo);
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) out[i * 4 + j] = lo[i][j];
}

/* rijvisc.zig:266-311 calcRijViscTotal: domain + one-cell ring, corners skipped */
static void calc_rij_visc_total(kzo_sim *s, double global_dt) {
  memset(s->rijvisc, 0, (size_t)s->ncell * 16 * sizeof(double));
  long nx = s->grid.nx, ny = s->grid.ny, nz = s->grid.nz;
  long xl = nx > 1 ? 1 : 0, yl = ny > 1 ? 1 : 0, zl = nz > 1 ? 1 : 0;
  for (long iz = -zl; iz < nz + zl; ++iz)
    for (long iy = -yl; iy < ny + yl; ++iy)
      for (long ix = -xl; ix < nx + xl; ++ix) {
        if (is_corner(s, ix, iy, iz)) continue;
        calc_rij_visc(s, ix, iy, iz, global_dt, s->rijvisc + cidx(s, ix, iy, iz) * 16);
      }
}

/* rijvisc.zig:313-353 faceAvg + addRadViscFlux (radvisc.zig:201-244): the
 * face-averaged, velocity-damped viscous flux added to the M1 rows */
static void add_rad_visc_flux(const kzo_sim *s, int dim, const long fc[3], const double *pp, const kzo_geo *g,
                              double *ff) {
  const double *a = s->rijvisc + cidx(s, fc[0], fc[1], fc[2]) * 16;
  long c[3] = {fc[0], fc[1], fc[2]};
  c[dim] -= 1;
  const double *b = s->rijvisc + cidx(s, c[0], c[1], c[2]) * 16;
  double rv[4][4];
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) rv[i][j] = 0.5 * (a[i * 4 + j] + b[i * 4 + j]);
  double uu[NV];
  kzo_p2u(pp, uu, g, s->ph.gam);
  int y_active = s->grid.ny > 1, z_active = s->grid.nz > 1;
  double gdetu = g->gdet;
  double vel[3] = {0, 0, 0};
  for (int id = 0; id < 3; ++id)
    for (int ii = 1; ii < 4; ++ii) {
      if (ii == 2 && !y_active) continue;
      if (ii == 3 && !z_active) continue;
      if (kzm_fabs(uu[EE + ii]) < 1.0e-10 * kzm_fabs(uu[EE])) continue;
      vel[id] = rv[id + 1][ii] / (uu[EE + ii] / gdetu) * kzm_sqrt(g->g[id + 1][id + 1]);
    }
  double maxv = kzm_fabs(vel[dim]);
  double dampfac = 1.0;
  if (maxv > s->ph.radvisc.maxvel) dampfac = s->ph.radvisc.maxvel / maxv;
  for (int nu = 0; nu < 4; ++nu) ff[EE + nu] += gdetu * dampfac * rv[dim + 1][nu];
}

/* ---- mean-field dynamo (sim/dynamo.zig) ---- */
static const double KZO_PI = 3.14159265358979323846;

/* dynamo.zig:226-256 calcScaleHeight (single rank): density-weighted RMS of
 * pi/2 - theta per radial column; column 0 keeps the raw sum (the reference's
 * gix > 0 quirk) */
static void calc_scale_height(kzo_sim *s) {
  long nx = s->grid.nx, ny = s->grid.ny, nz = s->grid.nz;
  for (long ix = 0; ix < nx; ++ix) {
    double sigma = 0, sth = 0;
    for (long iy = 0; iy < ny; ++iy)
      for (long iz = 0; iz < nz; ++iz) {
        double gd = s->g[cidx(s, ix, iy, iz) * 20 + 19];
        double th = bl_rec(s, ix, iy, iz)[35];
        double rho = s->p[cidx(s, ix, iy, iz) * NV + RHO];
        double dth = KZO_PI / 2.0 - th;
        sigma += rho * gd;
        sth += rho * gd * dth * dth;
      }
    s->scaleth[ix] = ix > 0 ? kzm_sqrt(sth / sigma) : sth;
  }
}

/* dynamo.zig:67-73 dampBphi: the DAMPBETA increment of B^phi */
static double damp_bphi(double alphabeta, double facradius, double faczh, double dt_over_pk, double beta,
                        double betasat, double bphi) {
  double dbphi = -alphabeta * facradius * faczh * dt_over_pk * kzm_fmax(0.0, beta - betasat) / betasat * bphi;
  if ((dbphi + bphi) * bphi < 0.0) dbphi = -bphi;
  return dbphi;
}

/* dynamo.zig:104-151 dynamoDeltaA: returns 0 and leaves *bphi_out = bphi
 * when r < 1.0001 r_horizon */
static double dynamo_delta_a(const kz_dynamo_params *dp, double a, double dt, double rhor, double risco, double r,
                             double th, double bphi, double bsq, double angle, double prermhd, double gg33,
                             double scaleth, double *bphi_out) {
  *bphi_out = bphi;
  if (r < 1.0001 * rhor) return 0.0;
  double omk = 1.0 / (a + kzm_sqrt(r * r * r));
  double pk = 2.0 * KZO_PI / omk;
  double facangle = 0;
  if (kzm_isfinite(angle)) {
    if (angle < -1.0) angle = -1.0;
    facangle = kzm_fmax(0.0, (dp->thetaangle - angle) / dp->thetaangle);
  }
  double facradius = step_function(r - 1.0 * risco, 0.1 * risco);
  double beta = 0.5 * bsq / prermhd;
  double hrdtheta;
  if (dp->calchronthego) {
    hrdtheta = scaleth;
    if (hrdtheta > 0.9 * KZO_PI / 2.0) hrdtheta = 0.9 * KZO_PI / 2.0;
  } else {
    hrdtheta = dp->expectedhr * KZO_PI / 2.0;
  }
  double zh = (KZO_PI / 2.0 - th) / hrdtheta;
  double faczh = kzm_fmax(0.0, 1.0 - zh * zh);
  double facmagnetization = faczh;
  double effalpha = dp->alphadynamo;
  if (dp->alphaflipssign) effalpha = -(KZO_PI / 2.0 - th) / (hrdtheta / 2.0) * dp->alphadynamo;
  double aphi = effalpha * (hrdtheta / (KZO_PI / 2.0)) / 0.4 * dt / pk * r * gg33 * bphi * facradius *
                facmagnetization * facangle;
  if (dp->dampbeta) *bphi_out = bphi + damp_bphi(dp->alphabeta, facradius, faczh, dt / pk, beta, dp->betasaturated, bphi);
  return aphi;
}

/* dynamo.zig:335-368 fieldAngle: -b^r b^phi sqrt(g_rr g_phph)/b^2 in BL */
static int field_angle(const double *pp, const kzo_geo *gm, const kzo_geo *gbl, const double *J, double *angle,
                       double *bsq) {
  double ppbl[NV];
  memcpy(ppbl, pp, sizeof ppbl);
  int e = trans_pmhd(ppbl, gm, gbl, J);
  if (e) return e;
  double ucon[4], ucov[4];
  kzo_ucon_ucov_from_prims(ppbl, gbl, ucon, ucov);
  double B[3] = {ppbl[B1], ppbl[B2], ppbl[B3]}, bcon[4], bcov[4], b2;
  kzo_bcon_bcov_bsq(B, ucon, ucov, gbl, bcon, bcov, &b2);
  double brbphi = kzm_sqrt(gbl->g[1][1] * gbl->g[3][3]) * bcon[1] * bcon[3];
  *angle = -brbphi / b2;
  *bsq = b2;
  return KZO_OK;
}

/* dynamo.zig:370-460 mimicDynamo: Delta A_phi over the domain + ring
 * (corners included), DAMPBETA on p[B3], curl, superpose on the domain */
static int mimic_dynamo(kzo_sim *s, double dt) {
  long nx = s->grid.nx, ny = s->grid.ny, nz = s->grid.nz;
  long xl = nx > 1 ? 1 : 0, yl = ny > 1 ? 1 : 0, zl = nz > 1 ? 1 : 0;
  memset(s->dyn_a, 0, (size_t)s->ncell * sizeof(double));
  const kz_dynamo_params *dp = &s->ph.dynamo_params;
  for (long iz = -zl; iz < nz + zl; ++iz)
    for (long iy = -yl; iy < ny + yl; ++iy)
      for (long ix = -xl; ix < nx + xl; ++ix) {
        kzo_geo geom, gbl;
        geo_c(s, ix, iy, iz, &geom);
        bl_geo(s, ix, iy, iz, &gbl);
        const double *rec = bl_rec(s, ix, iy, iz);
        double *pp = s->p + cidx(s, ix, iy, iz) * NV;
        double angle, bsq;
        int e = field_angle(pp, &geom, &gbl, rec + 36, &angle, &bsq);
        if (e) return e;
        double r = rec[34], th = rec[35];
        if (r < 1.0001 * s->rhor) continue;
        double prermhd = (s->ph.gam - 1.0) * pp[UU];
        double ehat = -kzo_calc_ff_rtt(pp, &geom, 0);
        prermhd += ehat / 3.0;
        long gix = ix < 0 ? 0 : (ix > nx - 1 ? nx - 1 : ix);
        double bphi_new;
        double aphi = dynamo_delta_a(dp, s->ph.bh_spin, dt, s->rhor, s->risco, r, th, pp[B3], bsq, angle, prermhd,
                                     geom.g[3][3], s->scaleth[gix], &bphi_new);
        s->dyn_a[cidx(s, ix, iy, iz)] = aphi;
        if (dp->dampbeta) pp[B3] = bphi_new;
      }
  /* ct.zig:228-264 cornerAverageA of (0, 0, Delta A_phi) */
  for (long iz = 0; iz <= nz; ++iz)
    for (long iy = 0; iy <= ny; ++iy)
      for (long ix = 0; ix <= nx; ++ix) {
        if (nz == 1 && iz > 0) continue;
        double *v = s->vecpot + cidx(s, ix, iy, iz) * 6;
        for (int c = 0; c < 3; ++c) {
#define DA(jx, jy, jz) (c == 2 ? s->dyn_a[cidx(s, jx, jy, jz)] : 0.0)
          if (nz == 1)
            v[c] = 1.0 / 4.0 * (DA(ix, iy, iz) + DA(ix, iy - 1, iz) + DA(ix - 1, iy, iz) + DA(ix - 1, iy - 1, iz));
          else
            v[c] = 1.0 / 8.0 * (DA(ix, iy, iz) + DA(ix, iy - 1, iz) + DA(ix - 1, iy, iz) + DA(ix - 1, iy - 1, iz) +
                                DA(ix, iy, iz - 1) + DA(ix, iy - 1, iz - 1) + DA(ix - 1, iy, iz - 1) +
                                DA(ix - 1, iy - 1, iz - 1));
#undef DA
        }
      }
  /* ct.zig:279-331 calcBfromACore then dynamo.zig:432-460 superpose */
  for (long iz = 0; iz < nz; ++iz)
    for (long iy = 0; iy < ny; ++iy)
      for (long ix = 0; ix < nx; ++ix) {
        double dA[4][4] = {{0}};
        for (int c = 0; c < 3; ++c) {
#define VA(jx, jy, jz) (s->vecpot[cidx(s, jx, jy, jz) * 6 + c])
          if (nz == 1) {
            dA[c + 1][2] = -(VA(ix, iy, iz) - VA(ix, iy + 1, iz) + VA(ix + 1, iy, iz) - VA(ix + 1, iy + 1, iz)) /
                           (2.0 * csize(s, iy, 1));
            dA[c + 1][1] = -(VA(ix, iy, iz) + VA(ix, iy + 1, iz) - VA(ix + 1, iy, iz) - VA(ix + 1, iy + 1, iz)) /
                           (2.0 * csize(s, ix, 0));
            dA[c + 1][3] = 0;
          } else {
            dA[c + 1][1] = (VA(ix + 1, iy, iz) - VA(ix, iy, iz) + VA(ix + 1, iy + 1, iz) - VA(ix, iy + 1, iz) +
                            VA(ix + 1, iy, iz + 1) - VA(ix, iy, iz + 1) + VA(ix + 1, iy + 1, iz + 1) -
                            VA(ix, iy + 1, iz + 1)) /
                           (4.0 * csize(s, ix, 0));
            dA[c + 1][2] = (VA(ix, iy + 1, iz) - VA(ix, iy, iz) + VA(ix + 1, iy + 1, iz) - VA(ix + 1, iy, iz) +
                            VA(ix, iy + 1, iz + 1) - VA(ix, iy, iz + 1) + VA(ix + 1, iy + 1, iz + 1) -
                            VA(ix + 1, iy, iz + 1)) /
                           (4.0 * csize(s, iy, 1));
            dA[c + 1][3] = (VA(ix, iy, iz + 1) - VA(ix, iy, iz) + VA(ix + 1, iy, iz + 1) - VA(ix + 1, iy, iz) +
                            VA(ix, iy + 1, iz + 1) - VA(ix, iy + 1, iz) + VA(ix + 1, iy + 1, iz + 1) -
                            VA(ix + 1, iy + 1, iz)) /
                           (4.0 * csize(s, iz, 2));
          }
#undef VA
        }
        double gdet = s->g[cidx(s, ix, iy, iz) * 20 + 19];
        double *v = s->vecpot + cidx(s, ix, iy, iz) * 6;
        v[3] = (dA[2][3] - dA[3][2]) / gdet;
        v[4] = (dA[3][1] - dA[1][3]) / gdet;
        v[5] = (dA[1][2] - dA[2][1]) / gdet;
      }
  for (long iz = 0; iz < nz; ++iz)
    for (long iy = 0; iy < ny; ++iy)
      for (long ix = 0; ix < nx; ++ix) {
        double *pp = s->p + cidx(s, ix, iy, iz) * NV;
        const double *v = s->vecpot + cidx(s, ix, iy, iz) * 6;
        pp[B1] += v[3];
        pp[B2] += v[4];
        pp[B3] += v[5];
        kzo_geo geom;
        geo_c(s, ix, iy, iz, &geom);
        double uu[NV];
        kzo_p2u(pp, uu, &geom, s->ph.gam);
        double *u = s->u + cidx(s, ix, iy, iz) * NV;
        u[B1] = uu[B1];
        u[B2] = uu[B2];
        u[B3] = uu[B3];
      }
  return KZO_OK;
}

/* ---- scalar diagnostics (io/scalars.zig, io/dump.zig:358-411) ---- */
/* frames.zig:300-310 transPallCoco to BL with the sidecar's point Jacobian:
 * transPmhdCocoJ, then transPradCoco (VELR -> VEL4 recomputing u^t, the
 * Jacobian, VEL4 -> VELR trusting u^t) */
static int trans_pall_bl(double *pp, const kzo_geo *gm, const kzo_geo *gbl, const double *J) {
  int e = trans_pmhd(pp, gm, gbl, J);
  if (e) return e;
  double ucon[4] = {0.0, pp[FX], pp[FY], pp[FZ]};
  e = kzo_convert(ucon, VELR, VEL4, gm, 0, ucon);
  if (e) return e;
  tvec(J, ucon, ucon);
  e = kzo_convert(ucon, VEL4, VELR, gbl, 1, ucon);
  if (e) return e;
  pp[FX] = ucon[1]; pp[FY] = ucon[2]; pp[FZ] = ucon[3];
  return KZO_OK;
}

/* The reductions of one scalars.dat row on one rank (scalars.zig:79-313
 * totalMass, mdot, lum, scaleHeightPartsAtIx + scaleHeightFinalize,
 * maxPmagPtot; dump.zig:358-380 collectDiag), in the reference's loop order.
 * ix_h, ix_l, ix_s: radial indices of the mass-flux, luminosity and
 * scale-height shells; dth, dph: BL theta / phi widths of shell ix_l's cells,
 * [nz][ny].  out: mass, mdot (net, not sign-flipped), radlum, totallum,
 * scaleheight, max pmag/ptot; cnt: cells with an HD fixup flag, cells with a
 * non-finite primitive.  Needs the BL sidecar. */
int kzo_sim_scalars(kzo_sim *s, long ix_h, long ix_l, long ix_s, const double *dth, const double *dph, double out[6],
                    uint64_t cnt[2]) {
  const long nx = s->grid.nx, ny = s->grid.ny, nz = s->grid.nz;
  const int twod = nz == 1;
  const double wedge = s->grid.maxz - s->grid.minz;
  if (!s->bl) return KZ_E_INVALID_CONFIG;
  /* totalMass (scalars.zig:79-101) */
  const double phifac = nz > 1 ? 2.0 * KZO_PI / wedge : 1.0;
  double mass = 0;
  for (long iz = 0; iz < nz; ++iz)
    for (long iy = 0; iy < ny; ++iy)
      for (long ix = 0; ix < nx; ++ix) {
        kzo_geo g;
        geo_c(s, ix, iy, iz, &g);
        double dx = csize(s, ix, 0), dy = csize(s, iy, 1), dz = csize(s, iz, 2) * phifac;
        mass += cell(s->p, s, ix, iy, iz)[RHO] * dx * dy * dz * g.gdet;
      }
  /* mdot (scalars.zig:108-133) */
  double md = 0;
  for (long iz = 0; iz < nz; ++iz)
    for (long iy = 0; iy < ny; ++iy) {
      const double *pp = cell(s->p, s, ix_h, iy, iz);
      kzo_geo g;
      geo_c(s, ix_h, iy, iz, &g);
      double v[4] = {0.0, pp[VX], pp[VY], pp[VZ]}, ucon[4];
      int e = kzo_convert(v, VELR, VEL4, &g, 0, ucon);
      if (e) return e;
      double rhouconr = pp[RHO] * ucon[1];
      double dy = csize(s, iy, 1);
      double dz = twod ? 2.0 * KZO_PI : csize(s, iz, 2) * (2.0 * KZO_PI / wedge);
      md += g.gdet * rhouconr * dy * dz;
    }
  /* lum, type 1 (scalars.zig:141-185) */
  double radlum = 0, totallum = 0;
  for (long iz = 0; iz < nz; ++iz)
    for (long iy = 0; iy < ny; ++iy) {
      double pp[NV];
      memcpy(pp, cell(s->p, s, ix_l, iy, iz), sizeof pp);
      kzo_geo g, gbl;
      geo_c(s, ix_l, iy, iz, &g);
      bl_geo(s, ix_l, iy, iz, &gbl);
      int e = trans_pall_bl(pp, &g, &gbl, bl_rec(s, ix_l, iy, iz) + 36);
      if (e) return e;
      double dthbl = dth[iz * ny + iy];
      double dphbl = twod ? 2.0 * KZO_PI : dph[iz * ny + iy] * (2.0 * KZO_PI / wedge);
      double v[4] = {0.0, pp[VX], pp[VY], pp[VZ]}, ucon[4];
      e = kzo_convert(v, VELR, VEL4, &gbl, 0, ucon);
      if (e) return e;
      double rhour = pp[RHO] * ucon[1];
      double tu[4][4], tl[4][4];
      e = kzo_calc_tij(pp, &gbl, s->ph.gam, tu);
      if (e) return e;
      kzo_lower_second(tu, &gbl, tl);
      double trt = tl[1][0];
      double ru[4][4], rl[4][4];
      kzo_calc_rij(pp, &gbl, ru);
      kzo_lower_second(ru, &gbl, rl);
      double rrt = -rl[1][0];
      if (rrt < 0) rrt = 0;
      radlum += gbl.gdet * rrt * dthbl * dphbl;
      totallum += gbl.gdet * (rhour + trt + rrt) * dthbl * dphbl;
    }
  /* scaleHeightPartsAtIx + scaleHeightFinalize (dynamo.zig:239-269) */
  double sig = 0, sth = 0;
  for (long iy = 0; iy < ny; ++iy)
    for (long iz = 0; iz < nz; ++iz) {
      kzo_geo g;
      geo_c(s, ix_s, iy, iz, &g);
      double gd = g.gdet, th = bl_rec(s, ix_s, iy, iz)[35];
      double rho = cell(s->p, s, ix_s, iy, iz)[RHO];
      double dthe = KZO_PI / 2.0 - th;
      sig += rho * gd;
      sth += rho * gd * dthe * dthe;
    }
  double hr = ix_s > 0 ? kzm_sqrt(sth / sig) : sth;
  /* maxPmagPtot (scalars.zig:275-313) and collectDiag */
  double maxb = 0;
  uint64_t nhd = 0, nnan = 0;
  for (long iz = 0; iz < nz; ++iz)
    for (long iy = 0; iy < ny; ++iy)
      for (long ix = 0; ix < nx; ++ix) {
        const double *pp = cell(s->p, s, ix, iy, iz);
        kzo_geo g;
        geo_c(s, ix, iy, iz, &g);
        double ucon[4], ucov[4];
        kzo_ucon_ucov_from_prims(pp, &g, ucon, ucov);
        double B[3] = {pp[B1], pp[B2], pp[B3]}, bcon[4], bcov[4], bsq;
        kzo_bcon_bcov_bsq(B, ucon, ucov, &g, bcon, bcov, &bsq);
        double pmag = bsq / 2.0;
        double ptot = (s->ph.gam - 1.0) * pp[UU];
        double rtt = kzo_calc_ff_rtt(pp, &g, 0);
        ptot += (-rtt) / 3.0;
        if (pmag / ptot > maxb) maxb = pmag / ptot;
        for (int iv = 0; iv < NV; ++iv)
          if (!kzm_isfinite(pp[iv])) {
            nnan += 1;
            break;
          }
        if (*flagp(s, KZ_FLAG_HD_FIXUP, ix, iy, iz) != 0) nhd += 1;
      }
  out[0] = mass; out[1] = md; out[2] = radlum; out[3] = totallum; out[4] = hr; out[5] = maxb;
  cnt[0] = nhd; cnt[1] = nnan;
  return KZO_OK;
}

/* dynamo.zig:462-469 applyDynamo */
static int apply_dynamo(kzo_sim *s, double dt) {
  calc_scale_height(s);
  int e = kzo_sim_set_bc(s, 0);
  if (e) return e;
  e = mimic_dynamo(s, dt);
  if (e) return e;
  return kzo_sim_calc_u2p(s);
}

/* ---- explicit operator (sim.zig:1209-1582) ---- */
static long cross_lo(const kzo_sim *s, int dim) { return ndim(s, dim) > 1 ? -1 : 0; }
static long cross_hi(const kzo_sim *s, int dim) { return ndim(s, dim) > 1 ? ndim(s, dim) + 1 : ndim(s, dim); }

static int sweep_face(kzo_sim *s, int dim, long i, const long cellc[3], uint64_t *nguard) {
  long n = ndim(s, dim);
  int dol = i >= 0, dor = i < n;
  double st[5][NV];
  long c[3] = {cellc[0], cellc[1], cellc[2]};
  for (int k = 0; k < 5; ++k) {
    c[dim] = i - 2 + k;
    memcpy(st[k], cell(s->p, s, c[0], c[1], c[2]), NV * sizeof(double));
  }
  double dx5[5];
  for (int k = 0; k < 5; ++k) dx5[k] = csize(s, i - 2 + k, dim);
  /* sim.zig:1316-1337: REDUCEORDERATBH (the cell's radial index inside the
     horizon) and REDUCEORDERAFTERFIXUP (a fixup flag set on the cell by the
     last u2p or implicit pass; ghost cells carry none) drop the order by one */
  int reduce = s->ph.reduceorderatbh && cellc[0] <= s->ph.reduce_order_ix_max;
  if (!reduce && s->ph.reduceorderafterfixup &&
      (*flagp(s, KZ_FLAG_HD_FIXUP, cellc[0], cellc[1], cellc[2]) != 0 ||
       *flagp(s, KZ_FLAG_RAD_FIXUP, cellc[0], cellc[1], cellc[2]) != 0 ||
       *flagp(s, KZ_FLAG_RADIMP_FIXUP, cellc[0], cellc[1], cellc[2]) != 0))
    reduce = 1;
  double pl[NV], pr[NV];
  for (int iv = 0; iv < NV; ++iv) {
    double u5[5] = {st[0][iv], st[1][iv], st[2][iv], st[3][iv], st[4][iv]};
    kzo_reconstruct(2 - reduce, s->ph.minmod_theta, u5, dx5, &pl[iv], &pr[iv]);
  }
  double ffl[NV], ffr[NV];
  if (dol) {
    kzo_geo g;
    geo_f(s, cellc[0], cellc[1], cellc[2], dim, &g);
    kzo_check_floors_mhd(pl, &g, s->ph.gam, &s->ph.floors, nguard);
    int e = kzo_f_flux_prime(pl, dim, &g, s->ph.gam, ffl);
    if (e) return e;
    if (s->ph.radviscosity) add_rad_visc_flux(s, dim, cellc, pl, &g, ffl); /* sim.zig:1349-1352 */
  }
  long cp[3] = {cellc[0], cellc[1], cellc[2]};
  cp[dim] = i + 1;
  if (dor) {
    kzo_geo g;
    geo_f(s, cp[0], cp[1], cp[2], dim, &g);
    kzo_check_floors_mhd(pr, &g, s->ph.gam, &s->ph.floors, nguard);
    int e = kzo_f_flux_prime(pr, dim, &g, s->ph.gam, ffr);
    if (e) return e;
    if (s->ph.radviscosity) add_rad_visc_flux(s, dim, cp, pr, &g, ffr); /* sim.zig:1361-1364 */
  }
  memcpy(face(s->pb_r[dim], s, dim, cellc[0], cellc[1], cellc[2]), pl, sizeof pl);
  if (dol) memcpy(face(s->fl_r[dim], s, dim, cellc[0], cellc[1], cellc[2]), ffl, sizeof ffl);
  memcpy(face(s->pb_l[dim], s, dim, cp[0], cp[1], cp[2]), pr, sizeof pr);
  if (dor) memcpy(face(s->fl_l[dim], s, dim, cp[0], cp[1], cp[2]), ffr, sizeof ffr);
  return KZO_OK;
}

static int sweep(kzo_sim *s, int dim) {
  long n = ndim(s, dim);
  if (n <= 1) return KZO_OK;
  int c0d = dim == 0 ? 1 : 0, c1d = dim == 2 ? 1 : 2;
  const long lo0 = cross_lo(s, c0d), lo1 = cross_lo(s, c1d);
  const long w1 = cross_hi(s, c1d) - lo1, nline = (cross_hi(s, c0d) - lo0) * w1;
  uint64_t nguard = 0;
  int err = KZO_OK;
  /* every line runs to its end; the first error in line order is reported */
  long err_line = nline;
#pragma omp parallel for num_threads(kzo_nthreads) schedule(dynamic, 1) reduction(+ : nguard)
  for (long line = 0; line < nline; ++line) {
    long c0 = lo0 + line / w1, c1 = lo1 + line % w1;
    for (long i = -1; i < n + 1; ++i) {
      long cl[3];
      cl[dim] = i; cl[c0d] = c0; cl[c1d] = c1;
      int e = sweep_face(s, dim, i, cl, &nguard);
      if (e) {
#pragma omp critical(kzo_sweep_err)
        if (line < err_line) {
          err_line = line;
          err = e;
        }
        break;
      }
    }
  }
  s->n_guard += nguard;
  return err;
}

static int flux_face(kzo_sim *s, int dim, const long cf[3]) {
  long cm[3] = {cf[0], cf[1], cf[2]};
  cm[dim] -= 1;
  double ap1 = *scp(s, ahd_m[dim], cf[0], cf[1], cf[2]);
  double am1 = *scp(s, ahd_m[dim], cm[0], cm[1], cm[2]);
  double pLl[NV], pRl[NV], uLl[NV], uRl[NV];
  memcpy(pLl, face(s->pb_l[dim], s, dim, cf[0], cf[1], cf[2]), sizeof pLl);
  memcpy(pRl, face(s->pb_r[dim], s, dim, cf[0], cf[1], cf[2]), sizeof pRl);
  kzo_geo g;
  geo_f(s, cf[0], cf[1], cf[2], dim, &g);
  kzo_p2u(pLl, uLl, &g, s->ph.gam);
  kzo_p2u(pRl, uRl, &g, s->ph.gam);
  double ag = kzm_fmax(ap1, am1);
  double rp1 = *scp(s, arad_m[dim], cf[0], cf[1], cf[2]);
  double rm1 = *scp(s, arad_m[dim], cm[0], cm[1], cm[2]);
  double agr = kzm_fmax(rp1, rm1);
  double *fll = face(s->fl_l[dim], s, dim, cf[0], cf[1], cf[2]);
  double *flr = face(s->fl_r[dim], s, dim, cf[0], cf[1], cf[2]);
  double fstar[NV];
  for (int iv = 0; iv < NV; ++iv) fstar[iv] = kzo_laxf(fll[iv], flr[iv], uLl[iv], uRl[iv], iv >= EE ? agr : ag);
  memcpy(face(s->flb[dim], s, dim, cf[0], cf[1], cf[2]), fstar, sizeof fstar);
  return KZO_OK;
}

static int fluxes_at_faces(kzo_sim *s) {
  for (int d = 0; d < 3; ++d) memset(s->flb[d], 0, (size_t)s->nface[d] * NV * sizeof(double));
  for (int dim = 0; dim < 3; ++dim) {
    long n = ndim(s, dim);
    if (n <= 1) continue;
    int c0d = dim == 0 ? 1 : 0, c1d = dim == 2 ? 1 : 2;
    const long lo0 = cross_lo(s, c0d), lo1 = cross_lo(s, c1d);
    const long w1 = cross_hi(s, c1d) - lo1, nline = (cross_hi(s, c0d) - lo0) * w1;
#pragma omp parallel for num_threads(kzo_nthreads) schedule(dynamic, 1)
    for (long line = 0; line < nline; ++line) {
      long c0 = lo0 + line / w1, c1 = lo1 + line % w1;
      for (long i = 0; i <= n; ++i) {
        long cf[3];
        cf[dim] = i; cf[c0d] = c0; cf[c1d] = c1;
        flux_face(s, dim, cf);
      }
    }
  }
  return KZO_OK;
}

/* ct.zig:336-381 calcDivB: corner divergence of gdet*B from cell-centred
 * values, over every domain cell into out[(iz*ny + iy)*nx + ix]; reads one
 * ghost layer below each active dimension. */
static double gdet_pb(const kzo_sim *s, int iv, long ix, long iy, long iz) {
  long ci = cidx(s, ix, iy, iz);
  return s->g[ci * 20 + 19] * s->p[ci * NV + iv];
}
void kzo_sim_calc_divb(kzo_sim *s, double *out) {
  const long nx = s->grid.nx, ny = s->grid.ny, nz = s->grid.nz;
  for (long iz = 0; iz < nz; ++iz)
    for (long iy = 0; iy < ny; ++iy)
      for (long ix = 0; ix < nx; ++ix) {
        double divb;
        if (nz == 1) {
          divb = (gdet_pb(s, B1, ix, iy, iz) + gdet_pb(s, B1, ix, iy - 1, iz) - gdet_pb(s, B1, ix - 1, iy, iz) -
                  gdet_pb(s, B1, ix - 1, iy - 1, iz)) /
                     (2.0 * (xc_(s, ix + 1) - xc_(s, ix))) +
                 (gdet_pb(s, B2, ix, iy, iz) + gdet_pb(s, B2, ix - 1, iy, iz) - gdet_pb(s, B2, ix, iy - 1, iz) -
                  gdet_pb(s, B2, ix - 1, iy - 1, iz)) /
                     (2.0 * (yc_(s, iy + 1) - yc_(s, iy)));
        } else {
          divb = (gdet_pb(s, B1, ix, iy, iz) + gdet_pb(s, B1, ix, iy - 1, iz) - gdet_pb(s, B1, ix - 1, iy, iz) -
                  gdet_pb(s, B1, ix - 1, iy - 1, iz) + gdet_pb(s, B1, ix, iy, iz - 1) +
                  gdet_pb(s, B1, ix, iy - 1, iz - 1) - gdet_pb(s, B1, ix - 1, iy, iz - 1) -
                  gdet_pb(s, B1, ix - 1, iy - 1, iz - 1)) /
                     (4.0 * (xc_(s, ix) - xc_(s, ix - 1))) +
                 (gdet_pb(s, B2, ix, iy, iz) + gdet_pb(s, B2, ix - 1, iy, iz) - gdet_pb(s, B2, ix, iy - 1, iz) -
                  gdet_pb(s, B2, ix - 1, iy - 1, iz) + gdet_pb(s, B2, ix, iy, iz - 1) +
                  gdet_pb(s, B2, ix - 1, iy, iz - 1) - gdet_pb(s, B2, ix, iy - 1, iz - 1) -
                  gdet_pb(s, B2, ix - 1, iy - 1, iz - 1)) /
                     (4.0 * (yc_(s, iy) - yc_(s, iy - 1))) +
                 (gdet_pb(s, B3, ix, iy, iz) + gdet_pb(s, B3, ix - 1, iy, iz) - gdet_pb(s, B3, ix, iy, iz - 1) -
                  gdet_pb(s, B3, ix - 1, iy, iz - 1) + gdet_pb(s, B3, ix, iy - 1, iz) +
                  gdet_pb(s, B3, ix - 1, iy - 1, iz) - gdet_pb(s, B3, ix, iy - 1, iz - 1) -
                  gdet_pb(s, B3, ix - 1, iy - 1, iz - 1)) /
                     (4.0 * (zc_(s, iz) - zc_(s, iz - 1)));
        }
        out[(iz * ny + iy) * nx + ix] = divb / s->g[cidx(s, ix, iy, iz) * 20 + 19];
      }
}

/* ct.zig:38-177 fluxCt */
static long emf_idx(const kzo_sim *s, long ix, long iy, long iz) {
  return (iz * (s->grid.ny + 1) + iy) * (long)(s->grid.nx + 1) + ix;
}
static long pin(long n, long i) { return n == 1 ? 0 : i; }
static double flbg(kzo_sim *s, int d, int iv, long ix, long iy, long iz) { return face(s->flb[d], s, d, ix, iy, iz)[iv]; }

void kzo_flux_ct(kzo_sim *s) {
  long nx = s->grid.nx, ny = s->grid.ny, nz = s->grid.nz;
  double coef1 = (ny > 1 && nz > 1) ? 0.25 : 0.5;
  double coef2 = (nx > 1 && nz > 1) ? 0.25 : 0.5;
  double coef3 = (nx > 1 && ny > 1) ? 0.25 : 0.5;
  for (long iz = 0; iz <= nz; ++iz)
    for (long iy = 0; iy <= ny; ++iy)
      for (long ix = 0; ix <= nx; ++ix) {
        long e = emf_idx(s, ix, iy, iz);
        if (ny > 1 || nz > 1) {
          double sum = 0;
          if (ny > 1) sum += flbg(s, 1, B3, pin(nx, ix), pin(ny, iy), pin(nz, iz)) + flbg(s, 1, B3, pin(nx, ix), pin(ny, iy), pin(nz, iz - 1));
          if (nz > 1) sum -= flbg(s, 2, B2, pin(nx, ix), pin(ny, iy), pin(nz, iz)) + flbg(s, 2, B2, pin(nx, ix), pin(ny, iy - 1), pin(nz, iz));
          s->emf[0][e] = coef1 * sum;
        } else s->emf[0][e] = 0;
        if (nx > 1 || nz > 1) {
          double sum = 0;
          if (nz > 1) sum += flbg(s, 2, B1, pin(nx, ix), pin(ny, iy), pin(nz, iz)) + flbg(s, 2, B1, pin(nx, ix - 1), pin(ny, iy), pin(nz, iz));
          if (nx > 1) sum -= flbg(s, 0, B3, ix, pin(ny, iy), pin(nz, iz)) + flbg(s, 0, B3, ix, pin(ny, iy), pin(nz, iz - 1));
          s->emf[1][e] = coef2 * sum;
        } else s->emf[1][e] = 0;
        if (nx > 1 || ny > 1) {
          double sum = 0;
          if (nx > 1) sum += flbg(s, 0, B2, ix, pin(ny, iy), pin(nz, iz)) + flbg(s, 0, B2, ix, pin(ny, iy - 1), pin(nz, iz));
          if (ny > 1) sum -= flbg(s, 1, B1, pin(nx, ix), pin(ny, iy), pin(nz, iz)) + flbg(s, 1, B1, pin(nx, ix - 1), pin(ny, iy), pin(nz, iz));
          s->emf[2][e] = coef3 * sum;
        } else s->emf[2][e] = 0;
      }
  for (long iz = 0; iz <= nz; ++iz)
    for (long iy = 0; iy <= ny; ++iy)
      for (long ix = 0; ix <= nx; ++ix) {
#define EMF(c, a, b, d) s->emf[(c)-1][emf_idx(s, a, b, d)]
        if (nx > 1 && iy < ny && iz < nz) {
          double *f = face(s->flb[0], s, 0, ix, iy, iz);
          f[B1] = 0;
          f[B2] = 0.5 * (EMF(3, ix, iy, iz) + EMF(3, ix, iy + 1, iz));
          f[B3] = -0.5 * (EMF(2, ix, iy, iz) + EMF(2, ix, iy, iz + 1));
        }
        if (ny > 1 && ix < nx && iz < nz) {
          double *f = face(s->flb[1], s, 1, ix, iy, iz);
          f[B1] = -0.5 * (EMF(3, ix, iy, iz) + EMF(3, ix + 1, iy, iz));
          f[B2] = 0;
          f[B3] = 0.5 * (EMF(1, ix, iy, iz) + EMF(1, ix, iy, iz + 1));
        }
        if (nz > 1 && ix < nx && iy < ny) {
          double *f = face(s->flb[2], s, 2, ix, iy, iz);
          f[B1] = 0.5 * (EMF(2, ix, iy, iz) + EMF(2, ix + 1, iy, iz));
          f[B2] = -0.5 * (EMF(1, ix, iy, iz) + EMF(1, ix, iy + 1, iz));
          f[B3] = 0;
        }
#undef EMF
      }
}

static void update_rows(kzo_sim *s, double dtin) {
  const long ny = s->grid.ny, nrow = s->grid.nz * ny;
#pragma omp parallel for num_threads(kzo_nthreads) schedule(dynamic, 1)
  for (long row = 0; row < nrow; ++row)
    for (long ix = 0; ix < s->grid.nx; ++ix) {
      long iz = row / ny, iy = row % ny;
      double ms[NV];
      metric_source(s, ix, iy, iz, ms);
      double dx = csize(s, ix, 0), dy = csize(s, iy, 1), dz = csize(s, iz, 2);
      double dt = dtin;
      double *flxl = face(s->flb[0], s, 0, ix, iy, iz), *flxr = face(s->flb[0], s, 0, ix + 1, iy, iz);
      double *flyl = face(s->flb[1], s, 1, ix, iy, iz), *flyr = face(s->flb[1], s, 1, ix, iy + 1, iz);
      /* z faces of a collapsed dim: the store is zeroed by fluxesAtFaces */
      double *flzl = face(s->flb[2], s, 2, ix, iy, iz), *flzr = face(s->flb[2], s, 2, ix, iy, iz + 1);
      double *uu = cell(s->u, s, ix, iy, iz);
      for (int iv = 0; iv < NV; ++iv) {
        double du = -(flxr[iv] - flxl[iv]) * dt / dx - (flyr[iv] - flyl[iv]) * dt / dy - (flzr[iv] - flzl[iv]) * dt / dz;
        uu[iv] = uu[iv] + du + ms[iv] * dt;
      }
    }
}

int kzo_sim_op_explicit(kzo_sim *s, double dt) {
  int e = calc_wavespeeds(s);
  if (e) return e;
  for (int dim = 0; dim < 3; ++dim) {
    e = sweep(s, dim);
    if (e) return e;
  }
  fluxes_at_faces(s);
  kzo_flux_ct(s);
  update_rows(s, dt);
  return kzo_sim_calc_u2p(s);
}

/* ---- implicit operator (sim.zig:1715-1776) ---- */
int kzo_sim_op_implicit(kzo_sim *s, double dtin) {
  if (!s->ph.has_opac) return KZO_OK;
  const long ny = s->grid.ny, nrow = s->grid.nz * ny;
  uint64_t n_iters = 0, n_solves = 0, n_fail = 0;
#pragma omp parallel for num_threads(kzo_nthreads) schedule(dynamic, 1) reduction(+ : n_iters, n_solves, n_fail)
  for (long row = 0; row < nrow; ++row) {
    long iz = row / ny, iy = row % ny;
    if (polar_row(s, iy)) continue;
    for (long ix = 0; ix < s->grid.nx; ++ix) {
      double uu[NV], pp[NV];
      memcpy(uu, cell(s->u, s, ix, iy, iz), sizeof uu);
      memcpy(pp, cell(s->p, s, ix, iy, iz), sizeof pp);
      kzo_geo g;
      geo_c(s, ix, iy, iz, &g);
      kz_implicit_result r = kzo_solve_implicit_lab(uu, pp, &g, dtin, s->ph.gam, &s->ph.rad, &s->ph.opac, &s->ph.implicit);
      if (r.ok) {
        memcpy(cell(s->u, s, ix, iy, iz), uu, sizeof uu);
        memcpy(cell(s->p, s, ix, iy, iz), pp, sizeof pp);
        *flagp(s, KZ_FLAG_RADIMP_FIXUP, ix, iy, iz) = 0;
        n_iters += (uint64_t)r.iters;
        n_solves += 1;
      } else {
        *flagp(s, KZ_FLAG_RADIMP_FIXUP, ix, iy, iz) = -1;
        n_fail += 1;
      }
    }
  }
  s->n_iters += n_iters;
  s->n_solves += n_solves;
  s->n_fail += n_fail;
  cell_fixup(s, KZ_FLAG_RADIMP_FIXUP);
  return KZO_OK;
}

/* ---- polar axis overwrite (polaraxis.zig:84-147) ---- */
int kzo_sim_do_correct(kzo_sim *s) {
  if (!s->ph.correct_polaraxis) return KZO_OK;
  long nc = s->ph.nccorrectpolar, ny = s->grid.ny;
  for (long iz = 0; iz < s->grid.nz; ++iz)
  for (long ix = 0; ix < s->grid.nx; ++ix)
    for (long ic = 0; ic < nc; ++ic) {
      long rows2[2] = {ic, ny - 1 - ic};
      for (int w = 0; w < 2; ++w) {
        long iy = rows2[w];
        long iysrc = iy < nc ? nc : ny - 1 - nc;
        double thaxis = yl_(s, iy < nc ? 0 : ny);
        double fac = kzm_fabs((yc_(s, iy) - thaxis) / (yc_(s, iysrc) - thaxis));
        double pp[NV], uu[NV];
        double *ps = cell(s->p, s, ix, iysrc, iz);
        memcpy(pp, cell(s->p, s, ix, iy, iz), sizeof pp);
        pp[RHO] = ps[RHO]; pp[UU] = ps[UU]; pp[ENTR] = ps[ENTR];
        pp[VX] = ps[VX]; pp[VZ] = ps[VZ]; pp[VY] = fac * ps[VY];
        pp[EE] = ps[EE]; pp[FX] = ps[FX]; pp[FZ] = ps[FZ]; pp[FY] = fac * ps[FY];
        kzo_geo g;
        geo_c(s, ix, iy, iz, &g);
        kzo_p2u(pp, uu, &g, s->ph.gam);
        memcpy(cell(s->p, s, ix, iy, iz), pp, sizeof pp);
        memcpy(cell(s->u, s, ix, iy, iz), uu, sizeof uu);
      }
    }
  return KZO_OK;
}

/* ---- stage arithmetic (sim.zig:1586-1709) ---- */
static void stage_deriv(kzo_sim *s, double *dst, const double *a, const double *b, double delta) {
  double f1 = 1.0 / delta, f2 = -1.0 / delta;
  for (long iz = 0; iz < s->grid.nz; ++iz)
  for (long iy = 0; iy < s->grid.ny; ++iy)
    for (long ix = 0; ix < s->grid.nx; ++ix) {
      long o = cidx(s, ix, iy, iz) * NV;
      for (int iv = 0; iv < NV; ++iv) dst[o + iv] = f1 * a[o + iv] + f2 * b[o + iv];
    }
}
static void stage_combine(kzo_sim *s, double *dst, const double *a, double f1, const double *b, double f2, const double *c) {
  for (long iz = 0; iz < s->grid.nz; ++iz)
  for (long iy = 0; iy < s->grid.ny; ++iy)
    for (long ix = 0; ix < s->grid.nx; ++ix) {
      long o = cidx(s, ix, iy, iz) * NV;
      for (int iv = 0; iv < NV; ++iv) dst[o + iv] = a[o + iv] + f1 * b[o + iv] + f2 * c[o + iv];
    }
}
static void copy_full(kzo_sim *s, double *dst, const double *src) { memcpy(dst, src, (size_t)s->ncell * NV * sizeof(double)); }

static void copy_entropy_count(kzo_sim *s) {
  for (long iz = 0; iz < s->grid.nz; ++iz)
  for (long iy = 0; iy < s->grid.ny; ++iy)
    for (long ix = 0; ix < s->grid.nx; ++ix)
      *flagp(s, KZ_FLAG_ENTROPY3, ix, iy, iz) = *flagp(s, KZ_FLAG_ENTROPY, ix, iy, iz);
}

static void update_entropy(kzo_sim *s) {
  for (long iz = 0; iz < s->grid.nz; ++iz)
  for (long iy = 0; iy < s->grid.ny; ++iy)
    for (long ix = 0; ix < s->grid.nx; ++ix) {
      double *pp = cell(s->p, s, ix, iy, iz);
      pp[ENTR] = kzo_s_from_u(pp[RHO], pp[UU], s->ph.gam);
      kzo_geo g;
      geo_c(s, ix, iy, iz, &g);
      kzo_p2u_mhd(pp, cell(s->u, s, ix, iy, iz), &g, s->ph.gam);
    }
}

/* The RK2IMEX gamma 1 - 1/sqrt(2), correctly rounded (sim.zig:1821 is a
 * comptime expression evaluated above f64 precision). */
static const double GAM_IMEX = 0x1.2bec333018867p-2;

/* sim.zig:1792-1899 step */
int kzo_sim_step(kzo_sim *s, const double *forced_dt) {
  double t = s->t;
  s->own_dt = 1.0 / s->tstepdenmax;
  double dt = forced_dt ? *forced_dt : s->own_dt;
  s->dt = dt;
  s->tstepdenmax = -1;
  s->tstepdenmin = 1.0e50;
  const double g = GAM_IMEX;
  int e;
  if ((s->ph.radviscosity || s->ph.dynamo) && !s->bl) return KZ_E_INVALID_CONFIG;
  if (s->ph.radviscosity) calc_rij_visc_total(s, dt); /* sim.zig:1819-1823 */
  save_timesteps(s);
  copy_full(s, s->ut0, s->u);
  if ((e = kzo_sim_op_implicit(s, dt * g))) return e;
  stage_deriv(s, s->drt1, s->u, s->ut0, dt * g);

  copy_full(s, s->ut1, s->u);
  if ((e = kzo_sim_calc_u2p(s))) return e;
  kzo_sim_do_correct(s);
  if ((e = kzo_sim_set_bc(s, 0))) return e;
  if ((e = kzo_sim_op_explicit(s, dt))) return e;
  if (s->ph.dynamo && (e = apply_dynamo(s, dt))) return e; /* sim.zig:1845-1849 */
  copy_entropy_count(s);
  stage_deriv(s, s->dut1, s->u, s->ut1, dt);

  stage_combine(s, s->u, s->ut0, dt, s->dut1, dt * (1.0 - 2.0 * g), s->drt1);

  copy_full(s, s->uforget, s->u);
  if ((e = kzo_sim_calc_u2p(s))) return e;
  kzo_sim_do_correct(s);
  if ((e = kzo_sim_op_implicit(s, g * dt))) return e;
  stage_deriv(s, s->drt2, s->u, s->uforget, dt * g);

  copy_full(s, s->ut2, s->u);
  if ((e = kzo_sim_calc_u2p(s))) return e;
  kzo_sim_do_correct(s);
  if ((e = kzo_sim_set_bc(s, 0))) return e;
  if ((e = kzo_sim_op_explicit(s, dt))) return e;
  if (s->ph.dynamo && (e = apply_dynamo(s, dt))) return e; /* sim.zig:1872-1876 */
  stage_deriv(s, s->dut2, s->u, s->ut2, dt);

  stage_combine(s, s->u, s->ut0, dt / 2.0, s->dut1, dt / 2.0, s->dut2);
  stage_combine(s, s->u, s->u, dt / 2.0, s->drt1, dt / 2.0, s->drt2);

  if ((e = kzo_sim_calc_u2p(s))) return e;
  kzo_sim_do_correct(s);
  if ((e = kzo_sim_set_bc(s, 0))) return e;
  s->t = t + dt;
  update_entropy(s);
  s->nstep += 1;
  return KZO_OK;
}

/* puffy.zig:892-1024 postinit (BETANORMFULL): scale B by sqrt(maxbeta/max
 * pmag/ptot) over the domain and refresh u. Ghosts keep the unscaled B. */
double kzo_sim_postinit(kzo_sim *s, double maxbeta) {
  double maxb = 0.0;
  for (long iz = 0; iz < s->grid.nz; ++iz)
  for (long iy = 0; iy < s->grid.ny; ++iy)
    for (long ix = 0; ix < s->grid.nx; ++ix) {
      const double *pp = cell(s->p, s, ix, iy, iz);
      kzo_geo g;
      geo_c(s, ix, iy, iz, &g);
      double ucon[4], ucov[4];
      kzo_ucon_ucov_from_prims(pp, &g, ucon, ucov);
      double B[3] = {pp[B1], pp[B2], pp[B3]}, bcon[4], bcov[4], bsq;
      kzo_bcon_bcov_bsq(B, ucon, ucov, &g, bcon, bcov, &bsq);
      double pmag = bsq / 2.0;
      double pgas = (s->ph.gam - 1.0) * pp[UU];
      double ptot = pgas;
      double rtt = kzo_calc_ff_rtt(pp, &g, 0);
      double ehat = -rtt;
      ptot += ehat / 3.0;
      if (pmag / ptot > maxb) maxb = pmag / ptot;
    }
  double fac = kzm_sqrt(maxbeta / maxb);
  for (long iz = 0; iz < s->grid.nz; ++iz)
  for (long iy = 0; iy < s->grid.ny; ++iy)
    for (long ix = 0; ix < s->grid.nx; ++ix) {
      double *pp = cell(s->p, s, ix, iy, iz);
      pp[B1] *= fac; pp[B2] *= fac; pp[B3] *= fac;
      p2u_cell(s, ix, iy, iz);
    }
  return fac;
}

/* test hooks: the per-step radiative-viscosity pass and one dynamo application */
int kzo_sim_calc_rij_visc_total(kzo_sim *s, double global_dt) {
  if (!s->bl) return KZ_E_INVALID_CONFIG;
  calc_rij_visc_total(s, global_dt);
  return KZO_OK;
}
void kzo_sim_get_rijvisc(kzo_sim *s, double *out) { memcpy(out, s->rijvisc, (size_t)s->ncell * 16 * sizeof(double)); }
int kzo_sim_apply_dynamo(kzo_sim *s, double dt) {
  if (!s->bl) return KZ_E_INVALID_CONFIG;
  return apply_dynamo(s, dt);
}
void kzo_sim_get_scaleth(kzo_sim *s, double *out) { memcpy(out, s->scaleth, (size_t)s->grid.nx * sizeof(double)); }
double kzo_step_function(double x, double x9) { return step_function(x, x9); }
double kzo_visc_coeff(double chi, double mindx, double r_bl, double rhor, double alpha, double global_dt) {
  return visc_coeff(chi, mindx, r_bl, rhor, alpha, global_dt);
}
void kzo_shear_from_gradients(const double *du16, const double *du2_16, const double *ucon, const double *ucov,
                              const kz_geom *g, const double *kr64, double *sig16) {
  kzo_geo gm;
  kzo_geo_from_kz(g, &gm);
  shear_from_gradients((const double(*)[4])du16, (const double(*)[4])du2_16, ucon, ucov, &gm, kr64,
                       (double(*)[4])sig16);
}
double kzo_dynamo_delta_a(const kz_dynamo_params *dp, double a, double dt, double rhor, double risco, double r,
                          double th, double bphi, double bsq, double angle, double prermhd, double gg33,
                          double scaleth, double *bphi_out) {
  return dynamo_delta_a(dp, a, dt, rhor, risco, r, th, bphi, bsq, angle, prermhd, gg33, scaleth, bphi_out);
}
double kzo_damp_bphi(double alphabeta, double facradius, double faczh, double dt_over_pk, double beta,
                     double betasat, double bphi) {
  return damp_bphi(alphabeta, facradius, faczh, dt_over_pk, beta, betasat, bphi);
}
int kzo_sim_shear_lab(kzo_sim *s, long ix, long iy, long iz, double *sig16) {
  calc_shear_lab(s, ix, iy, iz, (double(*)[4])sig16);
  return KZO_OK;
}
void kzo_sim_get_dyn_a(kzo_sim *s, double *out) { memcpy(out, s->dyn_a, (size_t)s->ncell * sizeof(double)); }
/* the dynamo core without the final calcU2p: scale height, setBc, mimicDynamo */
int kzo_sim_mimic_dynamo(kzo_sim *s, double dt) {
  if (!s->bl) return KZ_E_INVALID_CONFIG;
  calc_scale_height(s);
  int e = kzo_sim_set_bc(s, 0);
  if (e) return e;
  return mimic_dynamo(s, dt);
}
