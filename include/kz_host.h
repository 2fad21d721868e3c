/* kz_host.h — caller-side C-ABI (libkz_host.so): the tables and initial
 * state koral-zig builds on the host before the per-step hot path runs.
 * These replace the Zig calls the run driver makes once per run; none of
 * them is on the per-step path.
 */
#ifndef KZ_HOST_H
#define KZ_HOST_H

#include <stdint.h>
#include "kz_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* padded cell / face counts (koral/grid.zig:96-116, metric/precompute.zig:216-223) */
long kz_host_cell_count(const kz_grid *g);
long kz_host_face_count(const kz_grid *g, int dim);

/* MetricCache.init (koral/metric/precompute.zig:175-213): MKS2 blocks at
 * centres (g, gcon: 20/cell; kris: 64/cell, sqrt(-g)-corrected) and faces
 * (gb_d, gconb_d: 20/face), in the reference's layouts. */
int kz_host_metric_tables(const kz_grid *g, double a, double mksr0, double mksh0, double *gc, double *gcon,
                          double *kris, double *gb0, double *gconb0, double *gb1, double *gconb1, double *gb2,
                          double *gconb2);

/* metric.compute at one point (koral/metric/metric.zig:64-104);
 * coords 0 Minkowski, 1 BL, 2 KS, 3 MKS2; out: gg16 GG16 gdet dlgdet3 Kr64 gttpert */
int kz_host_metric_point(int coords, double a, double mksr0, double mksh0, const double *x4, double *out101);
/* kz_geom records (MetricCache.fillGeometry, koral/metric/precompute.zig:
 * 499-507) at n points x4[4 n] of coordinate system `coords`. */
int kz_host_geom_batch(int coords, double a, double mksr0, double mksh0, int64_t n, const double *x4,
                       kz_geom *out);

/* coco.dxdx (koral/metric/coco.zig:144-160): dir 0 MKS2->BL, 1 BL->MKS2 */
int kz_host_dxdx(int dir, double a, double mksr0, double mksh0, const double *x4, double *out16);

/* Tables for the PUFFY outer radial boundary (koral/problems/puffy/puffy.zig:
 * 1089-1166): 66 doubles per xhi ghost (iz, iy, i): gBL[16] GBL[16] r_ghost
 * r_bound J_mks2->bl[16] J_bl->mks2[16]. */
int kz_host_bc_xhi_tables(const kz_grid *g, double a, double mksr0, double mksh0, double *out);

/* The BL sidecar of MetricCache (koral/metric/precompute.zig:361-380,
 * :457-475) for the radiative viscosity and the dynamo: 52 doubles per padded
 * cell, gBL[16] GBL[16] alpha_BL gdet_BL r theta J_mks2->bl[16]. */
int kz_host_bl_tables(const kz_grid *g, double a, double mksr0, double mksh0, double *out);

/* PUFFY prepinit + init over the domain (puffy.zig:715-866): padded AoS
 * primitives with A_phi in the B3 slot and ENTR = S(rho, u). */
int kz_host_puffy_prepinit(const kz_grid *g, double a, double mksr0, double mksh0, double gam,
                           const kz_opac_params *op, const kz_rad_params *rp, double mass, double rhoatmmin,
                           double atm_tgas, double atm_trad_init, double lt_kappa, const double *gc,
                           const double *gcon, double *p);

/* calc_BfromA with overwrite, 2D and 3D (koral/sim/ct.zig:184-334) */
int kz_host_calc_b_from_a(const kz_grid *g, const double *gc, double *p);

/* seeded multiplicative perturbation of the domain state (see kz_host.cpp) */
double kz_host_uniform(uint64_t seed, uint64_t cell, uint64_t var);
int kz_host_perturb(const kz_grid *g, uint64_t seed, double amp, double gam, int rho_u_only, double *p);

/* sim.zig:481-495: the largest radial index (ghosts included) whose cell
 * centre lies inside the BL horizon, kz_physics.reduce_order_ix_max for
 * reduceorderatbh; INT32_MIN when there is none */
int32_t kz_host_reduce_order_ix_max(const kz_grid *g, double a, double mksr0, double mksh0);
/* BL radius of each radial index (scalars.zig:47-50 radiusBL), r[nx] */
int kz_host_radius_bl(const kz_grid *g, double a, double mksr0, double mksh0, double *r);
/* BL theta and phi widths of the cells of radial shell ix, [nz][ny] each
 * (scalars.zig:188-204 blThetaWidth / blPhiWidth) */
int kz_host_shell_widths_bl(const kz_grid *g, double a, double mksr0, double mksh0, long ix, double *dth,
                            double *dph);
int kz_host_geometry_bl(const kz_grid *g, double a, double mksr0, double mksh0, long ix, long iy, long iz,
                        kz_geom *out);

#ifdef __cplusplus
}
#endif
#endif /* KZ_HOST_H */
