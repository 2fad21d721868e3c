/* kz_step.h — the koral per-step hot path on one MI355X (libkz_hip.so).
 *
 * A kz_sim holds one run's device-resident state: primitives, conserved
 * variables, the RK2IMEX stage copies, face fluxes, EMFs, per-cell flags
 * and the metric tables, all in HBM.  Each entry point replaces one method
 * of the reference's Sim(config.puffy) (koral/sim.zig:93-1900); the caller
 * keeps the reference's array formats at this boundary (padded AoS fields
 * iv-fastest, koral/field.zig; MetricCache 20-double blocks,
 * koral/metric/precompute.zig:121-526) and the library converts to its own
 * SoA device layout inside.
 *
 * Return codes are the KZ_OK / KZ_E_* values of kz_types.h.  Calls on one
 * kz_sim are not thread-safe; different kz_sim handles are independent.
 */
#ifndef KZ_STEP_H
#define KZ_STEP_H

#include <stdint.h>
#include "kz_types.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct kz_sim kz_sim;

/* Sim.init (koral/sim.zig:309-500) with the MetricCache blocks computed by
 * the caller (kz_host_metric_tables) and the PUFFY outer-boundary table
 * (kz_host_bc_xhi_tables). Returns NULL and sets *err on failure:
 * KZ_E_UNSUPPORTED for a grid with a collapsed x or y dimension or fewer
 * than 3 ghosts, KZ_E_INVALID_CONFIG for sim.zig:319-363's preconditions
 * (including radiative viscosity without opacities) and for a b^2/rho floor
 * frame other than drift (0) or ZAMO (1), the two values of
 * koral/solve/invert.zig:35.  With radviscosity or dynamo set, the BL sidecar must be
 * loaded with kz_sim_set_bl_tables before the first step. */
kz_sim *kz_sim_create(const kz_grid *grid, const kz_physics *phys, const double *g, const double *gcon,
                      const double *kris, const double *gb0, const double *gconb0, const double *gb1,
                      const double *gconb1, const double *gb2, const double *gconb2, const double *bcx, int *err);
void kz_sim_destroy(kz_sim *s);

/* State exchange in the reference Field layout (padded AoS, 13 slots per
 * cell; flags 6 int32 per cell). NULL pointers are skipped. Synchronous. */
int kz_sim_set_state(kz_sim *s, const double *p, const double *u);
int kz_sim_get_state(kz_sim *s, double *p, double *u, int32_t *flags);

/* Sim.initCell's p2u over the domain (koral/sim.zig:672-685). */
int kz_sim_p2u_domain(kz_sim *s);
/* ct.calcDivB (koral/sim/ct.zig:336-381) at every domain cell of the current
 * primitives: corner divergence of gdet*B over gdet, out[nz][ny][nx].
 * Reads one ghost layer below each active dimension. Synchronous. */
int kz_sim_calc_divb(kz_sim *s, double *out);
/* Sim.setBc (koral/sim.zig:735-745 -> sim/bc.zig:40-73, PUFFY Bc.calc
 * koral/problems/puffy/puffy.zig:1080-1194, fillCorners2d bc.zig:286-377). */
int kz_sim_set_bc(kz_sim *s, int ifinit);
/* Physics.postinit beta normalisation (puffy.zig:892-1024); *fac_out gets
 * the B scale factor. */
int kz_sim_postinit(kz_sim *s, double maxbeta, double *fac_out);
/* Sim.initTimestepGuess (koral/sim.zig:693-718). */
int kz_sim_init_timestep_guess(kz_sim *s);
/* Restore the step clock (t, tstepdenmax, tstepdenmin) of a resumed run. */
int kz_sim_set_time(kz_sim *s, double t, double tstepdenmax, double tstepdenmin);

/* The passes of one step, each synchronous:
 * Sim.calcU2p (koral/sim.zig:924-1026, with cellFixup :1027-1133 and setBc),
 * Sim.opExplicit (koral/sim.zig:1506-1582: calcWavespeeds :747-886,
 *   sweeps + fluxesAtFaces :1209-1503, sim/ct.zig:38-177 fluxCt, update),
 * Sim.opImplicit (koral/sim.zig:1715-1776),
 * Sim.doCorrect (koral/sim.zig:1783-1790 -> sim/polaraxis.zig:84-147). */
int kz_calc_u2p(kz_sim *s);
int kz_op_explicit(kz_sim *s, double dt);
int kz_op_implicit(kz_sim *s, double dt);
int kz_do_correct(kz_sim *s);

/* The MetricCache BL sidecar (koral/metric/precompute.zig:148,379,457-470:
 * blGeom + jacMy2Bl) as kz_host_bl_tables builds it: KZ_BL_REC doubles per
 * padded cell (BL g[16], G[16], alpha, gdet, r, theta, J_mks2->bl[16]).
 * Radiative viscosity, the dynamo and kz_sim_scalars read it; a run with
 * neither of the first two allocates its copy on the first call. */
int kz_sim_set_bl_tables(kz_sim *s, const double *bl);
/* calcRijViscTotal (koral/sim/rijvisc.zig:266-311): the viscous R^i_j on the
 * domain and its one-cell ring (corners 0) for the step size global_dt;
 * kz_sim_get_rijvisc copies it out, 16 doubles per padded cell. */
int kz_calc_rij_visc_total(kz_sim *s, double global_dt);
int kz_sim_get_rijvisc(kz_sim *s, double *out);
/* applyDynamo (koral/sim/dynamo.zig:462-469): scale height, setBc,
 * mimicDynamo over dt, calcU2p. */
int kz_apply_dynamo(kz_sim *s, double dt);

/* Sim.step (koral/sim.zig:1792-1899): one RK2IMEX step; forced_dt NULL
 * takes the CFL step 1/tstepdenmax. Synchronous; returns the first error
 * any kernel of the step reported. */
int kz_step(kz_sim *s, const double *forced_dt);

/* Enqueue nsteps CFL steps on the sim's stream without waiting (the
 * timestep stays on the device between steps); kz_sim_sync waits and
 * returns the first device-side error. */
int kz_step_many(kz_sim *s, int nsteps);
int kz_sim_sync(kz_sim *s);
/* Capture nsteps CFL steps into a HIP graph (the launches kz_step_many
 * enqueues; event timing off) and replay it: kz_sim_replay enqueues the
 * graph on the sim's stream without waiting and advances the run exactly
 * like kz_step_many(nsteps).  For run servers that keep many runs resident
 * and advance them in batches; no counterpart in the reference, whose steps
 * are host-driven. */
int kz_sim_capture_steps(kz_sim *s, int nsteps);
int kz_sim_replay(kz_sim *s);
/* Launch shape of the implicit pass (no counterpart in the reference; the
 * results are identical either way): 0 chooses by size (default), 1 one
 * first-rung launch over both starting-block lists plus one launch for the
 * later rungs, 2 persistent work-stealing launches (each lane takes the next
 * cell when its Newton ends), which pay when the cells span many rounds of
 * waves. */
int kz_sim_set_implicit_launch(kz_sim *s, int mode);
/* 0 when everything enqueued on the sim's stream has finished, 1 while
 * work is pending (a non-blocking poll; kz_sim_sync reports errors). */
int kz_sim_query(kz_sim *s);

/* Run groups (no counterpart in the reference, which steps one run per
 * process): resident runs stepped together, the implicit cells of all of
 * them solved by one set of implicit launches per implicit pass, each run's
 * other kernels on its own stream.  Every run ends each step bit for bit
 * where kz_step_many would leave it.  The runs must agree bitwise in the
 * physics the implicit solve reads (gam, rad, opac, implicit, has_opac and
 * the polar-axis rows); grids and spins may differ.  A run must leave the
 * group (kz_group_set_runs without it) before kz_sim_destroy. */
typedef struct kz_group kz_group;
int kz_group_create(kz_group **out);
void kz_group_destroy(kz_group *g);
/* Membership: waits for the previous members' work, grows the group's
 * solve scratch when the cells outgrow it, drops a captured graph, and
 * records each member's device layout (so call it again after
 * kz_sim_set_bl_tables on a member). */
int kz_group_set_runs(kz_group *g, kz_sim *const *sims, int nsims);
/* as kz_sim_set_implicit_launch, for the group's implicit passes */
int kz_group_set_implicit_launch(kz_group *g, int mode);
/* enqueue nsteps CFL steps of every member without waiting */
int kz_group_step_many(kz_group *g, int nsteps);
/* capture nsteps steps of every member into one HIP graph; replay enqueues
 * it on the group's stream (event timing must be off on every member) */
int kz_group_capture_steps(kz_group *g, int nsteps);
int kz_group_replay(kz_group *g);
/* 0 when every enqueued step of every member has finished, 1 while pending */
int kz_group_query(kz_group *g);
/* wait, then the first device-side error of any member */
int kz_group_sync(kz_group *g);
/* One scalars.dat row of the current state (koral/io/dump.zig:383-411
 * scalarRow on one rank, over io/scalars.zig's reductions): mass, mdot
 * through radial shell ix_h (sign-flipped), radiative and total luminosity
 * through shell ix_l in the BL frame, scale height at ix_s, max pmag/ptot,
 * the fixup/NaN counts and the solver counters.  dth_bl, dph_bl: the BL
 * theta and phi widths of shell ix_l's cells, [nz][ny]
 * (kz_host_shell_widths_bl); the shell indices come from kz_host_radius_bl
 * (scalars.zig:56-66 radialShellIndex).  Needs kz_sim_set_bl_tables; the
 * sums run in the reference's serial order, so the row is bitwise a serial
 * evaluation's. */
int kz_sim_scalars(kz_sim *s, double dt, long ix_h, long ix_l, long ix_s, const double *dth_bl, const double *dph_bl,
                   kz_scalar_row *row);
/* Device memory this run holds (bytes), for admission by HBM footprint. */
int kz_sim_device_bytes(kz_sim *s, uint64_t *bytes);
/* The hipStream_t the sim enqueues on. */
void *kz_sim_stream(kz_sim *s);

/* The Sim fields the reference's self-goldens record (t, dt, own_dt,
 * tstepdenmax/min, nstep, implicit-solve counters). Synchronous. */
int kz_sim_counters(kz_sim *s, kz_counters *c);

/* KDMP v2 checkpoints (koral/io/dump.zig:19-160): kz_sim_write_kdmp
 * serializes the header (domain shape, nv, the run's t and nstep, out_idx)
 * and the domain primitives p[nz][ny][nx][nv] (serializePrimDump :101-135)
 * into buf (kz_sim_kdmp_size bytes; KZ_E_TRUNCATED if cap is smaller);
 * kz_sim_load_kdmp validates magic, version, shape and length
 * (KZ_E_BAD_MAGIC / _BAD_VERSION / _DIM_MISMATCH / _TRUNCATED, the
 * reference's ReadError), stores the primitives and u = p2u(p) in every
 * domain cell (loadPrimDump :137-160) and restores t and nstep (the driver's
 * restart, problems/puffy/main.zig:283-298); kz_sim_finish_init then fills
 * the ghosts and the startup dt guess (Sim.finishInit, koral/sim.zig:
 * 686-690).  kz_kdmp_parse_header is parseDumpHeader (:82-98). */
int kz_kdmp_parse_header(const uint8_t *buf, uint64_t len, kz_dump_header *h);
int kz_sim_kdmp_size(kz_sim *s, uint64_t *bytes);
int kz_sim_write_kdmp(kz_sim *s, uint32_t out_idx, uint8_t *buf, uint64_t cap, uint64_t *written);
int kz_sim_load_kdmp(kz_sim *s, const uint8_t *buf, uint64_t len, kz_dump_header *h);
int kz_sim_finish_init(kz_sim *s);

/* Kernel classes of a step, for event timing. */
enum {
  KZ_K_WAVESPEEDS = 0, /* calcWavespeeds */
  KZ_K_FACE_FLUX = 1,  /* fused PPM + floors + fFluxPrime + LAXF, per dim */
  KZ_K_CT = 2,         /* EMF + flux rebuild */
  KZ_K_UPDATE = 3,     /* flux divergence + metric source */
  KZ_K_U2P = 4,        /* u2p + floors */
  KZ_K_IMPLICIT = 5,   /* implicit four-force Newton */
  KZ_K_BC = 6,         /* boundary conditions */
  KZ_K_VISC = 7,       /* viscous R^i_j (radiative viscosity) */
  KZ_K_DYNAMO = 8,     /* scale height + Delta A_phi + curl (dynamo) */
  KZ_K_COUNT = 9
};
/* Time every launch of one kernel class with HIP events recorded on the
 * sim's stream (kernel = -1 turns timing off); kz_sim_kernel_time returns
 * the summed elapsed milliseconds and the launch count since the last
 * call and restarts the accumulation.  Profiling aid with no counterpart
 * in the reference (its per-phase timers are host wall clocks). */
int kz_sim_time_kernel(kz_sim *s, int kernel);
int kz_sim_kernel_time(kz_sim *s, double *ms_total, uint64_t *launches);

#ifdef __cplusplus
}
#endif

#endif /* KZ_STEP_H */
