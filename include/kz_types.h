/* kz_types.h — plain-old-data types exchanged across the koral step C-ABI.
 *
 * Every struct here is a flat restatement of a reference parameter/record
 * type; the citation names the Zig declaration whose fields it carries.
 * All reals are IEEE-754 binary64, all flags int32 (0/1), little-endian,
 * natural alignment (no packing pragmas) so ctypes and Zig `extern struct`
 * see the same layout.
 */
#ifndef KZ_TYPES_H
#define KZ_TYPES_H

#include <stdint.h>

#define KZ_NV 13
/* doubles per padded cell of the BL sidecar (kz_host_bl_tables) */
#define KZ_BL_REC 52
/* primitive / conserved slot map (koral/layout.zig:114-137, config.puffy) */
enum {
  KZ_RHO = 0, KZ_UU = 1, KZ_VX = 2, KZ_VY = 3, KZ_VZ = 4, KZ_ENTR = 5,
  KZ_B1 = 6, KZ_B2 = 7, KZ_B3 = 8, KZ_EE = 9, KZ_FX = 10, KZ_FY = 11, KZ_FZ = 12
};
#define KZ_NG 3 /* ghost depth for PPM (koral/config.zig:30-35) */

/* per-cell flag slots (koral/sim/storage.zig:11-18) */
enum {
  KZ_FLAG_ENTROPY = 0, KZ_FLAG_ENTROPY2 = 1, KZ_FLAG_ENTROPY3 = 2,
  KZ_FLAG_HD_FIXUP = 3, KZ_FLAG_RAD_FIXUP = 4, KZ_FLAG_RADIMP_FIXUP = 5,
  KZ_NFLAGS = 6
};

/* return codes of the pass-level entry points */
enum {
  KZ_OK = 0,
  KZ_E_SPACELIKE = -1,       /* relele.Error.SpacelikeVelocity */
  KZ_E_VELCONV = -2,         /* relele.Error.VelocityConversionFailed */
  KZ_E_NAN_IN_FLUX = -3,     /* relele.Error.NanInFlux */
  KZ_E_OOM = -4,
  KZ_E_INVALID_CONFIG = -5,  /* Sim.init preconditions (koral/sim.zig:319-363) */
  KZ_E_HIP = -6,             /* a HIP runtime call failed */
  KZ_E_UNSUPPORTED = -7,     /* a grid this build does not run: a collapsed x or y dimension, < 3 ghosts */
  KZ_E_BAD_MAGIC = -8,       /* dump.ReadError.BadMagic (koral/io/dump.zig:43) */
  KZ_E_BAD_VERSION = -9,     /* dump.ReadError.BadVersion */
  KZ_E_DIM_MISMATCH = -10,   /* dump.ReadError.DimMismatch */
  KZ_E_TRUNCATED = -11       /* dump.ReadError.Truncated */
};

/* Point geometry (koral/geometry.zig:29-49). gg/GG keep the reference's 4x5
 * blocks: column 4 holds dlgdet[0..2], gdet (gg) and 0,0,0,gttpert (GG). */
typedef struct kz_geom {
  double gg[4][5];
  double GG[4][5];
  double gdet;
  double alpha;
  double gttpert;
  double xx[4];
} kz_geom;

/* MHD floors (koral/solve/invert.zig:47-89). b2rhofloorframe: 0 = drift
 * frame (the validated build); 1 = ZAMO frame (invert.zig:649-684). */
typedef struct kz_floor_params {
  double rhofloor;
  double uurhoratiomin;
  double uurhoratiomax;
  double b2rhoratiomax;
  double b2uuratiomax;
  double gammamaxhd;
  double horizon_x1;
  int32_t b2rhofloorframe;
  int32_t isentropic_b2rhofloors;
  int32_t b2uufloor;
  int32_t _pad;
} kz_floor_params;

/* radiation caps/floors (koral/solve/invert_rad.zig:37-55) */
typedef struct kz_rad_params {
  double gammamaxrad;
  double eradfloor;
  double eerhoratiomin;
  double eerhoratiomax;
  double eeuuratiomin;
  double eeuuratiomax;
} kz_rad_params;

/* implicit solver knobs (koral/solve/implicit.zig:81-150) */
typedef struct kz_implicit_params {
  double conv;
  double entrconv;
  double eps;
  double convrel;
  double convrelerr;
  double convrelentr;
  double convrelentrerr;
  double damping_factor;
  double max_en_change_down;
  double max_en_change_up;
  double max_damping;
  double max_tgas_change;
  double max_trad_change;
  double threshold;
  double opdamp_factor;
  int32_t maxiter;
  int32_t start_with_bisect;
  int32_t scale_jacobian;
  int32_t allow_rad_ceiling;
  int32_t allow_entr_in_4dprim;
  int32_t opdamp_maxlevels;
  int32_t lag_opac;
  /* Not in the reference: nonzero rounds the finite-difference Jacobian's
   * quotients to FP32 and scales and LU-factors it in FP32 (the residuals,
   * the constraint inversions and the Newton update stay FP64); 0 is the
   * reference's all-FP64 solve. */
  int32_t jacobian_fp32;
} kz_implicit_params;

/* Opacity / four-force constants: thermo.Consts (koral/physics/thermo.zig:
 * 107-144), units.Units conversion factors used by the lane-generic helpers
 * (koral/units.zig), the composition (thermo.zig:39-73), the channel switches
 * (koral/physics/opacities.zig:65-82) and radforce.Params
 * (koral/physics/radforce.zig:74-151). */
typedef struct kz_opac_params {
  double hfrac, hefrac, a_inv_mean, z2diva_mean;
  double rhogu2cgs, numdensgu2cgs, kappacgs2gu;
  double k_boltz_cgs, h_cgs, sigma_rad_cgs, m_proton_cgs;
  double fourpi, sigma_rad_over_pi, four_sigmarad, one_over_four_sigmarad;
  double k_over_mecsq, kb_over_me, kb_over_mugas_mp, mugas_mp_over_kb;
  double one_over_mue_mp, one_over_log_2p6;
  double kappa_cgs2gu_factor;  /* (CCC*CCC)/(GGG*masscm), units.zig kappaCgs2GuG */
  double enden_gu2cgs_factor;  /* (CCC^4)/(GGG*masscm^2), units.zig endenGu2CgsG */
  double kappa_grey;           /* kappa_mode == 1: kappa = kappa_grey * rho */
  double kappaes_grey;         /* kappaes_mode == 2: kappaes = kappaes_grey * rho */
  double opdamp;               /* four-force scale (1.0 = undamped) */
  int32_t bremsstrahlung, synchrotron, kleinnishina, comptonization;
  int32_t synchrotron_bridge;  /* Ramesh NR bridge instead of the Terelfactor (opacities.zig:70-74) */
  int32_t kappa_mode;          /* 0 default (opacities), 1 grey */
  int32_t kappaes_mode;        /* 0 none, 1 PUFFY Klein-Nishina, 2 grey */
  int32_t _pad;
} kz_opac_params;

/* radiative shear viscosity knobs (koral/physics/radvisc.zig:44-53):
 * nu = alpha * mfp; the viscous flux is damped to speed <= maxvel */
typedef struct kz_radvisc_params {
  double alpha;
  double maxvel;
} kz_radvisc_params;

/* mean-field "mimic" dynamo knobs (koral/sim/dynamo.zig:54-64) */
typedef struct kz_dynamo_params {
  double alphadynamo, thetaangle, alphabeta, betasaturated, expectedhr;
  int32_t alphaflipssign, dampbeta, calchronthego, _pad;
} kz_dynamo_params;

/* Everything the per-cell kernels read (Sim.Options subset, koral/sim.zig:
 * 130-204, as set by Physics.toOptions, koral/problems/puffy/puffy.zig:210-242). */
typedef struct kz_physics {
  double gam;          /* adiabatic index */
  double tsteplim;     /* CFL factor */
  double minmod_theta; /* linear reconstruction theta */
  kz_floor_params floors;
  kz_rad_params rad;
  kz_opac_params opac;
  kz_implicit_params implicit;
  int32_t has_opac;          /* 0 = SKIPRADSOURCE (no implicit, thin wavespeeds) */
  int32_t do_fixups;         /* DOFIXUPS && DOU2PMHDFIXUPS */
  int32_t do_u2prad_fixups;
  int32_t do_radimp_fixups;
  int32_t correct_polaraxis;
  int32_t nccorrectpolar;
  int32_t reduceorderatbh;        /* cells with radial index <= reduce_order_ix_max reconstruct
                                     one order lower (sim.zig:1316-1320) */
  int32_t reduceorderafterfixup;  /* cells with a fixup flag set reconstruct one order lower
                                     (sim.zig:1321-1332) */
  int32_t dampradwavespeednearaxis;
  int32_t radviscosity;           /* radiative shear viscosity (needs the BL tables) */
  int32_t dynamo;                 /* mean-field dynamo after each explicit stage (BL tables) */
  int32_t reduce_order_ix_max;    /* the largest radial index, ghosts included, whose centre lies
                                     inside the BL horizon (sim.zig:481-495); read only with
                                     reduceorderatbh; kz_host_reduce_order_ix_max computes it */
  kz_radvisc_params radvisc;
  kz_dynamo_params dynamo_params;
  double bh_spin;                 /* metric spin a (Options.mp.a) */
} kz_physics;

/* Grid (koral/grid.zig:10-161): uniform internal coordinates, NG ghosts on
 * every active (n > 1) dimension. */
typedef struct kz_grid {
  int32_t nx, ny, nz, ng;
  double minx, maxx, miny, maxy, minz, maxz;
} kz_grid;

/* Implicit-solve outcome (koral/solve/implicit.zig:152-164). */
typedef struct kz_implicit_result {
  int32_t ok;
  int32_t rung;      /* 0..5, -1 on failure */
  int32_t iters;
  int32_t whichprim; /* 1 rad, 2 mhd */
  int32_t whicheq;   /* 0 energy, 1 entropy */
  int32_t whichframe;/* 0 lab, 1 ff */
} kz_implicit_result;

/* u2p outcome (koral/solve/invert.zig:406-413). */
typedef struct kz_u2p_result {
  int32_t ret;
  int32_t entropy_used;
  int32_t fixup;
  int32_t _pad;
} kz_u2p_result;

/* u2p_rad outcome (koral/solve/invert_rad.zig:57-62). */
typedef struct kz_u2prad_result {
  int32_t ret;
  int32_t corrected;
} kz_u2prad_result;

/* Per-step counters (the Sim fields the self-goldens record,
 * koral/testing/selfgolden.zig:52-88, plus the solve count). */
/* One row of scalars.dat (koral/io/dump.zig:414-436 ScalarRow): code units,
 * mdot sign-flipped so that accretion is positive. */
typedef struct kz_scalar_row {
  double t, dt;
  uint64_t nstep;
  double mass, mdot, radlum, totallum, scaleheight, max_pmag_ptot;
  uint64_t n_hd_fixup, n_radimp_fail, n_nan, n_floorguard;
} kz_scalar_row;

typedef struct kz_counters {
  double t, dt, own_dt, tstepdenmax, tstepdenmin;
  uint64_t nstep;
  uint64_t n_radimp_failures, n_radimp_iters, n_radimp_solves;
  uint64_t n_guard_recoveries;
} kz_counters;

/* KDMP v2 checkpoint header (koral/io/dump.zig:29-45 DumpHeader): the
 * domain cell counts, the primitive count, the clock, the step count and the
 * output frame index.  On disk: "KDMP", u32 version 2, u32 nx ny nz nv, f64 t,
 * u64 nstep, u32 out_idx, little-endian (44 bytes), then p[nz][ny][nx][nv]. */
typedef struct kz_dump_header {
  uint32_t nx, ny, nz, nv;
  double t;
  uint64_t nstep;
  uint32_t out_idx;
  uint32_t _pad;
} kz_dump_header;
#define KZ_KDMP_HEADER_BYTES 44

#endif /* KZ_TYPES_H */
