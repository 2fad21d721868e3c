"""ctypes mirrors of include/kz_types.h and the PUFFY parameter presets.

The presets restate the reference's validated constants:
  FloorParams.puffy      koral/solve/invert.zig:81-88
  RadParams.puffy        koral/solve/invert_rad.zig:48-54
  ImplicitParams.puffy   koral/solve/implicit.zig:138-149
  thermo.Consts.init     koral/physics/thermo.zig:107-144 (C global shapes)
  Composition.puffy      koral/physics/thermo.zig:53-57
  radforce.Params.puffy  koral/physics/radforce.zig:98-124
  Physics.toOptions      koral/problems/puffy/puffy.zig:210-242
Every derived constant is evaluated with the reference's expression shape
(same operand order and association) in IEEE binary64 — Python floats.
"""
import ctypes as C
import math

import numpy as np

NV = 13
RHO, UU, VX, VY, VZ, ENTR, B1, B2, B3, EE, FX, FY, FZ = range(13)
NG = 3
NFLAGS = 6
FLAG_ENTROPY, FLAG_ENTROPY2, FLAG_ENTROPY3, FLAG_HD_FIXUP, FLAG_RAD_FIXUP, FLAG_RADIMP_FIXUP = range(6)
GEOM_DOUBLES = 47  # kz_geom: gg[4][5] GG[4][5] gdet alpha gttpert xx[4]

ERRORS = {0: "ok", -1: "SpacelikeVelocity", -2: "VelocityConversionFailed", -3: "NanInFlux",
          -4: "OutOfMemory", -5: "InvalidConfig", -6: "HipError", -7: "Unsupported", -8: "BadMagic",
          -9: "BadVersion", -10: "DimMismatch", -11: "Truncated"}


class kz_floor_params(C.Structure):
    _fields_ = [("rhofloor", C.c_double), ("uurhoratiomin", C.c_double), ("uurhoratiomax", C.c_double),
                ("b2rhoratiomax", C.c_double), ("b2uuratiomax", C.c_double), ("gammamaxhd", C.c_double),
                ("horizon_x1", C.c_double), ("b2rhofloorframe", C.c_int32),
                ("isentropic_b2rhofloors", C.c_int32), ("b2uufloor", C.c_int32), ("_pad", C.c_int32)]


class kz_rad_params(C.Structure):
    _fields_ = [("gammamaxrad", C.c_double), ("eradfloor", C.c_double), ("eerhoratiomin", C.c_double),
                ("eerhoratiomax", C.c_double), ("eeuuratiomin", C.c_double), ("eeuuratiomax", C.c_double)]


class kz_implicit_params(C.Structure):
    _fields_ = [(n, C.c_double) for n in (
        "conv", "entrconv", "eps", "convrel", "convrelerr", "convrelentr", "convrelentrerr",
        "damping_factor", "max_en_change_down", "max_en_change_up", "max_damping",
        "max_tgas_change", "max_trad_change", "threshold", "opdamp_factor")] + \
        [(n, C.c_int32) for n in (
            "maxiter", "start_with_bisect", "scale_jacobian", "allow_rad_ceiling",
            "allow_entr_in_4dprim", "opdamp_maxlevels", "lag_opac", "jacobian_fp32")]


class kz_opac_params(C.Structure):
    _fields_ = [(n, C.c_double) for n in (
        "hfrac", "hefrac", "a_inv_mean", "z2diva_mean",
        "rhogu2cgs", "numdensgu2cgs", "kappacgs2gu",
        "k_boltz_cgs", "h_cgs", "sigma_rad_cgs", "m_proton_cgs",
        "fourpi", "sigma_rad_over_pi", "four_sigmarad", "one_over_four_sigmarad",
        "k_over_mecsq", "kb_over_me", "kb_over_mugas_mp", "mugas_mp_over_kb",
        "one_over_mue_mp", "one_over_log_2p6",
        "kappa_cgs2gu_factor", "enden_gu2cgs_factor", "kappa_grey", "kappaes_grey", "opdamp")] + \
        [(n, C.c_int32) for n in (
            "bremsstrahlung", "synchrotron", "kleinnishina", "comptonization",
            "synchrotron_bridge", "kappa_mode", "kappaes_mode", "_pad")]


class kz_dump_header(C.Structure):
    _fields_ = [("nx", C.c_uint32), ("ny", C.c_uint32), ("nz", C.c_uint32), ("nv", C.c_uint32), ("t", C.c_double),
                ("nstep", C.c_uint64), ("out_idx", C.c_uint32), ("_pad", C.c_uint32)]


KDMP_HEADER_BYTES = 44


class kz_radvisc_params(C.Structure):
    _fields_ = [("alpha", C.c_double), ("maxvel", C.c_double)]


class kz_dynamo_params(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("alphadynamo", "thetaangle", "alphabeta", "betasaturated",
                                          "expectedhr")] + \
        [(n, C.c_int32) for n in ("alphaflipssign", "dampbeta", "calchronthego", "_pad")]


class kz_physics(C.Structure):
    _fields_ = [("gam", C.c_double), ("tsteplim", C.c_double), ("minmod_theta", C.c_double),
                ("floors", kz_floor_params), ("rad", kz_rad_params), ("opac", kz_opac_params),
                ("implicit", kz_implicit_params)] + \
        [(n, C.c_int32) for n in (
            "has_opac", "do_fixups", "do_u2prad_fixups", "do_radimp_fixups", "correct_polaraxis",
            "nccorrectpolar", "reduceorderatbh", "reduceorderafterfixup", "dampradwavespeednearaxis",
            "radviscosity", "dynamo", "reduce_order_ix_max")] + \
        [("radvisc", kz_radvisc_params), ("dynamo_params", kz_dynamo_params), ("bh_spin", C.c_double)]


class kz_grid(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32), ("ng", C.c_int32),
                ("minx", C.c_double), ("maxx", C.c_double), ("miny", C.c_double), ("maxy", C.c_double),
                ("minz", C.c_double), ("maxz", C.c_double)]


class kz_implicit_result(C.Structure):
    _fields_ = [(n, C.c_int32) for n in ("ok", "rung", "iters", "whichprim", "whicheq", "whichframe")]


class kz_u2p_result(C.Structure):
    _fields_ = [(n, C.c_int32) for n in ("ret", "entropy_used", "fixup", "_pad")]


class kz_u2prad_result(C.Structure):
    _fields_ = [("ret", C.c_int32), ("corrected", C.c_int32)]


class kz_counters(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("t", "dt", "own_dt", "tstepdenmax", "tstepdenmin")] + \
        [(n, C.c_uint64) for n in ("nstep", "n_radimp_failures", "n_radimp_iters", "n_radimp_solves",
                                    "n_guard_recoveries")]


class kz_scalar_row(C.Structure):
    """One scalars.dat row (koral/io/dump.zig:414-436 ScalarRow)."""
    _fields_ = [("t", C.c_double), ("dt", C.c_double), ("nstep", C.c_uint64)] + \
        [(n, C.c_double) for n in ("mass", "mdot", "radlum", "totallum", "scaleheight", "max_pmag_ptot")] + \
        [(n, C.c_uint64) for n in ("n_hd_fixup", "n_radimp_fail", "n_nan", "n_floorguard")]


IMPLICIT_RESULT_DTYPE = np.dtype([(n, np.int32) for n in ("ok", "rung", "iters", "whichprim", "whicheq",
                                                          "whichframe")])
U2P_RESULT_DTYPE = np.dtype([(n, np.int32) for n in ("ret", "entropy_used", "fixup", "_pad")])
U2PRAD_RESULT_DTYPE = np.dtype([("ret", np.int32), ("corrected", np.int32)])

# ---- reference constants (koral/units.zig:14-34) ----------------------------
GGG = 6.674e-8 / 1.0
CCC = 2.9979246e10 / 1.0
MSUNCM = 147700.0
K_BOLTZ_CGS = 1.3806488e-16
M_ELECTR_CGS = 9.1094e-28
M_PROTON_CGS = 1.67262158e-24
SIGMA_RAD_CGS = 5.670367e-5
H_CGS = 6.6260755e-27


def floor_params_puffy():
    """FloorParams.puffy (koral/solve/invert.zig:81-88), drift frame."""
    return kz_floor_params(rhofloor=1.0e-30, uurhoratiomin=1.0e-8, uurhoratiomax=1.0e0,
                           b2rhoratiomax=50.0, b2uuratiomax=50.0, gammamaxhd=10.0,
                           horizon_x1=-math.inf, b2rhofloorframe=0, isentropic_b2rhofloors=0,
                           b2uufloor=1)


def rad_params_puffy():
    """RadParams.puffy (koral/solve/invert_rad.zig:48-54)."""
    return kz_rad_params(gammamaxrad=10.0, eradfloor=10.0 * 1.0e-80, eerhoratiomin=1.0e-20,
                         eerhoratiomax=1.0e4, eeuuratiomin=1.0e-20, eeuuratiomax=1.0e20)


def implicit_params_puffy():
    """ImplicitParams.puffy (koral/solve/implicit.zig:138-149 over :81-127 defaults)."""
    return kz_implicit_params(conv=1.0e-10, entrconv=1.0e-5, eps=1.0e-6, convrel=1.0e-8,
                              convrelerr=1.0e-4, convrelentr=1.0e-4, convrelentrerr=1.0e-2,
                              damping_factor=3.0, max_en_change_down=10.0, max_en_change_up=10.0,
                              max_damping=1.0e-3, max_tgas_change=10.0, max_trad_change=10.0,
                              threshold=1.0e-2, opdamp_factor=3.0, maxiter=40, start_with_bisect=1,
                              scale_jacobian=1, allow_rad_ceiling=1, allow_entr_in_4dprim=1,
                              opdamp_maxlevels=0, lag_opac=0)


def opac_params_puffy(mass=10.0):
    """thermo.Consts.init + radforce.Params.puffyMass (thermo.zig:107-144,
    radforce.zig:98-124): MASS solar masses, HFRAC=1, MU_GAS/MU_I/MU_E = 1/2/2,
    all channels on, Klein-Nishina scattering."""
    masscm = mass * MSUNCM
    hfrac, hefrac = 1.0, 0.0
    a_inv_mean, z2diva_mean = 0.0570490404519, 5.14445150955
    mu_gas, mu_i, mu_e = 1.0, 2.0, 2.0
    k_boltz = K_BOLTZ_CGS * GGG / CCC / CCC / CCC / CCC / masscm
    m_proton = M_PROTON_CGS * (GGG / (masscm * CCC * CCC))
    m_electr = M_ELECTR_CGS * (GGG / (masscm * CCC * CCC))
    sigma_rad = SIGMA_RAD_CGS * GGG / CCC / CCC / CCC / CCC / CCC * masscm * masscm
    four_sigmarad = 4.0 * sigma_rad
    pi = math.pi
    return kz_opac_params(
        hfrac=hfrac, hefrac=hefrac, a_inv_mean=a_inv_mean, z2diva_mean=z2diva_mean,
        rhogu2cgs=1.0 / GGG * CCC * CCC / masscm / masscm,
        numdensgu2cgs=1.0 / masscm / masscm / masscm,
        kappacgs2gu=1.0 / GGG * CCC * CCC / masscm,
        k_boltz_cgs=K_BOLTZ_CGS, h_cgs=H_CGS, sigma_rad_cgs=SIGMA_RAD_CGS, m_proton_cgs=M_PROTON_CGS,
        fourpi=4.0 * pi, sigma_rad_over_pi=sigma_rad / pi, four_sigmarad=four_sigmarad,
        one_over_four_sigmarad=1.0 / four_sigmarad, k_over_mecsq=1.69e-10,
        kb_over_me=k_boltz / m_electr, kb_over_mugas_mp=k_boltz / mu_gas / m_proton,
        mugas_mp_over_kb=1.0 / (k_boltz / mu_gas / m_proton),
        one_over_mue_mp=1.0 / mu_e / m_proton, one_over_log_2p6=1.0 / math.log(2.6),
        kappa_cgs2gu_factor=(CCC * CCC) / (GGG * masscm),
        enden_gu2cgs_factor=(CCC * CCC * CCC * CCC) / (GGG * masscm * masscm),
        kappa_grey=0.0, kappaes_grey=0.0, opdamp=1.0,
        bremsstrahlung=1, synchrotron=1, kleinnishina=1, comptonization=1, synchrotron_bridge=0,
        kappa_mode=0, kappaes_mode=1)


def radvisc_params_puffy():
    """radvisc.Params defaults = PUFFY define.h:96-102 (koral/physics/radvisc.zig:44-53)."""
    return kz_radvisc_params(alpha=0.1, maxvel=0.1)


def dynamo_params_puffy():
    """dynamo.Params defaults = PUFFY define.h:66-74 (koral/sim/dynamo.zig:54-64)."""
    return kz_dynamo_params(alphadynamo=2.0 * 0.314, thetaangle=0.25, alphabeta=2.0 * 6.28,
                            betasaturated=0.1, expectedhr=0.3, alphaflipssign=1, dampbeta=1,
                            calchronthego=1)


def physics_puffy(mass=10.0, radviscosity=False, dynamo=False, bh_spin=0.0):
    """The per-cell knobs of Physics.toOptions (koral/problems/puffy/puffy.zig:210-242).
    Radiative viscosity and the mimic dynamo (on in PUFFY production runs) are
    off unless asked for; bh_spin sets r_H and r_ISCO for them."""
    return kz_physics(gam=5.0 / 3.0, tsteplim=0.5, minmod_theta=1.5,
                      floors=floor_params_puffy(), rad=rad_params_puffy(),
                      opac=opac_params_puffy(mass), implicit=implicit_params_puffy(),
                      has_opac=1, do_fixups=1, do_u2prad_fixups=0, do_radimp_fixups=0,
                      correct_polaraxis=1, nccorrectpolar=2, reduceorderatbh=0,
                      reduceorderafterfixup=0, dampradwavespeednearaxis=0,
                      radviscosity=int(radviscosity), dynamo=int(dynamo),
                      radvisc=radvisc_params_puffy(), dynamo_params=dynamo_params_puffy(), bh_spin=bh_spin)


def rmin_for_spin(a):
    """puffy.zig:75-77 rminForSpin: RMIN(a) = 1.85 r_H(a) / r_H(0)."""
    return 1.85 * (1.0 + math.sqrt(1.0 - a * a)) / 2.0


def grid_puffy(nx, ny, nz=1, rmin=1.85, rmax=500.0, mksr0=0.1):
    """Physics.makeGridNz (koral/problems/puffy/puffy.zig:179-192)."""
    return kz_grid(nx=nx, ny=ny, nz=nz, ng=3,
                   minx=math.log(rmin - mksr0), maxx=math.log(rmax - mksr0),
                   miny=0.001, maxy=1.0 - 0.001, minz=-math.pi / 4.0, maxz=math.pi / 4.0)


def padded_shape(g):
    """(sz, sy, sx) of the padded cell array (koral/grid.zig:96-107)."""
    ngx = g.ng if g.nx > 1 else 0
    ngy = g.ng if g.ny > 1 else 0
    ngz = g.ng if g.nz > 1 else 0
    return (g.nz + 2 * ngz, g.ny + 2 * ngy, g.nx + 2 * ngx), (ngz, ngy, ngx)


def ptr(a, ctype=C.c_double):
    """Raw pointer to a contiguous numpy array (None passes NULL)."""
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"], "array must be C-contiguous"
    return a.ctypes.data_as(C.POINTER(ctype))
