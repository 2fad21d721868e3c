"""scalars.dat: the diagnostic time series of a PUFFY run.

The row's reductions run on the device (kz_sim_scalars) over the resident
state; this module supplies what the reference's driver supplies around
them (koral/problems/puffy/main.zig:28-31, io/dump.zig:383-446):

  * the diagnostic radii: the horizon r_H(a) for the mass flux, r = 5000
    (beyond RMAX, so the outermost shell) for the luminosity, r = 15 for the
    scale height, mapped to radial indices by radialShellIndex
    (scalars.zig:56-66: the first index whose BL radius exceeds the radius,
    else the last one);
  * the BL widths of the luminosity shell's cells (host library);
  * the header and the text of each row (appendScalarLine's precision).

The engine is anything with `scalars(dt, ix_h, ix_l, ix_s, dth, dph)`: the
HIP sim, or the CPU restatement in tests.
"""
import math

import numpy as np

from . import host

HEADER = "# t dt nstep mass mdot radlum totallum H/R maxPmag/Ptot n_hdfix n_radimpfail n_nan n_floorguard\n"
COLUMNS = ("t", "dt", "nstep", "mass", "mdot", "radlum", "totallum", "scaleheight", "max_pmag_ptot", "n_hd_fixup",
           "n_radimp_fail", "n_nan", "n_floorguard")
R_LUM = 5000.0
R_SCALE = 15.0


def r_horizon_bl(a):
    """metric.zig:20-23 rHorizonBL."""
    return 1.0 + math.sqrt(1.0 - a * a)


def radial_shell_index(r_bl, radius):
    """scalars.zig:56-66 radialShellIndex over the BL radii of the radial indices."""
    above = np.nonzero(np.asarray(r_bl) > radius)[0]
    return int(above[0]) if above.size else len(r_bl) - 1


class Diagnostics:
    """The shell indices and luminosity-shell widths of one grid, computed
    once; row(engine, dt) evaluates one scalars.dat row."""

    def __init__(self, grid, mp=host.MP_PUFFY, r_lum=R_LUM, r_scale=R_SCALE):
        r = host.radius_bl(grid, mp)
        self.ix_h = radial_shell_index(r, r_horizon_bl(mp["a"]))
        self.ix_l = radial_shell_index(r, r_lum)
        self.ix_s = radial_shell_index(r, r_scale)
        self.dth, self.dph = host.shell_widths_bl(grid, self.ix_l, mp)

    def row(self, engine, dt):
        return engine.scalars(dt, self.ix_h, self.ix_l, self.ix_s, self.dth, self.dph)


def _zig_e(x, prec):
    """A float as Zig's `{e:.N}` prints it: mantissa with N decimals, then
    `e` and the exponent without sign padding (1.5e3, 2.0e-7)."""
    if x != x:
        return "nan"
    if math.isinf(x):
        return "inf" if x > 0 else "-inf"
    m, e = f"{x:.{prec}e}".split("e")
    return f"{m}e{int(e)}"


def format_row(row):
    """dump.zig:438-446 appendScalarLine."""
    f = row
    return (f"{_zig_e(f['t'], 10)} {_zig_e(f['dt'], 6)} {int(f['nstep'])} {_zig_e(f['mass'], 10)} "
            f"{_zig_e(f['mdot'], 10)} {_zig_e(f['radlum'], 10)} {_zig_e(f['totallum'], 10)} "
            f"{_zig_e(f['scaleheight'], 6)} {_zig_e(f['max_pmag_ptot'], 6)} {int(f['n_hd_fixup'])} "
            f"{int(f['n_radimp_fail'])} {int(f['n_nan'])} {int(f['n_floorguard'])}\n")


class ScalarsLog:
    """The growing scalars.dat text; write() rewrites the file from it, as
    the reference's driver does at each output."""

    def __init__(self):
        self.text = HEADER
        self.rows = []

    def append(self, row):
        self.rows.append(dict(row))
        self.text += format_row(row)

    def write(self, path):
        with open(path, "w") as fh:
            fh.write(self.text)
