"""Readers for the reference's committed golden fixtures (tests/golden/).

Formats restated from koral/testing/golden.zig:32-369 (KGLD, KSTP, KINI) and
koral/testing/selfgolden.zig:20-130 (KSLF); all little-endian f64 payloads.
"""
import gzip
import struct
from pathlib import Path

import numpy as np

GOLDEN = Path(__file__).resolve().parent / "golden"


def _raw(path):
    b = Path(path).read_bytes()
    if b[:2] == b"\x1f\x8b":
        b = gzip.decompress(b)
    return b


def read_kgld(rel, nin, nout):
    """golden.zig:32-60 — "KGLD" u32 ver, u64 nrec, u32 nin, u32 nout, f64 recs."""
    b = _raw(GOLDEN / rel)
    assert b[:4] == b"KGLD"
    ver, nrec, fin, fout = struct.unpack_from("<IQII", b, 4)
    assert ver == 1 and fin == nin and fout == nout, (ver, fin, fout)
    data = np.frombuffer(b, "<f8", offset=24).reshape(nrec, nin + nout)
    return data[:, :nin].copy(), data[:, nin:].copy()


def sym10(vals):
    """golden.zig:255-266 — 10-value upper triangle -> symmetric 4x4."""
    m = np.zeros((4, 4))
    n = 0
    for i in range(4):
        for j in range(i, 4):
            m[i, j] = m[j, i] = vals[n]
            n += 1
    return m


def geom_from_record(vals, x=(0.0, 0.0, 0.0, 0.0)):
    """golden.zig:268-282 geomFromRecord -> kz_geom row (47 doubles)."""
    row = np.zeros(47)
    g = sym10(vals[0:10])
    G = sym10(vals[10:20])
    gg = np.zeros((4, 5))
    GG = np.zeros((4, 5))
    gg[:, :4] = g
    GG[:, :4] = G
    row[0:20] = gg.ravel()
    row[20:40] = GG.ravel()
    if len(vals) >= 23:
        row[40], row[41], row[42] = vals[20], vals[21], vals[22]
    else:
        row[40] = 1.0
        row[41] = np.sqrt(-1.0 / G[0, 0])
        row[42] = 0.0
    row[43:47] = x
    return row


def geoms(inp, n=23):
    return np.stack([geom_from_record(r[:n]) for r in inp])


def read_kstp(rel):
    """golden.zig:62-160 — forced-dt step records."""
    b = _raw(GOLDEN / rel)
    assert b[:4] == b"KSTP"
    ver, nx, ny, nz, nv, nrec = struct.unpack_from("<IIIIIQ", b, 4)
    ncell = nx * ny * nz
    per = (len(b) - 32) // 8 // nrec
    nflags = (per - 2 - 2 * ncell * nv) // ncell
    data = np.frombuffer(b, "<f8", offset=32).reshape(nrec, per)
    recs = []
    for r in data:
        t, dt = r[0], r[1]
        u = r[2:2 + ncell * nv].reshape(nz, ny, nx, nv)
        p = r[2 + ncell * nv:2 + 2 * ncell * nv].reshape(nz, ny, nx, nv)
        f = r[2 + 2 * ncell * nv:].reshape(nz, ny, nx, nflags)
        recs.append(dict(t=t, dt=dt, u=u, p=p, flags=f))
    return dict(nx=nx, ny=ny, nz=nz, nv=nv, nflags=nflags, recs=recs)


def read_kini(rel):
    """golden.zig:290-369 — grid snapshot of selected primitive slots."""
    b = _raw(GOLDEN / rel)
    assert b[:4] == b"KINI"
    ver, nx, ny, nz, gx, gy, gz, stride, nvars = struct.unpack_from("<Iiiiiiiii", b, 4)
    vars_ = list(struct.unpack_from("<%di" % nvars, b, 40))
    off = 40 + 4 * nvars
    nsx = nx + 2 * gx if gx > 0 else (nx + stride - 1) // stride
    nsy = ny + 2 * gy if gy > 0 else (ny + stride - 1) // stride
    data = np.frombuffer(b, "<f8", offset=off).reshape(nz, nsy, nsx, nvars)
    cx = (np.arange(nsx) - gx) if gx > 0 else np.arange(nsx) * stride
    cy = (np.arange(nsy) - gy) if gy > 0 else np.arange(nsy) * stride
    return dict(nx=nx, ny=ny, nz=nz, gx=gx, gy=gy, stride=stride, vars=vars_, data=data, cx=cx, cy=cy)


def read_kslf(rel):
    """selfgolden.zig:20-42 — KSLF v1."""
    b = _raw(GOLDEN / rel)
    assert b[:4] == b"KSLF"
    ver, nx, ny, nz, ng, nv = struct.unpack_from("<IIIIII", b, 4)
    ncell, = struct.unpack_from("<Q", b, 28)
    nflags, nsc, nfr, llen = struct.unpack_from("<IIII", b, 36)
    off = 52
    label = b[off:off + llen].decode()
    off += llen
    off = (off + 7) // 8 * 8
    a = np.frombuffer(b, "<f8", offset=off)
    k = 0
    scal = a[k:k + nsc * 7].reshape(nsc, 7)
    k += nsc * 7
    steps = a[k:k + nfr]
    k += nfr
    flen = 2 * ncell * nv + ncell * nflags
    fields = []
    for i in range(nfr):
        f = a[k + i * flen:k + (i + 1) * flen]
        u = f[:ncell * nv].reshape(-1, nv)
        p = f[ncell * nv:2 * ncell * nv].reshape(-1, nv)
        fl = f[2 * ncell * nv:].reshape(-1, nflags)
        fields.append(dict(step=int(steps[i]), u=u, p=p, flags=fl))
    return dict(nx=nx, ny=ny, nz=nz, ng=ng, nv=nv, ncell=ncell, nflags=nflags, label=label,
                scalars=scal, fields=fields)


class DevTracker:
    """golden.zig:165-195 — worst |C - ours| / max(1, |C|, |ours|)."""

    def __init__(self):
        self.max_dev = 0.0
        self.at = -1

    def add(self, c, z, irec=-1):
        c = np.asarray(c, float)
        z = np.asarray(z, float)
        # the reference's tracker skips NaN deviations (a NaN `>` is false);
        # skip only where the golden value itself is NaN, so a NaN of ours
        # still fails.
        c, z = np.broadcast_arrays(c, z)
        keep = ~np.isnan(c)
        c, z = c[keep], z[keep]
        dev = np.abs(c - z) / np.maximum(1.0, np.maximum(np.abs(c), np.abs(z)))
        m = float(np.max(dev)) if dev.size else 0.0
        if m > self.max_dev or np.isnan(m):
            self.max_dev = m
            self.at = irec
        return self

    def ok(self, bound):
        return self.max_dev <= bound


def kdmp_bytes(p_dom, t, nstep, out_idx):
    """io/dump.zig:101-135 serializePrimDump: the 44-byte v2 header then the
    domain primitives p[nz][ny][nx][nv], little-endian."""
    nz, ny, nx, nv = p_dom.shape
    head = b"KDMP" + struct.pack("<IIIIIdQI", 2, nx, ny, nz, nv, t, nstep, out_idx)
    return head + np.ascontiguousarray(p_dom, "<f8").tobytes()


def read_kdmp(b):
    """io/dump.zig:82-98 parseDumpHeader + the body."""
    assert b[:4] == b"KDMP"
    ver, nx, ny, nz, nv, t, nstep, out_idx = struct.unpack_from("<IIIIIdQI", b, 4)
    assert ver == 2
    p = np.frombuffer(b, "<f8", offset=44, count=nx * ny * nz * nv).reshape(nz, ny, nx, nv)
    return dict(nx=nx, ny=ny, nz=nz, nv=nv, t=t, nstep=nstep, out_idx=out_idx, p=p)
