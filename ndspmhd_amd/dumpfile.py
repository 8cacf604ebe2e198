"""Dump files in the reference's layout (src/readwrite_dumps.f90:32-200):
Fortran unformatted sequential records, each framed by its length in bytes
as a little-endian int32 before and after.  The first record is the header

    t (f8), npart, nprint (i4), gamma, hfact (f8), ndim, ndimV, ncolumns,
    iformat (i4), ibound(ndim) (i4), xmin(ndim), xmax(ndim) (f8),
    len(geom) (i4), geom (12 characters)

then one record per column of nprint values:

    x(1..ndim) vel(1..3) hh dens|rho uu pmass
    MHD (iformat 2):    alpha(1..3) Bfield(1..3) psi pr -drhodt/rho divB
                        curlB(1..3) gradh force(1..3)
    otherwise:          alpha(1..2) pr -drhodt/rho gradh force(1..3)
                        del2v(1..3) graddivv(1..3)
    one-fluid dust (iformat 5) appends dustfrac deltav(1..3) rhogas rhodust
    itype (i4) last, not counted in ncolumns.

The real particles are written (nprint = npart, the reference's idumpghost 0
layout); del2v is zero (no physical viscosity).  The Fortran host
(ndspmhd_amd/fortran, module ndsp_dumpfiles) writes the same bytes.
"""
from __future__ import annotations

import struct

import numpy as np

NDIMV = 3
GEOM = "cartesian".ljust(12)


def layout(ndim: int, imhd: int, idust: int) -> tuple:
    """(ncolumns, iformat, column names) for these options"""
    cols = ["x%d" % (d + 1) for d in range(ndim)] + ["v%d" % (d + 1) for d in range(NDIMV)]
    dusty = idust in (1, 3, 4)
    cols += ["hh", "rho" if dusty else "dens", "uu", "pmass"]
    if imhd != 0:
        iformat = 2
        cols += ["alpha%d" % d for d in (1, 2, 3)] + ["B%d" % d for d in (1, 2, 3)]
        cols += ["psi", "pr", "-drhodt/rho", "divB"] + ["curlB%d" % d for d in (1, 2, 3)]
        cols += ["gradh"] + ["force%d" % d for d in (1, 2, 3)]
    else:
        iformat = 1
        cols += ["alpha1", "alpha2", "pr", "-drhodt/rho", "gradh"]
        cols += ["force%d" % d for d in (1, 2, 3)] + ["del2v%d" % d for d in (1, 2, 3)]
        cols += ["graddivv%d" % d for d in (1, 2, 3)]
    if dusty:
        iformat = 5
        cols += ["dustfrac"] + ["deltav%d" % d for d in (1, 2, 3)] + ["rhogas", "rhodust"]
    return len(cols), iformat, cols


def _record(payload: bytes) -> bytes:
    n = struct.pack("<i", len(payload))
    return n + payload + n


def columns_from(fields: dict, ndim: int, imhd: int, idust: int) -> list:
    """the column arrays of a dump from per-particle fields in the reference
    layout (x (n,ndim), vel (n,3), ...), in file order"""
    _, _, names = layout(ndim, imhd, idust)
    n = fields["pmass"].shape[0]
    out = []
    for c in names:
        if c.startswith("x"):
            v = fields["x"].reshape(n, ndim)[:, int(c[1:]) - 1]
        elif c[0] == "v" and c[1:].isdigit():
            v = fields["vel"][:, int(c[1:]) - 1]
        elif c.startswith("alpha"):
            v = fields["alpha"][:, int(c[5:]) - 1]
        elif c[0] == "B" and c[1:].isdigit():
            v = fields["Bfield"][:, int(c[1:]) - 1]
        elif c == "-drhodt/rho":
            v = -fields["drhodt"] / fields["rho"]
        elif c[:-1] in ("curlB", "force", "graddivv", "deltav"):
            v = fields[c[:-1]][:, int(c[-1]) - 1]
        elif c.startswith("del2v"):
            v = np.zeros(n)
        else:
            v = fields[c]
        out.append(np.ascontiguousarray(v, dtype="<f8"))
    return out


def write_dump(path: str, t: float, opts, fields: dict):
    """write a dump of per-particle fields (names as the reference's arrays)"""
    ndim = opts.ndim
    ncol, iformat, _ = layout(ndim, opts.imhd, opts.idust)
    n = fields["pmass"].shape[0]
    hdr = struct.pack("<diiddiiii", t, n, n, opts.gamma, opts.hfact, ndim, NDIMV, ncol, iformat)
    hdr += struct.pack("<%di" % ndim, *opts.ibound[:ndim])
    hdr += struct.pack("<%dd" % ndim, *opts.xmin[:ndim])
    hdr += struct.pack("<%dd" % ndim, *opts.xmax[:ndim])
    hdr += struct.pack("<i", 12) + GEOM.encode()
    with open(path, "wb") as f:
        f.write(_record(hdr))
        for v in columns_from(fields, ndim, opts.imhd, opts.idust):
            f.write(_record(v.tobytes()))
        f.write(_record(np.ascontiguousarray(fields["itype"], dtype="<i4").tobytes()))


DUMP_FIELDS = ("x", "vel", "hh", "dens", "rho", "uu", "pmass", "alpha", "Bfield", "psi", "pr",
               "drhodt", "divB", "curlB", "gradh", "force", "graddivv", "dustfrac", "deltav",
               "rhogas", "rhodust")


def write_sim(path: str, sim, t: float):
    """dump of a Simulation's current state (fields downloaded from HBM)"""
    fields = {k: sim.get(k) for k in DUMP_FIELDS}
    fields["itype"] = np.asarray(sim.itype, dtype=np.int32)
    write_dump(path, t, sim.opts, fields)


def read_records(path: str) -> list:
    """the raw records of a Fortran unformatted sequential file"""
    raw = open(path, "rb").read()
    recs, off = [], 0
    while off < len(raw):
        (n,) = struct.unpack_from("<i", raw, off)
        body = raw[off + 4: off + 4 + n]
        (m,) = struct.unpack_from("<i", raw, off + 4 + n)
        if m != n:
            raise ValueError("record markers disagree at byte %d" % off)
        recs.append(body)
        off += n + 8
    return recs


def read_dump(path: str) -> dict:
    """header and named columns of a dump; 'state' holds what a restart
    needs (x, vel, hh, en = uu, pmass, alpha, Bevol = Bfield, psi,
    dustevol = dustfrac, itype)"""
    recs = read_records(path)
    h = recs[0]
    t, npart, nprint, gamma, hfact, ndim, ndimv, ncol, iformat = struct.unpack_from("<diiddiiii", h, 0)
    off = struct.calcsize("<diiddiiii")
    ibound = struct.unpack_from("<%di" % ndim, h, off)
    off += 4 * ndim
    xmin = struct.unpack_from("<%dd" % ndim, h, off)
    off += 8 * ndim
    xmax = struct.unpack_from("<%dd" % ndim, h, off)
    off += 8 * ndim
    (lg,) = struct.unpack_from("<i", h, off)
    geom = h[off + 4: off + 4 + lg].decode()
    imhd = 1 if iformat in (2, 4) else 0
    idust = 4 if iformat == 5 else 0
    ncol2, _, names = layout(ndim, imhd, idust)
    if ncol2 != ncol or len(recs) != ncol + 2:
        raise ValueError("dump has %d columns (%d records), layout expects %d" % (ncol, len(recs), ncol2))
    cols = {c: np.frombuffer(r, dtype="<f8").copy() for c, r in zip(names, recs[1:-1])}
    itype = np.frombuffer(recs[-1], dtype="<i4").copy()
    n = nprint
    st = {
        "x": np.stack([cols["x%d" % (d + 1)] for d in range(ndim)], axis=1),
        "vel": np.stack([cols["v%d" % (d + 1)] for d in range(NDIMV)], axis=1),
        "hh": cols["hh"], "en": cols["uu"].copy(), "pmass": cols["pmass"],
        "alpha": np.zeros((n, 3)), "Bevol": np.zeros((n, 3)), "psi": np.zeros(n),
        "dustevol": cols["dustfrac"].copy() if idust else np.zeros(n), "itype": itype,
    }
    for d in range(3 if imhd else 2):
        st["alpha"][:, d] = cols["alpha%d" % (d + 1)]
    if imhd:
        for d in range(3):
            st["Bevol"][:, d] = cols["B%d" % (d + 1)]
        st["psi"] = cols["psi"]
    return {"t": t, "npart": npart, "nprint": nprint, "gamma": gamma, "hfact": hfact, "ndim": ndim,
            "ndimV": ndimv, "ncolumns": ncol, "iformat": iformat, "ibound": ibound, "xmin": xmin,
            "xmax": xmax, "geom": geom, "columns": cols, "state": st}
