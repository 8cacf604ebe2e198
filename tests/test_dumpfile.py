"""Dump files (ndspmhd_amd.dumpfile): the record layout of the reference's
write_dump (src/readwrite_dumps.f90:32-200), checked on the CPU: column
counts against the reference's ncolumns formula, record framing, and a
write -> read round trip of every column."""
import numpy as np
import pytest

from ndspmhd_amd import dumpfile as df
from ndspmhd_amd.options import Options


def _ncolumns_reference(ndim, imhd, idust):
    """ncolumns as write_dump counts it (src/readwrite_dumps.f90:68-87) with
    the reference's allocations on this path: del2v always allocated
    (src/allocateND.f90:402-404), no gravity, cartesian, one dust species"""
    ndimV = 3
    n = ndim + 2 * ndimV + 4
    if imhd != 0:
        n += 8 + 2 * ndimV
    else:
        n += 5 + 2 * ndimV
    if idust in (1, 3, 4):
        n += ndimV + 2 * 1 + 1
    return n


@pytest.mark.parametrize("ndim", [1, 2, 3])
@pytest.mark.parametrize("imhd,idust", [(11, 0), (0, 0), (0, 4)])
def test_column_counts(ndim, imhd, idust):
    n, iformat, names = df.layout(ndim, imhd, idust)
    assert n == _ncolumns_reference(ndim, imhd, idust) == len(names)
    assert iformat == (5 if idust else 2 if imhd else 1)


def _fields(n, ndim, seed):
    rng = np.random.default_rng(seed)
    f = {k: rng.standard_normal(n) for k in ("hh", "dens", "rho", "uu", "pmass", "psi", "pr",
                                             "drhodt", "divB", "gradh", "dustfrac", "rhogas",
                                             "rhodust")}
    f["rho"] = np.abs(f["rho"]) + 0.5
    for k in ("vel", "alpha", "Bfield", "curlB", "force", "graddivv", "deltav"):
        f[k] = rng.standard_normal((n, 3))
    f["x"] = rng.standard_normal((n, ndim))
    f["itype"] = rng.integers(0, 3, n).astype(np.int32)
    return f


@pytest.mark.parametrize("ndim,imhd,idust", [(1, 11, 0), (2, 0, 0), (3, 0, 4), (3, 11, 0)])
def test_round_trip(tmp_path, ndim, imhd, idust):
    o = Options(ndim=ndim, imhd=imhd, idust=idust, ibound=(3, 2, 0)[:ndim] + (0,) * (3 - ndim),
                xmin=(-1.0, -2.0, -3.0), xmax=(1.0, 2.0, 3.0), gamma=1.4, hfact=1.2)
    f = _fields(37, ndim, ndim + imhd + idust)
    p = str(tmp_path / "d.dat")
    df.write_dump(p, 0.25, o, f)
    recs = df.read_records(p)
    ncol, iformat, names = df.layout(ndim, imhd, idust)
    assert len(recs) == ncol + 2
    assert all(len(r) == 37 * 8 for r in recs[1:-1]) and len(recs[-1]) == 37 * 4
    d = df.read_dump(p)
    assert d["t"] == 0.25 and d["npart"] == d["nprint"] == 37 and d["ndim"] == ndim
    assert d["iformat"] == iformat and d["geom"] == "cartesian   "
    assert d["ibound"] == tuple(o.ibound[:ndim]) and d["xmax"] == tuple(o.xmax[:ndim])
    st = d["state"]
    np.testing.assert_array_equal(st["x"], f["x"])
    np.testing.assert_array_equal(st["vel"], f["vel"])
    np.testing.assert_array_equal(st["en"], f["uu"])
    np.testing.assert_array_equal(st["itype"], f["itype"])
    np.testing.assert_array_equal(d["columns"]["-drhodt/rho"], -f["drhodt"] / f["rho"])
    if imhd:
        np.testing.assert_array_equal(st["Bevol"], f["Bfield"])
        np.testing.assert_array_equal(st["alpha"], f["alpha"])
    else:
        np.testing.assert_array_equal(st["alpha"][:, :2], f["alpha"][:, :2])
        np.testing.assert_array_equal(d["columns"]["del2v2"], 0.0)
    if idust:
        np.testing.assert_array_equal(st["dustevol"], f["dustfrac"])
        np.testing.assert_array_equal(d["columns"]["rho"], f["rho"])


def test_bad_record_marker_detected(tmp_path):
    o = Options(ndim=1, imhd=0)
    p = str(tmp_path / "d.dat")
    df.write_dump(p, 0.0, o, _fields(5, 1, 0))
    raw = bytearray(open(p, "rb").read())
    raw[-1] ^= 1
    open(p, "wb").write(bytes(raw))
    with pytest.raises(ValueError):
        df.read_records(p)
