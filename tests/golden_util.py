"""Loading of the golden vectors written by tests/golden/make_golden.py."""
import glob
import os

import numpy as np

from ndspmhd_amd.options import Options, IOPT, DOPT

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def names():
    return sorted(os.path.basename(p)[:-4] for p in glob.glob(os.path.join(GOLDEN, "*.npz")))


def opts_from(iopt, dopt) -> Options:
    o = Options()
    kw = {}
    for k in ("ikernel", "iener", "imhd", "imagforce", "idivbzero", "iav", "idust",
              "idrag_nature", "maxdensits", "idustevol", "ndim"):
        kw[k] = int(iopt[IOPT[k]])
    kw["iavlim"] = tuple(int(iopt[IOPT["iavlim%d" % d]]) for d in range(3))
    kw["ibound"] = tuple(int(iopt[IOPT["ibound%d" % d]]) for d in range(3))
    for k in ("hfact", "tolh", "gamma", "polyk", "Kdrag", "alphamin", "alphaumin",
              "alphaBmin", "beta", "avdecayconst", "psidecayfact", "C_cour", "C_force",
              "psep", "dt"):
        kw[k] = float(dopt[DOPT[k]])
    kw["xmin"] = tuple(float(dopt[DOPT["xmin%d" % d]]) for d in range(3))
    kw["xmax"] = tuple(float(dopt[DOPT["xmax%d" % d]]) for d in range(3))
    kw["Bconst"] = tuple(float(dopt[DOPT["Bconst%d" % d]]) for d in range(3))
    return o.replace(**kw)


def load(name):
    z = np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False)
    o = opts_from(z["in_iopt"], z["in_dopt"])
    st = {k[3:]: z[k] for k in z.files if k.startswith("in_") and k not in ("in_iopt", "in_dopt")}
    out = {k[4:]: z[k] for k in z.files if k.startswith("out_")}
    return o, st, out


def relerr(a, b):
    """max |a-b| / max |b| (norm-relative L-infinity)"""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64).reshape(a.shape)
    s = np.max(np.abs(b)) if b.size else 0.0
    d = np.max(np.abs(a - b)) if b.size else 0.0
    return d / s if s > 0 else d


def pairset(p):
    return set(map(tuple, np.asarray(p).reshape(-1, 3).tolist()))
