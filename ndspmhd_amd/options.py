"""Run options for the hot path, mirroring the reference's `options`/`artvi`/
`eos`/`timestep` module globals (src/variablesND.f90:135-155,33-37,255-273;
src/eos.f90:34-36) for the branches this library implements.

Defaults are the reference's (src/defaults.f90:29-120).  The two packed
vectors `iopt` (int32[32]) and `dopt` (float64[32]) are the layout shared by
the C-ABI (`ndsp_opts` in include/ndsp.h), the CPU oracle and the reference
harness input; their slots are listed in IOPT/DOPT below.
"""
from __future__ import annotations

import dataclasses
from dataclasses import dataclass, field

import numpy as np

# slot numbers (0-based) of the packed option vectors
IOPT = dict(ikernel=0, iener=1, imhd=2, imagforce=3, idivbzero=4, iav=5,
            iavlim0=6, iavlim1=7, iavlim2=8, idust=9, idrag_nature=10,
            ibound0=11, ibound1=12, ibound2=13, maxdensits=14, idustevol=15,
            ndim=16, cellscan=17)
DOPT = dict(hfact=0, tolh=1, gamma=2, polyk=3, Kdrag=4, alphamin=5,
            alphaumin=6, alphaBmin=7, beta=8, avdecayconst=9,
            psidecayfact=10, C_cour=11, C_force=12, xmin0=13, xmin1=14,
            xmin2=15, xmax0=16, xmax1=17, xmax2=18, psep=19, Bconst0=20,
            Bconst1=21, Bconst2=22, dt=23, damp=24)

# boundary codes (src/ghostND_mhd.f90:180-200, src/boundaryND.f90:66-95)
IBOUND_NONE, IBOUND_FIXED, IBOUND_REFLECT, IBOUND_PERIODIC = 0, 1, 2, 3


@dataclass
class Options:
    ndim: int = 3
    ikernel: int = 0            # 0 cubic M4, 3 quintic M6 (src/kernelND.f90)
    iener: int = 2              # 0 polytropic, 2 thermal energy
    imhd: int = 0               # 0 off, 11 evolve B
    imagforce: int = 2          # 2 tensor form
    idivbzero: int = 0          # 2 hyperbolic psi cleaning
    iav: int = 2
    iavlim: tuple = (2, 1, 0)
    idust: int = 0              # 4 one-fluid dust, in-loop diffusion
    idrag_nature: int = 0       # 2 constant stopping time
    idustevol: int = 0
    ibound: tuple = (3, 3, 3)
    maxdensits: int = 250
    hfact: float = 1.2
    tolh: float = 1.0e-3
    gamma: float = 5.0 / 3.0
    polyk: float = 1.0
    Kdrag: float = 0.0
    alphamin: float = 0.1
    alphaumin: float = 0.0
    alphaBmin: float = 1.0
    beta: float = 2.0
    avdecayconst: float = 0.1
    psidecayfact: float = 0.1
    C_cour: float = 0.3
    C_force: float = 0.25
    xmin: tuple = (0.0, 0.0, 0.0)
    xmax: tuple = (1.0, 1.0, 1.0)
    psep: float = 0.01
    Bconst: tuple = (0.0, 0.0, 0.0)
    dt: float = 0.0
    # device-only: 1 = scan the cell grid in every pass instead of reusing
    # per-particle neighbour lists (results are identical; lists are faster)
    cellscan: int = 0

    def replace(self, **kw) -> "Options":
        return dataclasses.replace(self, **kw)

    @property
    def radkern(self) -> float:
        return 3.0 if self.ikernel == 3 else 2.0

    def iopt(self) -> np.ndarray:
        a = np.zeros(32, dtype=np.int32)
        for k in ("ikernel", "iener", "imhd", "imagforce", "idivbzero", "iav",
                  "idust", "idrag_nature", "maxdensits", "idustevol", "ndim",
                  "cellscan"):
            a[IOPT[k]] = getattr(self, k)
        for d in range(3):
            a[IOPT["iavlim%d" % d]] = self.iavlim[d]
            a[IOPT["ibound%d" % d]] = self.ibound[d] if d < self.ndim else 0
        return a

    def dopt(self) -> np.ndarray:
        a = np.zeros(32, dtype=np.float64)
        for k in ("hfact", "tolh", "gamma", "polyk", "Kdrag", "alphamin",
                  "alphaumin", "alphaBmin", "beta", "avdecayconst",
                  "psidecayfact", "C_cour", "C_force", "psep", "dt"):
            a[DOPT[k]] = getattr(self, k)
        for d in range(3):
            a[DOPT["xmin%d" % d]] = self.xmin[d]
            a[DOPT["xmax%d" % d]] = self.xmax[d]
            a[DOPT["Bconst%d" % d]] = self.Bconst[d]
        return a


# per-particle state fields, in the reference's column-major layout
# (x(ndim,N) is a C array of shape (N, ndim)); itype is int32
STATE_FIELDS = ("x", "vel", "pmass", "hh", "en", "Bevol", "alpha", "psi",
                "dustevol", "itype")


def empty_state(n: int, ndim: int) -> dict:
    return dict(
        x=np.zeros((n, ndim)), vel=np.zeros((n, 3)), pmass=np.zeros(n),
        hh=np.zeros(n), en=np.zeros(n), Bevol=np.zeros((n, 3)),
        alpha=np.zeros((n, 3)), psi=np.zeros(n), dustevol=np.zeros(n),
        itype=np.zeros(n, dtype=np.int32))


def check_state(st: dict, ndim: int) -> int:
    n = st["pmass"].shape[0]
    shapes = dict(x=(n, ndim), vel=(n, 3), pmass=(n,), hh=(n,), en=(n,),
                  Bevol=(n, 3), alpha=(n, 3), psi=(n,), dustevol=(n,),
                  itype=(n,))
    for k, shp in shapes.items():
        a = st[k]
        if a.shape != shp:
            raise ValueError("state field %s has shape %s, expected %s"
                             % (k, a.shape, shp))
    return n
