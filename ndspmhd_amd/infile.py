"""Input files (.in) of the reference: reader, writer and the mapping onto
this library's options.

Format (src/utils_infiles.f90, src/readwrite_infile.f90:35-412): one option
per line, `name = value ! comment`; blank lines and lines starting with '#'
are skipped; a line holds at most 120 characters; the name is what lies left
of the first '=', the value what lies between it and the first '!' (both at
most 20 characters after trimming); a value with ':' is a time hh:mm:ss
(stored as seconds, rounded as es11.4 rounds).  Values are read as Fortran
list-directed input (reals may use a d exponent; logicals T/F/.true./.false.).
Options missing from the file keep their defaults (src/defaults.f90:29-120);
the reference then rewrites the file with all options and stops, which
`read_infile(..., strict=True)` mirrors by raising.

The writer produces the reference's layout (write_inopt formats,
src/utils_infiles.f90:83-233): `(a20,' = ',1x,i10,4x,'! ',a)` for integers,
l10 for logicals, and for reals es10.3 outside [1e-3, 1e4), f16.13 below
0.1 and g16.9 otherwise, trailing zeros stripped (keeping three after the
point, none from 10 up), right-justified in 10 characters.
"""
from __future__ import annotations

import math
import os
from dataclasses import dataclass, field

from .options import Options

# write_infile's groups, names, kinds (i integer, r real, l logical) and
# descriptions, in file order (src/readwrite_infile.f90:57-147)
GROUPS = [
    ("options affecting setup", [
        ("psep", "r", "particle separation")]),
    ("options affecting timestepping/output", [
        ("tmax", "r", "maximum time"),
        ("tout", "r", "time between outputs"),
        ("nmax", "i", "maximum number of timesteps"),
        ("nout", "i", "number of timesteps between outputs (-1=ignore)"),
        ("C_cour", "r", "Courant factor on timestep"),
        ("C_force", "r", "Factor in force timestep condition"),
        ("idumpghost", "i", "dump ghost particles? (0: no 1: yes)"),
        ("ind_timesteps", "i", "individual timesteps (0: no 1: yes)")]),
    ("options affecting density calculation", [
        ("icty", "i", "type of cty equation (0:direct sum 1:time deriv)"),
        ("ndirect", "i", "perform direct sum every n timesteps (if icty=1)"),
        ("maxdensits", "i", "maximum number of density iterations"),
        ("ihvar", "i", "type of variable smoothing length prediction"),
        ("hfact", "r", "h in units of mean particle spacing"),
        ("tolh", "r", "tolerance on h-rho iterations"),
        ("ikernel", "i", "kernel type (0: cubic spline, 3:quintic)"),
        ("ikernav", "i", "type of kernel averaging (1:average h, 2:average grad wab 3:springel/hernquist)"),
        ("usenumdens", "l", "Use number density formulation of gradh")]),
    ("options affecting equation of state, energy equation", [
        ("iener", "i", "type of energy equation"),
        ("gamma", "r", "adiabatic index"),
        ("polyk", "r", "polytropic constant if iener=0")]),
    ("options affecting momentum equation and shock-capturing terms", [
        ("iexternal_force", "i", "external force (1: toy star, 2:1/r^2 )"),
        ("iav", "i", "type of artificial viscosity"),
        ("alphamin", "r", "minimum alpha (viscosity)"),
        ("alphaumin", "r", "minimum alphau (conductivity)"),
        ("alphaBmin", "r", "minimum alphaB (resistivity)"),
        ("beta", "r", "beta in artificial viscosity"),
        ("iavlim", "i", "use viscosity switch"),
        ("iavlimu", "i", "use conductivity switch"),
        ("iavlimB", "i", "use resistivity switch"),
        ("avdecayconst", "r", "decay constant in av switch (0.1-0.2)"),
        ("damp", "r", "artificial damping (0.0 or few percent)"),
        ("dampz", "r", "artificial damping in z (0.0 or few percent)"),
        ("dampr", "r", "artificial damping in r (0.0 or few percent)"),
        ("ixsph", "i", "use XSPH (0:off 1:on)"),
        ("xsphfac", "r", "factor for XSPH correction"),
        ("iprterm", "i", "type of pressure term (0:normal 1:pa+pb/rhoa*rhob 2:hernquist/katz"),
        ("ismooth", "i", "smooth initial conditions?")]),
    ("options affecting magnetic fields", [
        ("imhd", "i", "MHD (0:no 1-10:B/rho >10:B <0:A -3:GEPs)"),
        ("imagforce", "i", "MHD force type(1:vector 2:tensor)"),
        ("idivbzero", "i", "divergence correction method (0:none 1:projection 2: hyperbolic/parabolic)"),
        ("psidecayfact", "r", "decay factor in hyperbolic/parabolic cleaning"),
        ("nsteps_remap", "i", "remapping interval (0:never)"),
        ("iuse_exact_derivs", "i", "use exact derivatives for MHD (0:off 1:on)"),
        ("iresist", "i", "resistivity (0:off 1:explicit 2:implicit 3:variable eta 4:two first derivs)"),
        ("etamhd", "r", "eta for resistivity"),
        ("iambipolar", "i", "ambipolar diffusion")]),
    ("options affecting dust", [
        ("idust", "i", "dust (0:off 1:one-f 2:two-f 3:diff-onef-1st 4:diff-onef-2ndderivs)"),
        ("idrag_nature", "i", "drag type (0=none 1=const K 2=const ts 3=Epstein)"),
        ("Kdrag", "r", "drag coeff (idrag=1) or ts (idrag=2) or grain size in cm (idrag=3)"),
        ("idustevol", "i", "dust variable [0=dustfrac(eps) 1=sqrt(rho*eps) 2=asin(eps)]"),
        ("use_smoothed_rhodust", "l", "use smoothed dust/gas density estimates?"),
        ("islope_limiter", "i",
         "reconstruction/slope limiter (-1=none,0=recon w/out limiter,1=van leer,2=van leer mc 3=minmod")]),
    ("options affecting self-gravity", [
        ("igravity", "i", "self-gravity"),
        ("hsoft", "r", "fixed softening length")]),
    ("options affecting physical viscosity", [
        ("ivisc", "i", "use physical viscosity?"),
        ("shearvisc", "r", "shear viscosity (nu)"),
        ("bulkvisc", "r", "bulk viscosity (zeta)"),
        ("icompute_d2v", "i", "compute velocity derivatives (0:none 1:brookshaw 2:direct)"),
        ("idiffuse", "i", "methods for testing diffusion terms (1=2nd derivs 2=two first derivs)"),
        ("k_iso", "r", "isotropic diffusion coefficient"),
        ("k_par", "r", "anisotropic diffusion coefficient")]),
    ("options affecting particle splitting", [
        ("isplitpart", "i", "particle splitting"),
        ("rhocrit", "r", "critical density for particle splitting")]),
    ("options affecting Quantum SPH", [
        ("iquantum", "i", "quantum SPH term")]),
]
KIND = {name: kind for _, items in GROUPS for name, kind, _ in items}

# set_default_options (src/defaults.f90:45-120); options it leaves unset are
# zero (static module storage)
DEFAULTS = dict(
    psep=0.01, tmax=1.0, tout=0.1, nmax=1000000, nout=-1, C_cour=0.3, C_force=0.25,
    idumpghost=1, ind_timesteps=0, icty=0, ndirect=1000000, maxdensits=250, ihvar=2,
    hfact=1.2, tolh=1.0e-3, ikernel=0, ikernav=3, usenumdens=False, iener=2,
    gamma=5.0 / 3.0, polyk=1.0, iexternal_force=0, iav=2, alphamin=0.1, alphaumin=0.0,
    alphaBmin=1.0, beta=2.0, iavlim=2, iavlimu=1, iavlimB=0, avdecayconst=0.1, damp=0.0,
    dampz=0.0, dampr=0.0, ixsph=0, xsphfac=0.0, iprterm=0, ismooth=0, imhd=0,
    imagforce=2, idivbzero=0, psidecayfact=0.1, nsteps_remap=0, iuse_exact_derivs=0,
    iresist=0, etamhd=0.0, iambipolar=0, idust=0, idrag_nature=0, Kdrag=0.0,
    idustevol=0, use_smoothed_rhodust=True, islope_limiter=-1, igravity=0, hsoft=0.0,
    ivisc=0, shearvisc=0.0, bulkvisc=0.0, icompute_d2v=0, idiffuse=0, k_iso=0.0,
    k_par=0.0, isplitpart=0, rhocrit=1.0e4, iquantum=0)


class InfileError(ValueError):
    pass


# ---------------------------------------------------------------- reading

def _parse_line(line: str):
    """(name, value string) of one option line (read_inopt_from_line,
    src/utils_infiles.f90:628-690), or None for a line without '='"""
    line = line[:120]
    ieq = line.find("=")
    if ieq < 0:
        return None
    name = line[:ieq].strip()[:20]
    icom = line.find("!")
    if icom < 0:
        icom = len(line.rstrip()) + 1
    val = line[ieq + 1:icom - 1 + 1].strip() if icom > ieq else ""
    if ":" in val:
        val = _time_string(val)
    return name, val[:20]


def _time_string(val: str) -> str:
    """hh:mm:ss (or hhh:mm:ss) as seconds, written as es11.4 writes them"""
    parts = val.split(":")
    try:
        nums = [int(p) for p in parts] + [0, 0]
    except ValueError:
        return val
    sec = (nums[0] * 60.0 + nums[1]) * 60.0 + nums[2]
    return _es(sec, 11, 4)


def _read_int(s: str) -> int:
    tok = s.replace(",", " ").split()
    if not tok:
        raise ValueError("empty")
    return int(tok[0])


def _read_real(s: str) -> float:
    tok = s.replace(",", " ").split()
    if not tok:
        raise ValueError("empty")
    t = tok[0].replace("d", "e").replace("D", "E")
    if t.endswith((".",)):
        t += "0"
    return float(t)


def _read_logical(s: str) -> bool:
    tok = s.replace(",", " ").split()
    if not tok:
        raise ValueError("empty")
    t = tok[0].lstrip(".").upper()
    if t[:1] == "T":
        return True
    if t[:1] == "F":
        return False
    raise ValueError(s)


_READ = {"i": _read_int, "r": _read_real, "l": _read_logical}


def _required(vals: dict):
    """the names read_infile counts as errors when absent: all but the
    switches iavlim*, and the MHD, gravity and dust options only when those
    are on (src/readwrite_infile.f90:208-302)"""
    for _, items in GROUPS:
        for name, _, _ in items:
            if name.startswith("iavlim"):
                continue
            if name in ("imagforce", "idivbzero", "psidecayfact", "iresist", "etamhd",
                        "iuse_exact_derivs", "nsteps_remap", "iambipolar") and vals["imhd"] == 0:
                continue
            if name == "hsoft" and vals["igravity"] == 0:
                continue
            if name in ("idrag_nature", "Kdrag", "idustevol", "use_smoothed_rhodust") and vals["idust"] == 0:
                continue
            yield name


@dataclass
class Infile:
    values: dict
    missing: list = field(default_factory=list)
    warnings: list = field(default_factory=list)

    def __getitem__(self, k):
        return self.values[k]


def parse_infile(text: str, strict: bool = False) -> Infile:
    db = {}
    for raw in text.splitlines():
        line = raw[:120]
        if not line.strip() or line[:1] == "#":
            continue
        p = _parse_line(line)
        if p is None:
            continue
        db.setdefault(p[0], p[1])
    vals = dict(DEFAULTS)
    bad = []
    # imhd, igravity and idust first: they decide which options are read
    for name in ["imhd", "igravity", "idust"] + [n for n in KIND if n not in ("imhd", "igravity", "idust")]:
        if name in db:
            try:
                vals[name] = _READ[KIND[name]](db[name])
            except ValueError:
                bad.append(name)
    missing = [n for n in _required(vals) if n not in db] + bad
    inf = Infile(vals, missing)
    if strict and missing:
        raise InfileError("missing or unreadable options in input file: " + ", ".join(missing))
    inf.warnings = check(vals)
    return inf


def read_infile(path: str, strict: bool = False) -> Infile:
    if not os.path.exists(path):
        raise InfileError("input file %s not found" % path)
    with open(path) as f:
        return parse_infile(f.read(), strict)


def check(v: dict) -> list:
    """the reference's option checks (src/readwrite_infile.f90:328-395):
    returns the warnings, raises InfileError where the reference stops"""
    w = []
    if v["psep"] < 1.0e-5:
        w.append("psep < 1.e-5")
    if v["tout"] > v["tmax"]:
        w.append("no output tout > tmax")
    if v["nout"] > v["nmax"]:
        w.append("no output nout > nmax")
    if v["nout"] == 0:
        raise InfileError("error in input: nout = 0")
    if v["gamma"] < 1.0:
        w.append("gamma < 1.0 ")
    if abs(v["gamma"] - 1.0) < 1.0e-3 and v["iener"] != 0:
        raise InfileError("must use iener = 0 for isothermal eos")
    if (v["iener"] > 0 and v["alphaumin"] < 0.0) or v["alphaBmin"] < 0.0:
        w.append("alphaumin or alphabmin < 0.")
    elif v["iener"] == 0 and v["polyk"] < 0.0:
        w.append("polyk < 0.")
    if v["iav"] != 0 and (v["alphamin"] < 0.0 or v["beta"] < 0.0):
        w.append("av alpha or beta < 0.")
    if v["iavlim"] > 0 and v["alphamin"] >= 1.0:
        w.append("using av limiter, but alphamin set > 1.0")
    if any(v[k] > 0 for k in ("iavlim", "iavlimu", "iavlimB")) and \
            (v["avdecayconst"] <= 0.01 or v["avdecayconst"] > 0.5):
        w.append("av decay constant not in range 0.01-0.5")
    if v["ikernav"] <= 0 or v["ikernav"] > 3:
        w.append("kernel averaging not set (ikernav)")
    if v["hfact"] <= 1.0 or v["hfact"] > 2.0:
        w.append("hfact too low/high (1.0 < hfact < 2.0)")
    if v["psidecayfact"] < 0.0:
        w.append("psidecayfact < 0.0")
    if v["tolh"] < 1.0e-12:
        raise InfileError("tolh really, really tiny (probably zero)!!")
    if v["iresist"] < 0 or v["iresist"] > 4:
        raise InfileError("invalid choice of resistivity formulation")
    if v["etamhd"] < 0.0:
        raise InfileError("eta < 0 in resistivity")
    if v["shearvisc"] < 0.0:
        raise InfileError("invalid choice of shear viscosity parameter")
    if v["bulkvisc"] < 0.0:
        raise InfileError("invalid choice of bulk viscosity parameter")
    if v["idust"] < 0 or v["idust"] > 4:
        raise InfileError("invalid choice of dust formulation")
    if v["iambipolar"] < 0 or v["iambipolar"] > 1:
        raise InfileError("invalid choice of ambipolar diffusion formulation")
    if v["isplitpart"] < 0:
        raise InfileError("invalid choice for isplitpart")
    if v["isplitpart"] >= 0 and v["rhocrit"] <= 0.0:
        w.append("critical density <= 0 in particle splitting")
    if v["nsteps_remap"] < 0:
        raise InfileError("nsteps_remap < 0 invalid remapping interval")
    return w


# ---------------------------------------------------------------- writing

def _es(x: float, w: int, d: int) -> str:
    """x under esW.D"""
    mant, ex = ("%.*E" % (d, x)).split("E")
    e = int(ex)
    s = mant + ("E%+03d" % e if abs(e) <= 99 else "%+04d" % e)
    return s.rjust(w)


def _g(x: float, w: int, d: int) -> str:
    """x under gW.D for 0.1 <= |x| < 10**d (F editing with d-k decimals,
    then four blanks)"""
    if x == 0.0:
        return ("%.*f" % (d - 1, 0.0)).rjust(w - 4) + "    "
    e = int(("%.*E" % (d - 1, abs(x))).split("E")[1])
    k = e + 1
    if not (0 <= k <= d):
        return _es(x, w, d)
    return ("%.*f" % (d - k, x)).rjust(w - 4) + "    "


def _strip_zeros(s: str, n: int) -> str:
    """blank the trailing zeros after the last '.', keeping n of them"""
    idot = s.rfind(".")
    if idot < 0:
        return s
    chars = list(s)
    end = len(s.rstrip())
    for i in range(end - 1, idot + n, -1):
        if chars[i] == "0":
            chars[i] = " "
        else:
            break
    return "".join(chars)


def _real_field(x: float):
    """(value text, spaces before '!') of a real as write_inopt_real8 writes it"""
    ax = abs(x)
    if (ax < 1.0e-3 and ax > 2.2250738585072014e-308) or ax >= 1.0e4:
        return _es(x, 10, 3), 4
    if ax < 1.0e-1:
        t = _strip_zeros("%16.13f" % x, 3).strip()
    elif ax >= 1.0e1:
        t = _strip_zeros(_g(x, 16, 9), 0).strip()
    else:
        t = _strip_zeros(_g(x, 16, 9), 3).strip()
    if len(t) > 10:
        return t, 2
    return t.rjust(10), 4


def format_option(name: str, value, descript: str) -> str:
    kind = KIND[name]
    if kind == "i":
        val, gap = "%10d" % int(value), 4
    elif kind == "l":
        val, gap = ("T" if value else "F").rjust(10), 4
    else:
        val, gap = _real_field(float(value))
    return "%20s = %s%s! %s" % (name, " " + val, " " * gap, descript)


def format_infile(values: dict) -> str:
    v = dict(DEFAULTS)
    v.update(values)
    lines = ["# input file for ndspmhd, options not present assume default values"]
    for title, items in GROUPS:
        lines += ["", "# " + title]
        for name, _, descript in items:
            lines.append(format_option(name, v[name], descript))
    return "\n".join(lines) + "\n"


def write_infile(path: str, values: dict):
    with open(path, "w") as f:
        f.write(format_infile(values))


# ---------------------------------------------------------------- mapping

# options whose reference branches this library does not implement, with
# the only value it accepts
UNSUPPORTED = dict(ind_timesteps=0, icty=0, ihvar=2, ikernav=3, usenumdens=False,
                   iexternal_force=0, damp=0.0, dampz=0.0, dampr=0.0, ixsph=0, iprterm=0,
                   iuse_exact_derivs=0, iresist=0, iambipolar=0, igravity=0, ivisc=0,
                   icompute_d2v=0, idiffuse=0, isplitpart=0, iquantum=0)


def to_options(inf, ndim: int, ibound=(2, 0, 0), xmin=(0.0, 0.0, 0.0), xmax=(1.0, 1.0, 1.0),
               dt: float = 0.0) -> Options:
    """this library's Options from an input file's values and the set-up's
    dimension and box (the reference's set-up routine supplies those)"""
    v = inf.values if isinstance(inf, Infile) else inf
    for k, ok in UNSUPPORTED.items():
        if v[k] != ok:
            raise InfileError("option %s = %r is not supported on this path (only %r)" % (k, v[k], ok))
    return Options(
        ndim=ndim, ikernel=v["ikernel"], iener=v["iener"], imhd=v["imhd"], imagforce=v["imagforce"],
        idivbzero=v["idivbzero"], iav=v["iav"], iavlim=(v["iavlim"], v["iavlimu"], v["iavlimB"]),
        idust=v["idust"], idrag_nature=v["idrag_nature"], idustevol=v["idustevol"],
        ibound=tuple(ibound), maxdensits=v["maxdensits"], hfact=v["hfact"], tolh=v["tolh"],
        gamma=v["gamma"], polyk=v["polyk"], Kdrag=v["Kdrag"], alphamin=v["alphamin"],
        alphaumin=v["alphaumin"], alphaBmin=v["alphaBmin"], beta=v["beta"],
        avdecayconst=v["avdecayconst"], psidecayfact=v["psidecayfact"], C_cour=v["C_cour"],
        C_force=v["C_force"], xmin=tuple(xmin), xmax=tuple(xmax), psep=v["psep"], dt=dt)


def values_from_options(o: Options, **run) -> dict:
    """input-file values for Options (and run controls such as tmax, tout,
    nmax, nout)"""
    v = dict(DEFAULTS)
    for k in ("ikernel", "iener", "imhd", "imagforce", "idivbzero", "iav", "idust",
              "idrag_nature", "idustevol", "maxdensits", "hfact", "tolh", "gamma", "polyk",
              "Kdrag", "alphamin", "alphaumin", "alphaBmin", "beta", "avdecayconst",
              "psidecayfact", "C_cour", "C_force", "psep"):
        v[k] = getattr(o, k)
    v["iavlim"], v["iavlimu"], v["iavlimB"] = o.iavlim
    v.update(run)
    return v
