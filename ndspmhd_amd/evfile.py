"""The .ev file (one line of global diagnostics per step) in the reference's
text layout, src/evwrite_mhd.f90:270-308.

Column sets (src/evwrite_mhd.f90:276-279, 296-297, 304-305):
MHD runs 25 columns, one-fluid dust runs 14, other runs 13.  The header is
written with the format ('#',24(a18,1x)), the MHD line with
(24(1pe18.10,1x),1pe8.2) and the others with (24(1pe18.10,1x)).  The
formatting below produces the text those formats produce: a18 right-justifies
each name in 18 characters; 1pe18.10 prints one digit before the point, ten
after and a two-digit exponent (three digits without the letter E beyond
+-99); 1x separates items; a record that runs out of items ends there (no
trailing blanks); and format reversion starts a new record when a header has
more names than the 24 the format holds (the MHD header's 25th name,
fracdivBok, goes on a line of its own, without '#').
"""
from __future__ import annotations

import math
from typing import Iterable, Sequence

COMMON = ("t", "ekin", "etherm", "emag", "epot", "etot", "momtot", "angtot",
          "rhomax", "rhomean", "dt")
MHD = COMMON + ("emagp", "crosshel", "betamhdmin", "betamhdav", "divBav", "divBmax",
                "divBtot", "fdotBav", "fdotBmax", "force_err_av", "force_err_max",
                "omegamhdav", "omegamhdmax", "fracdivBok")
DUST = COMMON + ("ekiny", "totmassgas", "totmassdust")
HYDRO = COMMON + ("ekiny", "dmomtot")


def columns_for(imhd: int, idust: int) -> tuple:
    """the column names evwrite writes for these options"""
    if imhd != 0:
        return MHD
    if idust in (1, 3, 4):
        return DUST
    return HYDRO


def fortran_e(x: float, w: int, d: int) -> str:
    """x under the edit descriptor 1pEw.d"""
    if math.isnan(x):
        return "NaN".rjust(w)
    if math.isinf(x):
        return ("Infinity" if x > 0 else "-Infinity").rjust(w)
    mant, ex = ("%.*E" % (d, x)).split("E")
    e = int(ex)
    s = mant + ("E%+03d" % e if abs(e) <= 99 else "%+04d" % e)
    return s.rjust(w) if len(s) <= w else "*" * w


def header_lines(names: Sequence[str]) -> list:
    """('#',24(a18,1x)) over the names: records of at most 24 names"""
    lines = []
    for k in range(0, len(names), 24):
        chunk = names[k:k + 24]
        lead = "#" if k == 0 else ""
        lines.append(lead + " ".join(n.rjust(18)[:18] for n in chunk))
    return lines


def data_line(vals: Iterable[float], mhd: bool) -> str:
    """(24(1pe18.10,1x),1pe8.2) for MHD lines, (24(1pe18.10,1x)) otherwise"""
    vals = [float(v) for v in vals]
    if mhd:
        head = vals[:24]
        return "".join(fortran_e(v, 18, 10) + " " for v in head) + fortran_e(vals[24], 8, 2)
    return " ".join(fortran_e(v, 18, 10) for v in vals)


class EvFile:
    """Appends .ev lines for one simulation: the header once, then one line
    per call (the reference writes the header on its first call, keeping a
    saved flag, src/evwrite_mhd.f90:275-280)."""

    def __init__(self, path: str, imhd: int, idust: int):
        self.path = path
        self.names = columns_for(imhd, idust)
        self.mhd = imhd != 0
        self._fh = open(path, "w")
        self._started = False

    def write(self, vals: Sequence[float]):
        if len(vals) != len(self.names):
            raise ValueError("expected %d columns, got %d" % (len(self.names), len(vals)))
        if not self._started:
            for line in header_lines(self.names):
                self._fh.write(line + "\n")
            self._started = True
        self._fh.write(data_line(vals, self.mhd) + "\n")
        self._fh.flush()

    def write_sim(self, sim, t: float, dt: float):
        """evaluate and append the line of a Simulation at time t, timestep dt"""
        vals = sim.ev_columns(t, dt)
        self.write(vals)
        return dict(zip(self.names, vals))

    def close(self):
        if self._fh:
            self._fh.close()
            self._fh = None


def parse(text: str) -> tuple:
    """(names, rows) of .ev text: names from the header records, one float
    row per data line"""
    names, rows = [], []
    for line in text.splitlines():
        if not line.strip():
            continue
        toks = line.split()
        if line.startswith("#"):
            names += line[1:].split()
        elif all(_is_number(t) for t in toks):
            rows.append([_to_float(t) for t in toks])
        else:
            names += toks
    return names, rows


def _to_float(t: str) -> float:
    if "E" in t.upper():
        return float(t)
    # three-digit exponents are written without the letter E (1.0+100)
    k = max(t.rfind("+"), t.rfind("-"))
    return float(t[:k] + "E" + t[k:]) if k > 0 else float(t)


def _is_number(t: str) -> bool:
    try:
        _to_float(t)
        return True
    except ValueError:
        return False
