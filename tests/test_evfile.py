"""The .ev text layout (ndspmhd_amd.evfile) against the text the reference's
own evwrite wrote for the golden states (src/evwrite_mhd.f90:270-308): the
three column sets, the two-record MHD header, 1pe18.10 / 1pe8.2 numbers."""
import numpy as np
import pytest

import golden_util as gu
from ndspmhd_amd import evfile as ev


def _kind(o):
    return o.imhd, o.idust


@pytest.mark.parametrize("name", [n for n in gu.names() if n.endswith("_nr")])
def test_reference_text_reproduced(name):
    o, st, ref = gu.load(name)
    text = str(ref["ev"])
    names, rows = ev.parse(text)
    assert tuple(names) == ev.columns_for(*_kind(o))
    assert len(rows) == 1 and len(rows[0]) == len(names)
    mine = "\n".join(ev.header_lines(names)) + "\n" + ev.data_line(rows[0], o.imhd != 0) + "\n"
    assert mine == text


def test_column_sets_cover_reference_variants():
    got = {ev.columns_for(*_kind(gu.load(n)[0])) for n in gu.names()}
    assert got == {ev.MHD, ev.DUST, ev.HYDRO}
    assert len(ev.MHD) == 25 and len(ev.DUST) == 14 and len(ev.HYDRO) == 13


@pytest.mark.parametrize("x,w,d,s", [
    (0.0, 18, 10, "  0.0000000000E+00"),
    (-2.9047921579e-03, 18, 10, " -2.9047921579E-03"),
    (1.5e123, 18, 10, "  1.5000000000+123"),
    (-2.5e-150, 18, 10, " -2.5000000000-150"),
    (3.125, 8, 2, "3.12E+00"),
    (100.0, 8, 2, "1.00E+02"),
])
def test_edit_descriptor(x, w, d, s):
    assert ev.fortran_e(x, w, d) == s


def test_parse_three_digit_exponents():
    line = ev.data_line([1.5e123, -2.5e-150] + [0.0] * 11, False)
    _, rows = ev.parse(line)
    assert rows[0][:2] == [1.5e123, -2.5e-150]


def test_planted_column_swap_is_caught():
    """a swapped pair of columns (a plausible writer mistake) changes the text"""
    o, st, ref = gu.load("mhd3d_nr")
    names, rows = ev.parse(str(ref["ev"]))
    r = list(rows[0])
    r[1], r[2] = r[2], r[1]
    assert ev.data_line(r, True) not in str(ref["ev"])
