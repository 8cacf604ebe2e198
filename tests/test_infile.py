"""Input files (ndspmhd_amd.infile) against the text the reference's own
write_infile wrote (golden vectors; src/readwrite_infile.f90:35-147,
src/utils_infiles.f90:83-260), and the reader's rules."""
import numpy as np
import pytest

import golden_util as gu
from ndspmhd_amd import infile as inf


@pytest.mark.parametrize("name", gu.names())
def test_reference_text_reproduced(name):
    o, st, ref = gu.load(name)
    for key in ("infile", "infile2"):
        text = str(ref[key])
        p = inf.parse_infile(text, strict=True)
        assert inf.format_infile(p.values) == text, key


@pytest.mark.parametrize("name", gu.names())
def test_values_read_are_the_harness_options(name):
    """the reference wrote the options the fixture ran with; reading its
    file gives them back (to its printed precision)"""
    o, st, ref = gu.load(name)
    v = inf.parse_infile(str(ref["infile"]), strict=True).values
    for k in ("ikernel", "iener", "imhd", "imagforce", "idivbzero", "iav", "idust",
              "maxdensits", "idustevol"):
        if k in ("imagforce", "idivbzero") and o.imhd == 0:
            continue
        assert v[k] == getattr(o, k), k
    assert (v["iavlim"], v["iavlimu"], v["iavlimB"]) == tuple(o.iavlim)
    for k in ("hfact", "tolh", "gamma", "alphamin", "alphaumin", "alphaBmin", "beta",
              "avdecayconst", "C_cour", "C_force"):
        assert v[k] == pytest.approx(getattr(o, k), rel=1e-8), k
    back = inf.to_options(v, o.ndim, o.ibound, o.xmin, o.xmax)
    assert back.iavlim == o.iavlim and back.imhd == o.imhd


def test_varied_values_parse():
    o, st, ref = gu.load("mhd3d_nr")
    v = inf.parse_infile(str(ref["infile2"]), strict=True).values
    assert v["psep"] == 36.87 and v["tmax"] == 0.075 and v["C_cour"] == -0.3
    assert v["hsoft"] == 1.235e4 and v["etamhd"] == 1e-12 and v["rhocrit"] == 99.5
    assert v["tout"] == 1.23456789 and v["alphamin"] == 0.123456789


def test_reader_rules():
    text = "\n".join([
        "# comment line",
        "",
        "   psep = 2.5d-2   ! d exponent",
        "   tmax = 01:02:03 ! a time",
        "   nout = 5",
        "   usenumdens = .false.",
        "   hfact = 1.3 ! first one wins",
        "   hfact = 1.9",
        "   no equals sign here",
    ])
    f = inf.parse_infile(text)
    assert f.values["psep"] == 0.025
    assert f.values["tmax"] == (1 * 60 + 2) * 60 + 3
    assert f.values["nout"] == 5 and f.values["usenumdens"] is False
    assert f.values["hfact"] == 1.3
    assert "iener" in f.missing and "iavlim" not in f.missing
    # MHD options are only required when MHD is on
    assert "imagforce" not in f.missing
    with pytest.raises(inf.InfileError):
        inf.parse_infile(text, strict=True)


def test_checks_stop_where_the_reference_stops():
    v = dict(inf.DEFAULTS)
    v["nout"] = 0
    with pytest.raises(inf.InfileError):
        inf.check(v)
    v = dict(inf.DEFAULTS)
    v.update(gamma=1.0, iener=2)
    with pytest.raises(inf.InfileError):
        inf.check(v)
    v = dict(inf.DEFAULTS)
    v.update(tout=5.0)
    assert "no output tout > tmax" in inf.check(v)


def test_unsupported_options_rejected():
    v = dict(inf.DEFAULTS)
    v["ivisc"] = 1
    with pytest.raises(inf.InfileError, match="ivisc"):
        inf.to_options(v, 1)


def test_planted_format_mistake_is_caught():
    """a real written with one digit fewer (a plausible formatting slip)
    no longer matches the reference's text"""
    o, st, ref = gu.load("mhd3d_nr")
    text = str(ref["infile"])
    v = inf.parse_infile(text).values
    good = inf.format_option("gamma", v["gamma"], "adiabatic index")
    assert good in text
    bad = good.replace("1.66666667", " 1.6666667")
    assert bad not in text
