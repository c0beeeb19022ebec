"""The FITS writer / reader of workspace I/O, pinned by the FITS standard's fixed format
(2880-byte blocks, 80-column cards with the value indicator in columns 9-10, logical and
integer values right-justified to column 30, quoted strings from column 11, big-endian data)
and by round trips."""

import numpy as np
import pytest

from gmaster_amd import fits


def _write_sample(path):
    img = np.arange(12, dtype=np.float64).reshape(3, 4) * 0.5 - 1.25
    fits.write(path, img, {"EXTNAME": "WSP_PRIMARY", "LMAX": 7, "IS_TEB": False, "WAWB": 0.125, "NOTE": "it's"},
               [("TAB", [("L", np.arange(5, dtype=np.int32)), ("X", np.linspace(0, 1, 5))], {"ELL_MAX": 4}),
                ("BIG", [("K", np.array([2 ** 40, -3], dtype=np.int64)), ("F", np.array([1.5, -2], np.float32))],
                 {})])
    return img


def test_round_trip(tmp_path):
    p = str(tmp_path / "a.fits")
    img = _write_sample(p)
    hdus = fits.read(p)
    assert len(hdus) == 3
    h, d = fits.find(hdus, "WSP_PRIMARY")
    np.testing.assert_array_equal(d, img)
    assert h["LMAX"] == 7 and h["IS_TEB"] is False and h["WAWB"] == 0.125 and h["NOTE"] == "it's"
    h, t = fits.find(hdus, "TAB")
    np.testing.assert_array_equal(t["L"], np.arange(5))
    np.testing.assert_array_equal(t["X"], np.linspace(0, 1, 5))
    assert t["L"].dtype.kind == "i" and h["ELL_MAX"] == 4
    _, t = fits.find(hdus, "BIG")
    assert list(t["K"]) == [2 ** 40, -3] and list(t["F"]) == [1.5, -2.0]
    with pytest.raises(KeyError):
        fits.find(hdus, "NOPE")


def test_fixed_format_layout(tmp_path):
    p = str(tmp_path / "a.fits")
    img = _write_sample(p)
    raw = open(p, "rb").read()
    assert len(raw) % 2880 == 0
    cards = [raw[i:i + 80].decode("ascii") for i in range(0, 2880, 80)]
    assert cards[0] == "SIMPLE  =" + " " * 20 + "T" + " " * 50
    assert cards[1] == "BITPIX  =" + " " * 18 + "-64" + " " * 50
    assert cards[3].startswith("NAXIS1  =") and cards[3][10:30].strip() == "4" and cards[3][29] == "4"
    end = next(i for i, c in enumerate(cards) if c.startswith("END"))
    assert cards[end] == "END" + " " * 77 and all(c == " " * 80 for c in cards[end + 1:])
    # primary data: big-endian doubles right after the one header block, zero padded
    np.testing.assert_array_equal(np.frombuffer(raw[2880:2880 + 96], dtype=">f8").reshape(3, 4), img)
    assert raw[2880 + 96:5760] == b"\0" * (5760 - 2976)
    # first extension header: XTENSION card, quoted string padded to at least 8 characters
    ext = [raw[5760 + i:5760 + i + 80].decode("ascii") for i in range(0, 2880, 80)]
    assert ext[0] == "XTENSION= 'BINTABLE'" + " " * 60
    assert ext[3] == "NAXIS1  =" + " " * 19 + "12" + " " * 50  # J + D = 4 + 8 bytes per row
    assert "TFORM1  = '1J      '" in "".join(ext)
    # table rows big-endian, fields packed in order
    row0 = raw[8640:8640 + 12]
    assert np.frombuffer(row0[:4], ">i4")[0] == 0 and np.frombuffer(row0[4:], ">f8")[0] == 0.0


def test_value_parsing():
    assert fits._parse_value("                   T / a comment") is True
    assert fits._parse_value("                  42") == 42
    assert fits._parse_value("              1.5E-3 / x") == 1.5e-3
    assert fits._parse_value("'O''Neil '") == "O'Neil"
    assert fits._card("WAWB", 0.125)[10:30].strip() == "0.125"
    assert fits._card("LMAX_MASK", 47).rstrip() == "HIERARCH LMAX_MASK = 47"
    assert fits._card("L_TOEPLITZ", -1).rstrip() == "HIERARCH L_TOEPLITZ = -1"


def test_long_keywords_round_trip(tmp_path):
    p = str(tmp_path / "h.fits")
    fits.write(p, np.zeros((1, 1)), {"EXTNAME": "WSP_PRIMARY", "LMAX_MASK": 47, "L_TOEPLITZ": -1, "PURE_E1": True},
               [])
    h, _ = fits.read(p)[0]
    assert h["LMAX_MASK"] == 47 and h["L_TOEPLITZ"] == -1 and h["PURE_E1"] is True
