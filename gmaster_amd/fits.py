"""Minimal FITS reader / writer for NaMaster-compatible workspace files.

Only what workspace I/O needs (FITS standard 4.0): a primary image HDU (BITPIX -64) and
binary-table extensions with scalar columns of type J (int32), K (int64), E (float32) and
D (float64).  Headers are 80-character cards in 2880-byte blocks; data are big-endian and
padded to 2880 bytes.  The layout of the workspace file follows
``gmaster/workspaces.py:2162-2255`` and ``gmaster/bins.py:275-302``: a primary image
``WSP_PRIMARY`` holding the dense coupling matrix and the workspace keywords, then tables
``BEAMS`` (L, BEAM1, BEAM2), ``PCL_MASKS`` (L, PCL_MASKS) and ``BANDPOWERS`` (BAND, ELLS,
WEIGHTS, F_ELL; keywords ELL_MAX, N_BANDS).
"""

import numpy as np

BLOCK = 2880
_TFORM = {"J": ">i4", "K": ">i8", "E": ">f4", "D": ">f8", "I": ">i2"}
_CODE = {np.dtype(np.int32): "J", np.dtype(np.int64): "K", np.dtype(np.float32): "E", np.dtype(np.float64): "D",
         np.dtype(np.int16): "I"}


def _card(key, value=None):
    """One 80-column card; keywords longer than 8 characters use the HIERARCH convention
    (``HIERARCH KEY = value``, as cfitsio writes and reads them)."""
    key = key.upper()
    if value is None:
        if len(key) > 8:
            raise ValueError(f"FITS keyword too long: {key}")
        return key.ljust(80)
    if len(key) > 8:
        card = f"HIERARCH {key} = {_value(value).strip()}"
        if len(card) > 80:
            raise ValueError(f"FITS card too long: {card}")
        return card.ljust(80)
    card = f"{key:<8}= {_value(value)}"
    if len(card) > 80:
        raise ValueError(f"FITS card too long: {card}")
    return card.ljust(80)


def _value(value):
    if isinstance(value, (bool, np.bool_)):
        v = ("T" if value else "F").rjust(20)
    elif isinstance(value, (int, np.integer)):
        v = str(int(value)).rjust(20)
    elif isinstance(value, (float, np.floating)):
        s = repr(float(value)).upper()
        if "." not in s and "E" not in s and "INF" not in s and "NAN" not in s:
            s += ".0"
        v = s.rjust(20)
    else:
        s = str(value).replace("'", "''")
        v = ("'" + s.ljust(8) + "'").ljust(20)
    return v


def _header_bytes(cards):
    text = "".join(cards) + "END".ljust(80)
    pad = (-len(text)) % BLOCK
    return (text + " " * pad).encode("ascii")


def _pad(data):
    return data + b"\0" * ((-len(data)) % BLOCK)


def write(path, image, image_header, tables):
    """``image``: 2-D float64 array (primary HDU); ``image_header``: dict of keywords (EXTNAME
    included); ``tables``: list of ``(extname, [(column, 1-D array)], header dict)``."""
    image = np.ascontiguousarray(image, dtype=np.float64)
    if image.ndim != 2:
        raise ValueError("the primary image must be 2-D")
    cards = [_card("SIMPLE", True), _card("BITPIX", -64), _card("NAXIS", 2), _card("NAXIS1", image.shape[1]),
             _card("NAXIS2", image.shape[0]), _card("EXTEND", True)]
    cards += [_card(k, v) for k, v in image_header.items()]
    out = [_header_bytes(cards), _pad(image.astype(">f8").tobytes())]
    for extname, cols, header in tables:
        arrays = [np.asarray(a) for _, a in cols]
        n = len(arrays[0])
        if any(a.ndim != 1 or len(a) != n for a in arrays):
            raise ValueError("table columns must be 1-D and of equal length")
        codes = [_CODE[a.dtype] for a in arrays]
        rec = np.empty(n, dtype=[(name, _TFORM[c]) for (name, _), c in zip(cols, codes)])
        for (name, _), a in zip(cols, arrays):
            rec[name] = a
        cards = [_card("XTENSION", "BINTABLE"), _card("BITPIX", 8), _card("NAXIS", 2),
                 _card("NAXIS1", rec.dtype.itemsize), _card("NAXIS2", n), _card("PCOUNT", 0), _card("GCOUNT", 1),
                 _card("TFIELDS", len(cols))]
        for i, ((name, _), c) in enumerate(zip(cols, codes), start=1):
            cards += [_card(f"TTYPE{i}", name), _card(f"TFORM{i}", "1" + c)]
        cards.append(_card("EXTNAME", extname))
        cards += [_card(k, v) for k, v in (header or {}).items()]
        out += [_header_bytes(cards), _pad(rec.tobytes())]
    with open(path, "wb") as f:
        for chunk in out:
            f.write(chunk)


def _parse_value(raw):
    raw = raw.strip()
    if raw.startswith("'"):
        end = 1
        while True:
            end = raw.index("'", end)
            if end + 1 < len(raw) and raw[end + 1] == "'":
                end += 2
                continue
            break
        return raw[1:end].replace("''", "'").rstrip()
    val = raw.split("/", 1)[0].strip()
    if val == "T":
        return True
    if val == "F":
        return False
    try:
        return int(val)
    except ValueError:
        return float(val.replace("D", "E"))


def read(path):
    """All HDUs as ``[(header dict, data)]``: data is a 2-D float array for images (or None
    without data) and a dict of columns for binary tables."""
    buf = open(path, "rb").read()
    pos, hdus = 0, []
    while pos < len(buf):
        header = {}
        done = False
        while not done:
            block = buf[pos:pos + BLOCK].decode("ascii")
            if len(block) < BLOCK:
                raise ValueError("truncated FITS header")
            pos += BLOCK
            for i in range(0, BLOCK, 80):
                card = block[i:i + 80]
                key = card[:8].strip()
                if key == "END":
                    done = True
                    break
                if key == "HIERARCH" and "=" in card:
                    k, v = card[8:].split("=", 1)
                    header[k.strip().upper()] = _parse_value(v)
                elif card[8:10] == "= ":
                    header[key] = _parse_value(card[10:])
        naxis = int(header.get("NAXIS", 0))
        dims = [int(header[f"NAXIS{i}"]) for i in range(1, naxis + 1)]
        if header.get("XTENSION", "").strip() == "BINTABLE":
            nbytes = dims[0] * dims[1] + int(header.get("PCOUNT", 0))
            fields = []
            for i in range(1, int(header["TFIELDS"]) + 1):
                form = str(header[f"TFORM{i}"]).strip()
                rep = int(form[:-1]) if form[:-1] else 1
                code = form[-1]
                if code not in _TFORM:
                    raise ValueError(f"unsupported FITS column type {form}")
                dt = _TFORM[code] if rep == 1 else (_TFORM[code], (rep,))
                fields.append((str(header[f"TTYPE{i}"]).strip(), dt))
            rec = np.frombuffer(buf[pos:pos + dims[0] * dims[1]], dtype=fields, count=dims[1])
            data = {name: np.array(rec[name]) for name, _ in fields}
        else:
            bitpix = int(header["BITPIX"])
            count = int(np.prod(dims)) if dims else 0
            nbytes = count * abs(bitpix) // 8
            dt = {-64: ">f8", -32: ">f4", 8: ">u1", 16: ">i2", 32: ">i4", 64: ">i8"}[bitpix]
            data = None
            if count:
                data = np.frombuffer(buf[pos:pos + nbytes], dtype=dt).reshape(dims[::-1]).astype(np.float64)
        pos += nbytes + ((-nbytes) % BLOCK)
        hdus.append((header, data))
    return hdus


def find(hdus, extname):
    for header, data in hdus:
        if str(header.get("EXTNAME", "")).strip().upper() == extname.upper():
            return header, data
    raise KeyError(f"no HDU named {extname}")
