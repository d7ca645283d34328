"""Minimal FITS image I/O for the example data ProFit ships (KiDS and SDSS
cut-outs: a primary image HDU with BITPIX -32 or 32; the Offset example keeps
its images in IMAGE extensions after an empty primary HDU).

FITS layout (FITS Standard 4.0): a header of 80-character ASCII cards in
2880-byte blocks, ended by an END card; then the data, big-endian, NAXIS1
varying fastest, padded to a multiple of 2880 bytes; extensions follow with
the same structure.  Values are scaled as BZERO + BSCALE * stored.  Images
come back as numpy arrays indexed [y, x] (NAXIS2 rows of NAXIS1), the layout
of libprofit's Image (x fastest).
"""
from __future__ import annotations

import numpy as np

BLOCK = 2880
CARD = 80
_DTYPES = {8: ">u1", 16: ">i2", 32: ">i4", 64: ">i8", -32: ">f4", -64: ">f8"}


class FitsError(ValueError):
    pass


def _parse_value(raw):
    s = raw.split("/")[0].strip() if not raw.strip().startswith("'") else raw.strip()
    if s.startswith("'"):
        end = s.find("'", 1)
        while end != -1 and end + 1 < len(s) and s[end + 1] == "'":  # doubled quote
            end = s.find("'", end + 2)
        return s[1:end].replace("''", "'").rstrip() if end != -1 else s[1:]
    if s in ("T", "F"):
        return s == "T"
    try:
        return int(s)
    except ValueError:
        pass
    try:
        return float(s.replace("D", "E"))
    except ValueError:
        return s


def _read_header(buf, pos):
    """Header cards from byte offset pos; returns (dict, data offset)."""
    hdr = {}
    while True:
        if pos + BLOCK > len(buf):
            raise FitsError("truncated FITS header")
        block = buf[pos:pos + BLOCK]
        pos += BLOCK
        for k in range(0, BLOCK, CARD):
            card = block[k:k + CARD].decode("ascii", errors="replace")
            key = card[:8].strip()
            if key == "END":
                return hdr, pos
            if card[8:10] == "= " and key:
                hdr[key] = _parse_value(card[10:])


def _data_bytes(hdr):
    naxis = int(hdr.get("NAXIS", 0))
    if naxis == 0:
        return 0
    n = 1
    for a in range(1, naxis + 1):
        n *= int(hdr[f"NAXIS{a}"])
    gcount, pcount = int(hdr.get("GCOUNT", 1)), int(hdr.get("PCOUNT", 0))
    return abs(int(hdr["BITPIX"])) // 8 * gcount * (pcount + n)


def read_hdus(path):
    """Every HDU of a FITS file as (header dict, numpy array or None)."""
    with open(path, "rb") as f:
        buf = f.read()
    if not buf.startswith(b"SIMPLE  ="):
        raise FitsError(f"{path}: not a FITS file")
    out = []
    pos = 0
    while pos < len(buf):
        if not buf[pos:pos + 8].strip():  # trailing padding
            break
        hdr, pos = _read_header(buf, pos)
        nbytes = _data_bytes(hdr)
        data = None
        naxis = int(hdr.get("NAXIS", 0))
        ext = str(hdr.get("XTENSION", "IMAGE")).strip()
        if naxis > 0 and ext == "IMAGE":
            bitpix = int(hdr["BITPIX"])
            if bitpix not in _DTYPES:
                raise FitsError(f"unsupported BITPIX {bitpix}")
            shape = tuple(int(hdr[f"NAXIS{a}"]) for a in range(naxis, 0, -1))
            if pos + nbytes > len(buf):
                raise FitsError("truncated FITS data")
            raw = np.frombuffer(buf, dtype=_DTYPES[bitpix], count=int(np.prod(shape)), offset=pos).reshape(shape)
            bscale, bzero = float(hdr.get("BSCALE", 1.0)), float(hdr.get("BZERO", 0.0))
            if bitpix > 0 and bscale == 1.0 and bzero == 0.0:
                data = raw.astype(raw.dtype.newbyteorder("="))
            else:
                data = bzero + bscale * raw.astype(np.float64)
        out.append((hdr, data))
        pos += (nbytes + BLOCK - 1) // BLOCK * BLOCK
    return out


def read_image(path, hdu=None):
    """The image of HDU `hdu`, or of the first HDU that holds one."""
    hdus = read_hdus(path)
    if hdu is not None:
        img = hdus[hdu][1]
        if img is None:
            raise FitsError(f"{path}: HDU {hdu} holds no image")
        return img
    for _, img in hdus:
        if img is not None:
            return img
    raise FitsError(f"{path}: no image HDU")


def _card(key, value, comment=""):
    if isinstance(value, bool):
        v = f"{'T' if value else 'F':>20}"
    elif isinstance(value, (int, np.integer)):
        v = f"{int(value):>20d}"
    elif isinstance(value, float):
        v = f"{value:>20.14G}"
    else:
        v = "'" + str(value).replace("'", "''").ljust(8) + "'"
    s = f"{key:<8}= {v}"
    if comment:
        s += f" / {comment}"
    return s[:CARD].ljust(CARD)


def write_image(path, data, bitpix=-32, extra=None, extensions=()):
    """Writes `data` ([y, x]) as the primary HDU (and 2-D `extensions` as
    IMAGE extensions).  Integer BITPIX stores values as given."""
    def hdu_bytes(arr, primary, bp):
        arr = np.asarray(arr)
        cards = [_card("SIMPLE", True) if primary else _card("XTENSION", "IMAGE"),
                 _card("BITPIX", bp), _card("NAXIS", arr.ndim)]
        for a in range(arr.ndim):
            cards.append(_card(f"NAXIS{a + 1}", arr.shape[arr.ndim - 1 - a]))
        if primary:
            cards.append(_card("EXTEND", True))
        else:
            cards += [_card("PCOUNT", 0), _card("GCOUNT", 1)]
        for k, v in (extra or {}).items():
            if primary:
                cards.append(_card(k, v))
        cards.append("END".ljust(CARD))
        head = "".join(cards).encode("ascii")
        head += b" " * (-len(head) % BLOCK)
        body = arr.astype(_DTYPES[bp]).tobytes() if arr.size else b""
        body += b"\0" * (-len(body) % BLOCK)
        return head + body
    out = hdu_bytes(data if data is not None else np.zeros((0,)), True, bitpix) if data is not None else None
    if data is None:
        cards = [_card("SIMPLE", True), _card("BITPIX", 16), _card("NAXIS", 0), _card("EXTEND", True),
                 "END".ljust(CARD)]
        head = "".join(cards).encode("ascii")
        out = head + b" " * (-len(head) % BLOCK)
    for e in extensions:
        out += hdu_bytes(e, False, bitpix)
    with open(path, "wb") as f:
        f.write(out)
