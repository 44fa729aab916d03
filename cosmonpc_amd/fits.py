"""Minimal FITS binary-table reader and writer (FITS standard 4.0, section 7.3).

The reference reads survey catalogs with ``fitsio`` (``fits_reader``,
``catalog_processor.py:130-176``): the first extension when the file has one
(``fits[1] if len(fits) > 1 else fits[0]``), the requested columns that exist,
as a numpy structured array with the file's column types.  ``fitsio`` is not
part of this framework, so the subset of the format that catalogs use is
implemented here: 2880-byte blocks of 80-character header cards, a BINTABLE
extension with scalar columns of TFORM codes L, B, I, J, K, E, D and A
(character strings), big-endian row-major data.
"""
import numpy as np

BLOCK = 2880
CARD = 80

# TFORM letter -> big-endian numpy type (A handled separately)
_TFORM = {"L": "i1", "B": "u1", "I": ">i2", "J": ">i4", "K": ">i8", "E": ">f4", "D": ">f8"}
_NUMPY = {"b": "L", "u1": "B", "i2": "I", "i4": "J", "i8": "K", "f4": "E", "f8": "D"}


def _parse_header(buf, off):
    """Cards of one HDU header starting at ``off``; returns (dict, order, data offset)."""
    cards, order = {}, []
    pos = off
    while True:
        if pos + BLOCK > len(buf):
            raise ValueError("truncated FITS header")
        block = buf[pos:pos + BLOCK]
        pos += BLOCK
        done = False
        for i in range(0, BLOCK, CARD):
            card = block[i:i + CARD].decode("ascii", errors="replace")
            key = card[:8].strip()
            if key == "END":
                done = True
                break
            if card[8:10] != "= " or not key:
                continue
            val = card[10:]
            if val.lstrip().startswith("'"):
                s = val.lstrip()[1:]
                out, j = [], 0
                while j < len(s):
                    if s[j] == "'":
                        if j + 1 < len(s) and s[j + 1] == "'":
                            out.append("'")
                            j += 2
                            continue
                        break
                    out.append(s[j])
                    j += 1
                v = "".join(out).rstrip()
            else:
                v = val.split("/")[0].strip()
                if v in ("T", "F"):
                    v = v == "T"
                else:
                    try:
                        v = int(v)
                    except ValueError:
                        try:
                            v = float(v.replace("D", "E"))
                        except ValueError:
                            pass
            cards[key] = v
            order.append(key)
        if done:
            return cards, order, pos


def _data_bytes(h):
    naxis = int(h.get("NAXIS", 0))
    if naxis == 0:
        return 0
    n = abs(int(h.get("BITPIX", 8))) // 8
    for i in range(1, naxis + 1):
        n *= int(h[f"NAXIS{i}"])
    return (n + int(h.get("PCOUNT", 0))) * int(h.get("GCOUNT", 1))


def _hdus(buf):
    off, out = 0, []
    while off < len(buf):
        if not buf[off:off + 8].strip():
            break
        h, order, data = _parse_header(buf, off)
        nb = _data_bytes(h)
        out.append((h, data, nb))
        off = data + ((nb + BLOCK - 1) // BLOCK) * BLOCK
    return out


def _column_dtype(tform):
    t = tform.strip()
    i = 0
    while i < len(t) and t[i].isdigit():
        i += 1
    rep = int(t[:i]) if i else 1
    code = t[i]
    if code == "A":
        return f"S{rep}", rep
    if code not in _TFORM:
        raise ValueError(f"unsupported FITS column format {tform!r}")
    if rep != 1:
        raise ValueError(f"vector FITS columns are not supported ({tform!r})")
    dt = np.dtype(_TFORM[code])
    return dt, dt.itemsize


def read_table(path, columns=None):
    """Columns of the table HDU (``fits[1]`` if present, else ``fits[0]``) as a
    structured array with native-endian column types; ``columns`` selects the
    names that exist, in the given order."""
    with open(path, "rb") as f:
        buf = f.read()
    hdus = _hdus(buf)
    if not hdus:
        raise ValueError(f"{path}: not a FITS file")
    h, data, nb = hdus[1] if len(hdus) > 1 else hdus[0]
    if h.get("XTENSION", "") != "BINTABLE":
        raise ValueError(f"{path}: HDU is not a binary table")
    ncol = int(h["TFIELDS"])
    rowlen = int(h["NAXIS1"])
    nrows = int(h["NAXIS2"])
    names, formats, offsets = [], [], []
    off = 0
    for i in range(1, ncol + 1):
        dt, size = _column_dtype(h[f"TFORM{i}"])
        names.append(str(h.get(f"TTYPE{i}", f"col{i}")).strip())
        formats.append(dt)
        offsets.append(off)
        off += size
    if off != rowlen:
        raise ValueError(f"{path}: row length {rowlen} does not match the column formats ({off})")
    raw = np.frombuffer(buf, dtype=np.dtype({"names": names, "formats": formats, "offsets": offsets,
                                             "itemsize": rowlen}), count=nrows, offset=data)
    want = [c for c in (columns if columns is not None else names) if c in names]
    out = np.empty(nrows, dtype=[(c, raw.dtype[c].newbyteorder("=") if raw.dtype[c].kind != "S" else raw.dtype[c])
                                 for c in want])
    for c in want:
        out[c] = raw[c]
    return out


def table_columns(path):
    """Column names of the table HDU."""
    with open(path, "rb") as f:
        buf = f.read()
    hdus = _hdus(buf)
    h = hdus[1][0] if len(hdus) > 1 else hdus[0][0]
    return [str(h.get(f"TTYPE{i}", f"col{i}")).strip() for i in range(1, int(h["TFIELDS"]) + 1)]


def _card(key, value, comment=""):
    if isinstance(value, bool):
        v = f"{'T' if value else 'F':>20}"
    elif isinstance(value, (int, np.integer)):
        v = f"{int(value):>20d}"
    elif isinstance(value, str):
        s = "'" + value.replace("'", "''").ljust(8) + "'"
        v = f"{s:<20}"
    else:
        raise TypeError(value)
    c = f"{key:<8}= {v}"
    if comment:
        c += " / " + comment
    return c[:CARD].ljust(CARD)


def write_table(path, columns):
    """Write ``columns`` (dict name -> 1-D array, or a structured array) as a
    primary HDU plus one BINTABLE extension."""
    if isinstance(columns, np.ndarray) and columns.dtype.names:
        columns = {n: columns[n] for n in columns.dtype.names}
    names = list(columns)
    arrs = [np.asarray(columns[n]) for n in names]
    nrows = len(arrs[0]) if arrs else 0
    if any(len(a) != nrows for a in arrs):
        raise ValueError("columns must have equal lengths")
    formats, dts = [], []
    for a in arrs:
        if a.dtype.kind == "S":
            formats.append(f"{a.dtype.itemsize}A")
            dts.append(a.dtype)
            continue
        key = "b" if a.dtype.kind == "b" else a.dtype.str[1:]
        if key not in _NUMPY:
            raise ValueError(f"unsupported column dtype {a.dtype}")
        formats.append(_NUMPY[key])
        dts.append(np.dtype("i1") if key == "b" else a.dtype.newbyteorder(">"))
    rec = np.empty(nrows, dtype=[(n, d) for n, d in zip(names, dts)])
    for n, a in zip(names, arrs):
        rec[n] = np.where(a, ord("T"), ord("F")) if a.dtype.kind == "b" else a
    rowlen = rec.dtype.itemsize
    prim = [_card("SIMPLE", True), _card("BITPIX", 8), _card("NAXIS", 0), _card("EXTEND", True), "END".ljust(CARD)]
    ext = [_card("XTENSION", "BINTABLE"), _card("BITPIX", 8), _card("NAXIS", 2), _card("NAXIS1", rowlen),
           _card("NAXIS2", nrows), _card("PCOUNT", 0), _card("GCOUNT", 1), _card("TFIELDS", len(names))]
    for i, (n, f) in enumerate(zip(names, formats), 1):
        ext += [_card(f"TTYPE{i}", n), _card(f"TFORM{i}", f)]
    ext.append("END".ljust(CARD))

    def pad(b, fill):
        return b + fill * ((-len(b)) % BLOCK)

    with open(path, "wb") as f:
        f.write(pad("".join(prim).encode("ascii"), b" "))
        f.write(pad("".join(ext).encode("ascii"), b" "))
        f.write(pad(rec.tobytes(), b"\0"))
