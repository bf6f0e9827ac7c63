"""Minimal FITS reader / writer (numpy only) for redrock's template files.

astropy and fitsio are not dependencies here; this module reads what the
chi2-scan path needs from a FITS file: primary and IMAGE HDUs (BITPIX 8, 16,
32, 64, -32, -64 with BSCALE/BZERO) and BINTABLE HDUs with fixed-width
columns (L, B, I, J, K, A, E, D, with repeat counts and TZERO for unsigned
integers), plus header cards (strings with '' escapes and CONTINUE, logicals,
integers, floats).  It writes primary, IMAGE and BINTABLE HDUs in the same
subset, so template files and result tables can be produced and read back.

    hdus = read('rrtemplate-galaxy-v2.6.fits')
    hdus['BASIS_VECTORS'].header['CRVAL1'], hdus['BASIS_VECTORS'].data
"""
from __future__ import annotations

import numpy as np

BLOCK = 2880
CARD = 80

_BITPIX = {8: '>u1', 16: '>i2', 32: '>i4', 64: '>i8', -32: '>f4', -64: '>f8'}
_TFORM = {'L': ('S1', 1), 'B': ('>u1', 1), 'I': ('>i2', 2), 'J': ('>i4', 4), 'K': ('>i8', 8),
          'E': ('>f4', 4), 'D': ('>f8', 8)}


class HDU(object):
    """One header-data unit: header (dict, card order kept), data, name."""

    def __init__(self, header, data, name=''):
        self.header = header
        self.data = data
        self.name = name

    def __repr__(self):
        shape = getattr(self.data, 'shape', None) if not isinstance(self.data, dict) else \
            {k: v.shape for k, v in self.data.items()}
        return f'HDU({self.name!r}, {shape})'


class HDUList(list):
    """List of HDUs, indexable by position or (case-insensitive) EXTNAME."""

    def __getitem__(self, key):
        if isinstance(key, str):
            for h in self:
                if h.name.upper() == key.upper():
                    return h
            raise KeyError(key)
        return list.__getitem__(self, key)

    def __contains__(self, key):
        if isinstance(key, str):
            return any(h.name.upper() == key.upper() for h in self)
        return list.__contains__(self, key)


# ---------------------------------------------------------------- reading
def _parse_value(text):
    """Value field of a card (columns 11-80) -> python value."""
    s = text.strip()
    if not s:
        return None
    if s[0] == "'":
        out, i = [], 1
        while i < len(s):
            if s[i] == "'":
                if i + 1 < len(s) and s[i + 1] == "'":
                    out.append("'")
                    i += 2
                    continue
                break
            out.append(s[i])
            i += 1
        return ''.join(out).rstrip()
    v = s.split('/')[0].strip()
    if v == 'T':
        return True
    if v == 'F':
        return False
    try:
        return int(v)
    except ValueError:
        pass
    try:
        return float(v.replace('D', 'E').replace('d', 'e'))
    except ValueError:
        return v


def _read_header(buf, off):
    header = dict()
    last = None
    while True:
        if off + BLOCK > len(buf):
            raise ValueError('truncated FITS header')
        block = buf[off:off + BLOCK].decode('ascii', errors='replace')
        off += BLOCK
        for i in range(BLOCK // CARD):
            card = block[i * CARD:(i + 1) * CARD]
            key = card[:8].strip()
            if key == 'END':
                return header, off
            if key == 'CONTINUE' and last is not None and isinstance(header[last], str):
                more = _parse_value(card[8:])
                if header[last].endswith('&'):
                    header[last] = header[last][:-1] + (more or '')
                continue
            if card[8:10] == '= ' and key:
                header[key] = _parse_value(card[10:])
                last = key


def _image(header, buf, off):
    naxis = int(header.get('NAXIS', 0))
    shape = tuple(int(header[f'NAXIS{i}']) for i in range(naxis, 0, -1))
    dt = np.dtype(_BITPIX[int(header['BITPIX'])])
    n = int(np.prod(shape)) if naxis else 0
    nbytes = n * dt.itemsize
    if n == 0:
        return None, nbytes
    data = np.frombuffer(buf, dtype=dt, count=n, offset=off).reshape(shape)
    data = data.astype(dt.newbyteorder('='))
    bscale, bzero = header.get('BSCALE', 1), header.get('BZERO', 0)
    if bscale != 1 or bzero != 0:
        data = data * bscale + bzero
    return data, nbytes


def _tform(form):
    form = form.strip()
    i = 0
    while i < len(form) and form[i].isdigit():
        i += 1
    rep = int(form[:i]) if i else 1
    code = form[i]
    if code == 'A':
        return f'S{rep}', 1, 'A', rep
    if code not in _TFORM:
        raise NotImplementedError(f'BINTABLE column format {form!r}')
    dt, _ = _TFORM[code]
    return dt, rep, code, rep


def _table(header, buf, off):
    rowbytes, nrows = int(header['NAXIS1']), int(header['NAXIS2'])
    nfields = int(header['TFIELDS'])
    names, formats, offsets, kinds = [], [], [], []
    pos = 0
    for k in range(1, nfields + 1):
        dt, rep, code, width = _tform(header[f'TFORM{k}'])
        name = header.get(f'TTYPE{k}', f'COL{k}')
        names.append(name)
        if code == 'A':
            formats.append(dt)
            size = width
        else:
            formats.append((dt, (rep,)) if rep != 1 else dt)
            size = np.dtype(dt).itemsize * rep
        offsets.append(pos)
        kinds.append((code, k))
        pos += size
    if pos != rowbytes:
        raise ValueError(f'BINTABLE row size {rowbytes} != sum of columns {pos}')
    dt = np.dtype({'names': names, 'formats': formats, 'offsets': offsets, 'itemsize': rowbytes})
    rec = np.frombuffer(buf, dtype=dt, count=nrows, offset=off)
    out = dict()
    for name, (code, k) in zip(names, kinds):
        col = rec[name]
        if code == 'A':
            col = np.char.rstrip(col.astype(str))
        elif code == 'L':
            col = col == b'T'
        else:
            col = col.astype(col.dtype.newbyteorder('='))
            tz = header.get(f'TZERO{k}', 0)
            ts = header.get(f'TSCAL{k}', 1)
            if tz == 2 ** 63 and code == 'K':
                col = (col.view(np.uint64) ^ np.uint64(1 << 63))
            elif tz == 2 ** 31 and code == 'J':
                col = (col.astype(np.int64) + 2 ** 31).astype(np.uint32)
            elif tz == 2 ** 15 and code == 'I':
                col = (col.astype(np.int32) + 2 ** 15).astype(np.uint16)
            elif ts != 1 or tz != 0:
                col = col * ts + tz
        out[name] = col
    return out, rowbytes * nrows + int(header.get('PCOUNT', 0))


def read(path):
    """All HDUs of a FITS file as an HDUList."""
    with open(path, 'rb') as fh:
        buf = fh.read()
    hdus = HDUList()
    off = 0
    while off < len(buf):
        if not buf[off:off + 8].strip(b' \x00'):
            break
        header, off = _read_header(buf, off)
        kind = header.get('XTENSION', 'IMAGE' if not hdus else None)
        if not hdus:
            if not header.get('SIMPLE', False):
                raise ValueError(f'{path}: not a FITS file')
            data, nbytes = _image(header, buf, off)
            name = header.get('EXTNAME', 'PRIMARY')
        elif kind == 'IMAGE':
            data, nbytes = _image(header, buf, off)
            name = header.get('EXTNAME', '')
        elif kind == 'BINTABLE':
            data, nbytes = _table(header, buf, off)
            name = header.get('EXTNAME', '')
        else:
            raise NotImplementedError(f'{path}: XTENSION {kind!r}')
        hdus.append(HDU(header, data, str(name).strip()))
        off += (nbytes + BLOCK - 1) // BLOCK * BLOCK
    return hdus


# ---------------------------------------------------------------- writing
def _card(key, value):
    if isinstance(value, bool):
        v = f'{"T" if value else "F":>20}'
    elif isinstance(value, (int, np.integer)):
        v = f'{int(value):>20}'
    elif isinstance(value, (float, np.floating)):
        v = f'{float(value):>20.16G}'
        if '.' not in v and 'E' not in v:
            v = f'{float(value):>20.1f}'
    else:
        s = str(value).replace("'", "''")
        v = f"'{s:<8}'"
    card = f'{key:<8}= {v}'
    if len(card) > CARD:
        raise ValueError(f'header card too long: {key}')
    return f'{card:<{CARD}}'


def _header_bytes(cards):
    text = ''.join(_card(k, v) for k, v in cards) + f'{"END":<{CARD}}'
    text += ' ' * ((-len(text)) % BLOCK)
    return text.encode('ascii')


def _pad(b, fill=b'\x00'):
    return b + fill * ((-len(b)) % BLOCK)


def image_hdu(data, name=None, header=None):
    """(cards, bytes) of an IMAGE extension holding `data` as big-endian float64/int."""
    a = np.asarray(data)
    if a.dtype.kind == 'f':
        a, bitpix = a.astype('>f8'), -64
    else:
        a, bitpix = a.astype('>i8'), 64
    cards = [('XTENSION', 'IMAGE'), ('BITPIX', bitpix), ('NAXIS', a.ndim)]
    cards += [(f'NAXIS{i + 1}', n) for i, n in enumerate(a.shape[::-1])]
    cards += [('PCOUNT', 0), ('GCOUNT', 1)]
    if name:
        cards.append(('EXTNAME', name))
    cards += list((header or dict()).items())
    return cards, a.tobytes()


def table_hdu(columns, name=None, header=None):
    """(cards, bytes) of a BINTABLE extension from {name: array} (1-D or 2-D)."""
    cols = [(k, np.asarray(v)) for k, v in columns.items()]
    nrows = len(cols[0][1]) if cols else 0
    fields = []
    cards_cols = []
    for i, (k, v) in enumerate(cols, start=1):
        if len(v) != nrows:
            raise ValueError('columns of different lengths')
        rep = int(np.prod(v.shape[1:])) if v.ndim > 1 else 1
        if v.dtype.kind in 'US':
            width = max(1, int(max((len(str(x)) for x in v), default=1)))
            fields.append((k, f'S{width}'))
            cards_cols += [(f'TTYPE{i}', k), (f'TFORM{i}', f'{width}A')]
        elif v.dtype.kind == 'b':
            fields.append((k, 'S1') if rep == 1 else (k, 'S1', (rep,)))
            cards_cols += [(f'TTYPE{i}', k), (f'TFORM{i}', f'{rep}L')]
        elif v.dtype.kind in 'iu':
            fields.append((k, '>i8') if rep == 1 else (k, '>i8', (rep,)))
            cards_cols += [(f'TTYPE{i}', k), (f'TFORM{i}', f'{rep}K')]
        else:
            fields.append((k, '>f8') if rep == 1 else (k, '>f8', (rep,)))
            cards_cols += [(f'TTYPE{i}', k), (f'TFORM{i}', f'{rep}D')]
    dt = np.dtype(fields)
    rec = np.zeros(nrows, dtype=dt)
    for k, v in cols:
        if v.dtype.kind in 'US':
            rec[k] = np.char.encode(v.astype(str), 'ascii')
        elif v.dtype.kind == 'b':
            rec[k] = np.where(v, b'T', b'F')
        else:
            rec[k] = v.reshape(rec[k].shape)
    cards = [('XTENSION', 'BINTABLE'), ('BITPIX', 8), ('NAXIS', 2), ('NAXIS1', dt.itemsize),
             ('NAXIS2', nrows), ('PCOUNT', 0), ('GCOUNT', 1), ('TFIELDS', len(cols))] + cards_cols
    if name:
        cards.append(('EXTNAME', name))
    cards += list((header or dict()).items())
    return cards, rec.tobytes()


def write(path, hdus, primary_header=None):
    """Write an empty primary HDU (with primary_header cards) and `hdus`
    [(cards, bytes) from image_hdu / table_hdu]."""
    out = [_header_bytes([('SIMPLE', True), ('BITPIX', 8), ('NAXIS', 0), ('EXTEND', True)]
                         + list((primary_header or dict()).items()))]
    for cards, data in hdus:
        out.append(_header_bytes(cards))
        out.append(_pad(data))
    with open(path, 'wb') as fh:
        fh.write(b''.join(out))
