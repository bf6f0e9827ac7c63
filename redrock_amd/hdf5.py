"""Minimal HDF5 reader / writer (numpy only) for redrock's details files.

h5py is not a dependency here.  This module covers the subset of the HDF5
file format (format specification version 2/3, "earliest" library version)
that h5py writes for redrock's details file (results.py:18-97):

  * superblock version 0, 8-byte offsets and lengths;
  * groups stored as symbol tables (version-1 B-tree of symbol table nodes
    plus a local heap of link names), version-1 object headers with
    continuation blocks;
  * datasets with contiguous or compact layout (chunked layout is read too,
    without filters), dataspaces of any rank, and datatypes: integers,
    IEEE floats, fixed-length strings, compounds (members of any of these,
    including array members) and arrays.

Reading returns plain numpy arrays (structured arrays for compounds); the
writer produces the same structures, so a file written here reads back here
and has the layout h5py itself uses for such data.

    f = read('rrdetails.h5')          # Group: dict-like, nested
    f['zscan/GALAXY/zchi2']           # numpy array
"""
from __future__ import annotations

import struct

import numpy as np

SIGNATURE = b'\x89HDF\r\n\x1a\n'
UNDEF = 0xFFFFFFFFFFFFFFFF
LEAF_K = 4          # symbol table node holds 2 * LEAF_K entries
INTERNAL_K = 16     # group B-tree node holds 2 * INTERNAL_K children


class Group(dict):
    """A group: name -> Group or numpy array; '/'-separated paths index into
    subgroups."""

    def __getitem__(self, key):
        parts = [p for p in key.split('/') if p]
        node = self
        for p in parts:
            node = dict.__getitem__(node, p)
        return node

    def __contains__(self, key):
        try:
            self[key]
            return True
        except (KeyError, TypeError):
            return False


# ---------------------------------------------------------------- reading
class _Reader(object):
    def __init__(self, buf):
        self.b = buf
        if buf[:8] != SIGNATURE:
            raise ValueError('not an HDF5 file')
        ver = buf[8]
        if ver not in (0, 1):
            raise NotImplementedError(f'HDF5 superblock version {ver}')
        self.so, self.sl = buf[13], buf[14]
        if (self.so, self.sl) != (8, 8):
            raise NotImplementedError('HDF5 offsets/lengths other than 8 bytes')
        p = 24 + (4 if ver == 1 else 0)
        self.base = self.u(p, 8)
        root = p + 32
        self.root_header = self.u(root + 8, 8)

    def u(self, off, n):
        return int.from_bytes(self.b[off:off + n], 'little')

    # object headers --------------------------------------------------
    def messages(self, addr):
        b = self.b
        if b[addr:addr + 4] == b'OHDR':
            raise NotImplementedError('version-2 object headers')
        if b[addr] != 1:
            raise ValueError(f'object header version {b[addr]} at {addr}')
        nmsg = self.u(addr + 2, 2)
        size = self.u(addr + 8, 4)
        blocks = [(addr + 16, size)]
        out = []
        while blocks and len(out) < nmsg:
            start, length = blocks.pop(0)
            p, end = start, start + length
            while p + 8 <= end and len(out) < nmsg:
                mtype, msize, flags = self.u(p, 2), self.u(p + 2, 2), b[p + 4]
                body = p + 8
                if mtype == 0x10:
                    blocks.append((self.u(body, 8), self.u(body + 8, 8)))
                out.append((mtype, body, msize, flags))
                p = body + msize
        return out

    def obj(self, addr):
        msgs = self.messages(addr)
        kinds = {m[0] for m in msgs}
        if 0x11 in kinds:
            body = [m for m in msgs if m[0] == 0x11][0][1]
            return self.group(self.u(body, 8), self.u(body + 8, 8))
        if 0x08 in kinds:
            return self.dataset(msgs)
        if 0x06 in kinds or 0x02 in kinds or 0x0A in kinds:
            raise NotImplementedError('link-message (new-style) groups')
        return Group()

    # groups ------------------------------------------------------------
    def heap(self, addr):
        if self.b[addr:addr + 4] != b'HEAP':
            raise ValueError(f'no local heap at {addr}')
        return self.u(addr + 24, 8)

    def name(self, heap_data, off):
        p = heap_data + off
        e = self.b.index(b'\x00', p)
        return self.b[p:e].decode('ascii')

    def group(self, btree, heap):
        data = self.heap(heap)
        g = Group()
        for name, hdr in self._btree_entries(btree, data):
            dict.__setitem__(g, name, self.obj(hdr))
        return g

    def _btree_entries(self, addr, heap_data):
        b = self.b
        if b[addr:addr + 4] != b'TREE':
            raise ValueError(f'no B-tree node at {addr}')
        ntype, level, used = b[addr + 4], b[addr + 5], self.u(addr + 6, 2)
        if ntype != 0:
            raise ValueError('group B-tree expected')
        p = addr + 24
        children = []
        for i in range(used):
            children.append(self.u(p + 8, 8))     # skip key i
            p += 16
        out = []
        for c in children:
            if level > 0:
                out.extend(self._btree_entries(c, heap_data))
                continue
            if b[c:c + 4] != b'SNOD':
                raise ValueError(f'no symbol table node at {c}')
            n = self.u(c + 6, 2)
            for k in range(n):
                e = c + 8 + 40 * k
                out.append((self.name(heap_data, self.u(e, 8)), self.u(e + 8, 8)))
        return out

    # datasets ----------------------------------------------------------
    def dataspace(self, p):
        ver, rank, flags = self.b[p], self.b[p + 1], self.b[p + 2]
        q = p + (8 if ver == 1 else 4)
        if ver == 2 and self.b[p + 3] == 0:       # scalar
            return ()
        if ver == 2 and self.b[p + 3] == 2:       # null
            return None
        return tuple(self.u(q + 8 * i, 8) for i in range(rank))

    def datatype(self, p):
        """numpy dtype of the datatype message at p, and its encoded length."""
        b = self.b
        cv = b[p]
        cls, ver = cv & 0x0F, cv >> 4
        bits = b[p + 1] | (b[p + 2] << 8) | (b[p + 3] << 16)
        size = self.u(p + 4, 4)
        q = p + 8
        order = '>' if bits & 1 else '<'
        if cls == 0:
            kind = 'i' if bits & 0x08 else 'u'
            return np.dtype(f'{order}{kind}{size}'), 8 + 4
        if cls == 1:
            return np.dtype(f'{order}f{size}'), 8 + 12
        if cls == 3:
            return np.dtype(f'S{size}'), 8
        if cls == 6:
            nmemb = bits & 0xFFFF
            names, formats, offsets = [], [], []
            for _ in range(nmemb):
                e = b.index(b'\x00', q)
                name = b[q:e].decode('ascii')
                nl = e - q + 1
                q += (nl + 7) // 8 * 8 if ver < 3 else nl
                if ver < 3:
                    off = self.u(q, 4)
                    q += 4
                else:
                    nb = 1 if size < 256 else 2 if size < 65536 else 3 if size < 2 ** 24 else 4
                    off = self.u(q, nb)
                    q += nb
                dims = ()
                if ver == 1:
                    nd = b[q]
                    dims = tuple(self.u(q + 12 + 4 * i, 4) for i in range(nd))
                    q += 28
                mt, ml = self.datatype(q)
                q += ml
                names.append(name)
                formats.append((mt, dims) if dims else mt)
                offsets.append(off)
            dt = np.dtype({'names': names, 'formats': formats, 'offsets': offsets, 'itemsize': size})
            return dt, q - p
        if cls == 10:
            nd = b[q]
            if ver == 2:
                dims = tuple(self.u(q + 4 + 4 * i, 4) for i in range(nd))
                q += 4 + 8 * nd
            else:
                dims = tuple(self.u(q + 1 + 4 * i, 4) for i in range(nd))
                q += 1 + 4 * nd
            bt, bl = self.datatype(q)
            return np.dtype((bt, dims)), q + bl - p
        if cls == 8:      # enum: h5py's FALSE/TRUE enum is numpy bool, others their base
            nmemb = bits & 0xFFFF
            bt, bl = self.datatype(q)
            q += bl
            names = []
            for _ in range(nmemb):
                e = b.index(b'\x00', q)
                names.append(b[q:e])
                nl = e - q + 1
                q += (nl + 7) // 8 * 8 if ver < 3 else nl
            q += nmemb * size
            if sorted(names) == [b'FALSE', b'TRUE'] and size == 1:
                return np.dtype('?'), q - p
            return bt, q - p
        raise NotImplementedError(f'HDF5 datatype class {cls}')

    def dataset(self, msgs):
        shape = dtype = layout = None
        for mtype, body, msize, flags in msgs:
            if mtype == 0x01:
                shape = self.dataspace(body)
            elif mtype == 0x03:
                dtype, _ = self.datatype(body)
            elif mtype == 0x08:
                layout = body
            elif mtype == 0x0B:
                raise NotImplementedError('filtered (compressed) datasets')
        if shape is None:
            return np.zeros(0, dtype=dtype)
        n = int(np.prod(shape)) if shape else 1
        nbytes = n * dtype.itemsize
        b = self.b
        ver = b[layout]
        if ver == 3:
            cls = b[layout + 1]
            if cls == 0:
                sz = self.u(layout + 2, 2)
                raw = b[layout + 4:layout + 4 + sz]
            elif cls == 1:
                addr, sz = self.u(layout + 2, 8), self.u(layout + 10, 8)
                raw = bytes(nbytes) if addr == UNDEF else b[addr:addr + nbytes]
            else:
                rank = b[layout + 2] - 1
                addr = self.u(layout + 3, 8)
                chunk = tuple(self.u(layout + 11 + 4 * i, 4) for i in range(rank))
                return self.chunked(addr, shape, chunk, dtype)
        else:
            nd, cls = b[layout + 1], b[layout + 2]
            p = layout + 8
            if cls == 0:
                p += 4 * nd
                sz = self.u(p, 4)
                raw = b[p + 4:p + 4 + sz]
            else:
                addr = self.u(p, 8)
                raw = bytes(nbytes) if addr == UNDEF else b[addr:addr + nbytes]
        arr = np.frombuffer(raw, dtype=dtype, count=n).reshape(shape)
        return _native(arr)

    def chunked(self, addr, shape, chunk, dtype):
        out = np.zeros(shape, dtype=dtype)
        if addr == UNDEF:
            return _native(out)
        rank = len(shape)
        for offs, caddr, csize in self._chunks(addr, rank):
            raw = self.b[caddr:caddr + csize]
            block = np.frombuffer(raw, dtype=dtype, count=int(np.prod(chunk))).reshape(chunk)
            sl = tuple(slice(o, min(o + c, s)) for o, c, s in zip(offs, chunk, shape))
            out[sl] = block[tuple(slice(0, x.stop - x.start) for x in sl)]
        return _native(out)

    def _chunks(self, addr, rank):
        b = self.b
        if b[addr:addr + 4] != b'TREE' or b[addr + 4] != 1:
            raise ValueError('chunk B-tree expected')
        level, used = b[addr + 5], self.u(addr + 6, 2)
        ksize = 8 + 8 * (rank + 1)
        p = addr + 24
        out = []
        for i in range(used):
            csize = self.u(p, 4)
            offs = tuple(self.u(p + 8 + 8 * d, 8) for d in range(rank))
            child = self.u(p + ksize, 8)
            if level > 0:
                out.extend(self._chunks(child, rank))
            else:
                out.append((offs, child, csize))
            p += ksize + 8
        return out


def _native(arr):
    """Copy with native byte order (fields of structured arrays too)."""
    dt = arr.dtype
    if dt.fields is None:
        if dt.subdtype is not None:
            return arr.astype(np.dtype((dt.subdtype[0].newbyteorder('='), dt.subdtype[1])))
        return arr.astype(dt.newbyteorder('=')) if dt.kind in 'iuf' else arr.copy()
    fmt = []
    for name in dt.names:
        ft = dt.fields[name][0]
        if ft.subdtype is not None:
            base, dims = ft.subdtype
            fmt.append((name, base.newbyteorder('=') if base.kind in 'iuf' else base, dims))
        else:
            fmt.append((name, ft.newbyteorder('=') if ft.kind in 'iuf' else ft))
    out = np.empty(arr.shape, dtype=fmt)
    for name in dt.names:
        out[name] = arr[name]
    return out


def read(path):
    """The whole file as a nested Group of numpy arrays."""
    with open(path, 'rb') as fh:
        buf = fh.read()
    r = _Reader(buf)
    return r.obj(r.root_header)


# ---------------------------------------------------------------- writing
def _pad8(b):
    return b + bytes((-len(b)) % 8)


def _dtype_message(dt):
    """Encoded datatype message (little-endian) for numpy dtype dt, in the
    versions the HDF5 library writes by default."""
    if dt.fields is not None:
        arrays = any(dt.fields[n][0].subdtype is not None for n in dt.names)
        ver = 2 if arrays else 1
        body = b''
        for name in dt.names:
            ft, off = dt.fields[name][:2]
            body += _pad8(name.encode('ascii') + b'\x00') + struct.pack('<I', off)
            if ver == 1:
                body += bytes(28)
            body += _dtype_message(ft)
        n = len(dt.names)
        return struct.pack('<BBBBI', (ver << 4) | 6, n & 0xFF, n >> 8, 0, dt.itemsize) + body
    if dt.subdtype is not None:
        base, dims = dt.subdtype
        body = struct.pack('<B3x', len(dims)) + b''.join(struct.pack('<I', d) for d in dims)
        body += b''.join(struct.pack('<I', i) for i in range(len(dims)))
        return struct.pack('<BBBBI', (2 << 4) | 10, 0, 0, 0, dt.itemsize) + body + _dtype_message(base)
    if dt.kind == 'b':
        base = struct.pack('<BBBBI', (1 << 4) | 0, 0x08, 0, 0, 1) + struct.pack('<HH', 0, 8)
        names = _pad8(b'FALSE\x00') + _pad8(b'TRUE\x00')
        return struct.pack('<BBBBI', (1 << 4) | 8, 2, 0, 0, 1) + base + names + b'\x00\x01'
    if dt.kind in 'iu':
        if dt.byteorder == '>':
            raise ValueError('big-endian integers are written little-endian; convert first')
        return (struct.pack('<BBBBI', (1 << 4) | 0, 0x08 if dt.kind == 'i' else 0, 0, 0, dt.itemsize)
                + struct.pack('<HH', 0, 8 * dt.itemsize))
    if dt.kind == 'f':
        if dt.itemsize == 4:
            props = struct.pack('<HHBBBBI', 0, 32, 23, 8, 0, 23, 127)
            return struct.pack('<BBBBI', (1 << 4) | 1, 0x20, 31, 0, 4) + props
        if dt.itemsize == 8:
            props = struct.pack('<HHBBBBI', 0, 64, 52, 11, 0, 52, 1023)
            return struct.pack('<BBBBI', (1 << 4) | 1, 0x20, 63, 0, 8) + props
    if dt.kind == 'S':
        return struct.pack('<BBBBI', (1 << 4) | 3, 0x01, 0, 0, dt.itemsize)
    raise TypeError(f'no HDF5 datatype for {dt} (encode unicode strings to bytes first)')


def _little(dt):
    """dt with every numeric field little-endian."""
    if dt.fields is not None:
        return np.dtype({'names': list(dt.names),
                         'formats': [_little(dt.fields[n][0]) for n in dt.names],
                         'offsets': [dt.fields[n][1] for n in dt.names],
                         'itemsize': dt.itemsize})
    if dt.subdtype is not None:
        return np.dtype((_little(dt.subdtype[0]), dt.subdtype[1]))
    return dt.newbyteorder('<') if dt.kind in 'iuf' else dt


class _Writer(object):
    def __init__(self):
        self.buf = bytearray(96)

    def alloc(self, data):
        addr = len(self.buf)
        self.buf += data
        return addr

    def reserve(self, n):
        return self.alloc(bytes(n))

    def put(self, addr, data):
        self.buf[addr:addr + len(data)] = data

    def object_header(self, messages, min_size=0):
        body = b''
        for mtype, data, flags in messages:
            data = _pad8(data)
            body += struct.pack('<HHB3x', mtype, len(data), flags) + data
        nmsg = len(messages)
        if len(body) < min_size:
            body += struct.pack('<HHB3x', 0, min_size - len(body) - 8, 0) + bytes(min_size - len(body) - 8)
            nmsg += 1
        return self.alloc(struct.pack('<BBHII4x', 1, 0, nmsg, 1, len(body)) + body)

    def dataset(self, arr):
        arr = np.asarray(arr)
        if arr.dtype.kind == 'U':
            raise TypeError('encode unicode strings to bytes before writing')
        dt = _little(arr.dtype)
        arr = np.array(arr, dtype=dt, order='C')          # keeps 0-d arrays 0-d
        raw = arr.tobytes()
        addr = self.alloc(raw) if raw else UNDEF
        space = struct.pack('<BBBB4x', 1, arr.ndim, 1, 0)
        space += b''.join(struct.pack('<Q', n) for n in arr.shape) * 2
        fill = struct.pack('<BBBBI', 2, 2, 2, 1, 0)
        layout = struct.pack('<BBQQ', 3, 1, addr, len(raw))
        return self.object_header([(0x01, space, 0), (0x03, _dtype_message(dt), 1),
                                   (0x05, fill, 1), (0x08, layout, 0)], min_size=256)

    def group(self, tree):
        """Write a group (children first); returns (header, btree, heap) addresses."""
        names = sorted(tree.keys())
        entries = []
        for name in names:
            v = tree[name]
            if isinstance(v, dict):
                hdr, bt, hp = self.group(v)
                entries.append((hdr, 1, struct.pack('<QQ', bt, hp)))
            else:
                entries.append((self.dataset(v), 0, bytes(16)))
        # local heap of link names: "" at offset 0, then the names, then one free block
        data = bytearray(8)
        offs = []
        for name in names:
            offs.append(len(data))
            data += _pad8(name.encode('ascii') + b'\x00')
        free = len(data)
        data += struct.pack('<QQ', 1, 16)
        heap = self.alloc(b'HEAP' + bytes(4) + struct.pack('<QQQ', len(data), free, len(self.buf) + 32))
        self.alloc(bytes(data))
        # symbol table nodes of <= 2 * LEAF_K entries, sorted by name
        level = []
        per = 2 * LEAF_K
        for i in range(0, len(names), per):
            chunk = list(range(i, min(i + per, len(names))))
            node = b'SNOD' + struct.pack('<BBH', 1, 0, len(chunk))
            for k in chunk:
                hdr, cache, scratch = entries[k]
                node += struct.pack('<QQI4x', offs[k], hdr, cache) + scratch
            node += bytes(8 + per * 40 - len(node))
            left = offs[i - 1] if i else 0
            level.append((self.alloc(node), left, offs[chunk[-1]]))
        btree = self.btree(level, 0)
        hdr = self.object_header([(0x11, struct.pack('<QQ', btree, heap), 0)])
        return hdr, btree, heap

    def btree(self, children, lev):
        """Group B-tree over children [(address, left key, right key)]; returns the root."""
        fan = 2 * INTERNAL_K
        size = 24 + (fan + 1) * 8 + fan * 8
        groups = [children[i:i + fan] for i in range(0, len(children), fan)] or [[]]
        addrs = [self.reserve(size) for _ in groups]
        parents = []
        for n, (addr, kids) in enumerate(zip(addrs, groups)):
            left = addrs[n - 1] if n else UNDEF
            right = addrs[n + 1] if n + 1 < len(addrs) else UNDEF
            node = b'TREE' + struct.pack('<BBHQQ', 0, lev, len(kids), left, right)
            node += struct.pack('<Q', kids[0][1] if kids else 0)
            for a, _, r in kids:
                node += struct.pack('<QQ', a, r)
            self.put(addr, node)
            parents.append((addr, kids[0][1] if kids else 0, kids[-1][2] if kids else 0))
        if len(parents) == 1:
            return parents[0][0]
        return self.btree(parents, lev + 1)


def write(path, tree):
    """Write a nested dict (name -> dict or numpy array) as an HDF5 file."""
    w = _Writer()
    hdr, bt, hp = w.group(tree)
    sb = SIGNATURE + bytes([0, 0, 0, 0, 0, 8, 8, 0]) + struct.pack('<HHI', LEAF_K, INTERNAL_K, 0)
    sb += struct.pack('<QQQQ', 0, UNDEF, len(w.buf), UNDEF)
    sb += struct.pack('<QQI4xQQ', 0, hdr, 1, bt, hp)
    w.put(0, sb)
    with open(path, 'wb') as fh:
        fh.write(bytes(w.buf))
