"""Reader for the JLD2 input files of the ladder run (the reference reads them with
``jldopen`` in src/IO.jl:67-87 and src/runHelpers.jl:89-111).

JLD2 files are HDF5 files behind a 512-byte user block.  This reader covers the
subset of HDF5 that JLD2 writes for plain arrays and numbers:

* superblock version 2 or 3 (8-byte offsets and lengths);
* version-2 object headers ("OHDR") with continuation blocks ("OCHK");
* groups as link messages in the object header (hard links);
* dataspaces (scalar and simple), fixed-point and IEEE float datatypes,
  compound datatypes (a committed ``{re, im}`` pair of float64 is read as
  complex128), shared (committed) datatype messages;
* compact and contiguous data layouts (no chunking, no filters).

Julia arrays are column-major and JLD2 writes the HDF5 dimensions reversed, so
an array comes back with Julia's shape, Fortran-ordered.  Anything outside the
subset raises ``ValueError`` naming what was found.
"""
from __future__ import annotations

import struct
from pathlib import Path

import numpy as np

_SIG_HDF5 = b"\x89HDF\r\n\x1a\n"
_MSG_DATASPACE, _MSG_LINKINFO, _MSG_DATATYPE, _MSG_LINK, _MSG_LAYOUT, _MSG_CONT = 1, 2, 3, 6, 8, 16


class JLD2File:
    """An open JLD2 file: ``keys()``, ``f[name]`` (numpy array or Python scalar)."""

    def __init__(self, path):
        self.path = Path(path)
        self.d = self.path.read_bytes()
        self.base = self._find_superblock()
        sb = self.d[self.base:self.base + 48]
        ver, so, sl = sb[8], sb[9], sb[10]
        if ver not in (2, 3) or so != 8 or sl != 8:
            raise ValueError(f"{self.path}: HDF5 superblock version {ver} with {so}/{sl}-byte offsets/lengths "
                             "is outside the JLD2 subset read here")
        base_addr, _ext, _eof, root = struct.unpack("<4Q", sb[12:44])
        if base_addr != self.base:
            raise ValueError(f"{self.path}: superblock base address {base_addr} != its position {self.base}")
        self._root = root
        self._links = dict(self._group_links(root))
        self._dtype_cache = {}

    # -- structure ---------------------------------------------------------------------------------
    def _find_superblock(self):
        for off in (0, 512, 1024, 2048, 4096):
            if self.d[off:off + 8] == _SIG_HDF5:
                return off
        raise ValueError(f"{self.path}: no HDF5 superblock at 0/512/1024/2048/4096")

    def _messages(self, addr):
        d, p = self.d, self.base + addr
        if d[p:p + 4] != b"OHDR":
            raise ValueError(f"{self.path}: object header at {addr} is not a version-2 header")
        flags = d[p + 5]
        q = p + 6
        if flags & 0x20:
            q += 16          # access/modification/change/birth times
        if flags & 0x10:
            q += 4           # attribute phase-change values
        nsz = 1 << (flags & 3)
        size = int.from_bytes(d[q:q + nsz], "little")
        q += nsz
        blocks, out = [(q, q + size)], []
        while blocks:
            s, e = blocks.pop(0)
            while s + 4 <= e - 4:      # the last 4 bytes of a block are its checksum
                mtype, msize, mflags = d[s], struct.unpack("<H", d[s + 1:s + 3])[0], d[s + 3]
                s += 4
                if flags & 0x04:
                    s += 2          # creation order
                body = d[s:s + msize]
                s += msize
                if mtype == _MSG_CONT:
                    caddr, clen = struct.unpack("<QQ", body[:16])
                    cp = self.base + caddr
                    if d[cp:cp + 4] != b"OCHK":
                        raise ValueError(f"{self.path}: continuation block at {caddr} has no OCHK signature")
                    blocks.append((cp + 4, cp + clen))
                out.append((mtype, mflags, body))
        return out

    def _group_links(self, addr):
        for mtype, _f, b in self._messages(addr):
            if mtype == _MSG_LINKINFO:
                fheap = struct.unpack("<Q", b[2 + (8 if b[1] & 1 else 0):][:8])[0]
                if fheap != 0xFFFFFFFFFFFFFFFF:
                    raise ValueError(f"{self.path}: group at {addr} keeps its links in a fractal heap")
            if mtype != _MSG_LINK:
                continue
            fl, q = b[1], 2
            ltype = 0
            if fl & 0x08:
                ltype = b[q]
                q += 1
            if fl & 0x04:
                q += 8
            if fl & 0x10:
                q += 1
            nl = 1 << (fl & 3)
            n = int.from_bytes(b[q:q + nl], "little")
            q += nl
            name = b[q:q + n].decode("utf-8")
            q += n
            if ltype == 0:
                yield name, struct.unpack("<Q", b[q:q + 8])[0]

    # -- datatypes ---------------------------------------------------------------------------------
    def _datatype(self, b):
        """numpy dtype of a datatype message body; returns (dtype, bytes consumed)."""
        cls, ver = b[0] & 0x0F, b[0] >> 4
        bits = b[1] | (b[2] << 8) | (b[3] << 16)
        size = struct.unpack("<I", b[4:8])[0]
        if cls == 0:                                     # fixed point
            if bits & 1:
                raise ValueError("big-endian integers are not read")
            return np.dtype(("<i" if bits & 8 else "<u") + str(size)), 12
        if cls == 1:                                     # IEEE float
            if bits & 1 or size not in (2, 4, 8):
                raise ValueError(f"float datatype of size {size} / byte order {bits & 1} is not read")
            return np.dtype("<f" + str(size)), 20
        if cls == 6:                                     # compound
            nmem = bits & 0xFFFF
            q, names, fmts, offs = 8, [], [], []
            osz = 1 if size < 256 else 2 if size < 65536 else 4 if size < 2 ** 32 else 8
            for _ in range(nmem):
                end = b.index(b"\x00", q)
                name = b[q:end].decode("utf-8")
                q = end + 1
                if ver < 3:
                    q = 8 + ((q - 8 + 7) // 8) * 8       # names padded to 8 bytes
                    off = struct.unpack("<I", b[q:q + 4])[0]
                    q += 4
                    if ver == 1:
                        q += 28                          # dimensionality, permutation, reserved, dims
                else:
                    off = int.from_bytes(b[q:q + osz], "little")
                    q += osz
                dt, used = self._datatype(b[q:])
                q += used
                names.append(name)
                fmts.append(dt)
                offs.append(off)
            if names == ["re", "im"] and fmts[0] == fmts[1] and offs == [0, fmts[0].itemsize] and fmts[0].kind == "f":
                return np.dtype("<c" + str(2 * fmts[0].itemsize)), q
            return np.dtype({"names": names, "formats": fmts, "offsets": offs, "itemsize": size}), q
        raise ValueError(f"datatype class {cls} is outside the JLD2 subset read here")

    def _message_dtype(self, mflags, body):
        if mflags & 0x02:                                # shared: a committed datatype
            ver, typ = body[0], body[1]
            addr = struct.unpack("<Q", body[2:10] if ver == 3 else body[8:16])[0]
            if ver == 3 and typ != 2:
                raise ValueError(f"shared datatype message of type {typ} is not read")
            if addr not in self._dtype_cache:
                msgs = [m for m in self._messages(addr) if m[0] == _MSG_DATATYPE]
                if not msgs:
                    raise ValueError(f"committed datatype at {addr} has no datatype message")
                self._dtype_cache[addr] = self._datatype(msgs[0][2])[0]
            return self._dtype_cache[addr]
        return self._datatype(body)[0]

    # -- datasets ----------------------------------------------------------------------------------
    def keys(self):
        return [k for k in self._links if k != "_types"]

    def __contains__(self, name):
        return name in self._links

    def __getitem__(self, name):
        addr = self._links[name]
        shape = dtype = raw = None
        for mtype, mflags, b in self._messages(addr):
            if mtype == _MSG_DATASPACE:
                ver, nd, fl = b[0], b[1], b[2]
                if ver == 1:
                    shape = struct.unpack(f"<{nd}Q", b[8:8 + 8 * nd])
                else:
                    if b[3] == 2:
                        raise ValueError(f"{name}: null dataspace")
                    shape = struct.unpack(f"<{nd}Q", b[4:4 + 8 * nd])
            elif mtype == _MSG_DATATYPE:
                dtype = self._message_dtype(mflags, b)
            elif mtype == _MSG_LAYOUT:
                ver, cls = b[0], b[1]
                if ver not in (3, 4):
                    raise ValueError(f"{name}: layout message version {ver}")
                if cls == 0:
                    n = struct.unpack("<H", b[2:4])[0]
                    raw = b[4:4 + n]
                elif cls == 1:
                    a, n = struct.unpack("<QQ", b[2:18])
                    raw = self.d[self.base + a:self.base + a + n]
                else:
                    raise ValueError(f"{name}: chunked or virtual layout (class {cls}) is not read")
            elif mtype == 11:
                raise ValueError(f"{name}: filtered (compressed) data is not read")
        if shape is None or dtype is None or raw is None:
            raise ValueError(f"{name}: dataset without dataspace, datatype or layout")
        n = int(np.prod(shape)) if shape else 1
        arr = np.frombuffer(raw, dtype=dtype, count=n)
        if not shape:
            v = arr[0]
            return v.item() if dtype.kind in "iufc" else v
        # HDF5 dims are Julia's reversed; the bytes are Julia's column-major memory
        return np.asfortranarray(arr.reshape(shape).T)


def jldopen(path) -> JLD2File:
    return JLD2File(path)
