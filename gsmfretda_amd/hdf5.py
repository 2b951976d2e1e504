"""Burst data in gSMFRETda's HDF5 schema, without libhdf5.

The reference reads its input with HighFive (src/loadHdf5.cpp:37-74):
`/sub_bursts_l/{start, stop, istart, istop}` (int64 per burst), `times`
(int64 macrotime ticks per photon), `mask_ad`, `mask_dd` (uint8 per photon),
`T_burst_duration`, `SgDivSr` (float per burst) and the scalars `clk_p`,
`bg_ad_rate`, `bg_dd_rate`; photon counts are istop - istart + 1
(src/loadHdf5.cpp:101-108).  Neither libhdf5 nor h5py is available to this
package, so this module carries a reader for the parts of the HDF5 file
format such files use, written from the format specification:

* superblock versions 0-3; object header versions 1 and 2 with continuation
  blocks;
* groups as symbol tables (version-1 B-tree, local heap, symbol nodes) or as
  compact link messages;
* datasets with simple or scalar dataspaces, fixed-point (8-64 bit) and
  IEEE floating-point (16-64 bit) datatypes of either byte order, compact,
  contiguous or chunked (version-1 B-tree) storage, and the deflate and
  shuffle filters;

and a minimal writer (superblock 0, symbol-table groups, contiguous little
endian datasets) for exporting bursts in the same schema.
"""
import struct
import zlib

import numpy as np

SIGNATURE = b"\x89HDF\r\n\x1a\n"
UNDEF = 0xFFFFFFFFFFFFFFFF


class H5Error(RuntimeError):
    pass


class _Dataset:
    def __init__(self, shape, dtype, layout, filters, fill):
        self.shape, self.dtype, self.layout, self.filters, self.fill = (
            shape, dtype, layout, filters, fill)


class H5File:
    """Read-only access to the datasets of an HDF5 file by path."""

    def __init__(self, path):
        with open(path, "rb") as f:
            self.buf = f.read()
        self.base = self._find_superblock()
        self._parse_superblock()

    # -- low level -------------------------------------------------------------
    def _find_superblock(self):
        off = 0
        while off + 8 <= len(self.buf):
            if self.buf[off:off + 8] == SIGNATURE:
                return off
            off = 512 if off == 0 else off * 2
        raise H5Error("not an HDF5 file (no superblock signature)")

    def _u(self, off, size):
        return int.from_bytes(self.buf[off:off + size], "little")

    def _addr(self, off):
        return self._u(off, self.so)

    def _len(self, off):
        return self._u(off, self.sl)

    def _parse_superblock(self):
        b = self.base
        ver = self.buf[b + 8]
        if ver in (0, 1):
            self.so, self.sl = self.buf[b + 13], self.buf[b + 14]
            p = b + 24 + (4 if ver == 1 else 0)
            p += 4 * self.so  # base, free space, end of file, driver info addresses
            # root group symbol table entry
            self.root = self._addr(p + self.so)
            self.root_scratch = (self._u(p + 2 * self.so, 4), p + 2 * self.so + 8)
        elif ver in (2, 3):
            self.so, self.sl = self.buf[b + 9], self.buf[b + 10]
            p = b + 12 + 3 * self.so
            self.root = self._addr(p)
            self.root_scratch = (0, None)
        else:
            raise H5Error(f"superblock version {ver} not supported")

    # -- object headers ----------------------------------------------------------
    def _messages(self, addr):
        """(type, data offset, size) of every message of the object header at addr."""
        if addr >= len(self.buf):
            raise H5Error(f"object header address {addr} outside the file")
        out = []
        if self.buf[addr:addr + 4] == b"OHDR":
            flags = self.buf[addr + 5]
            p = addr + 6
            if flags & 0x20:
                p += 16
            if flags & 0x10:
                p += 4
            nsz = 1 << (flags & 3)
            size = self._u(p, nsz)
            p += nsz
            blocks = [(p, p + size, True)]
            while blocks:
                start, end, first = blocks.pop(0)
                q = start if first else start + 4  # "OCHK"
                while q + 4 <= end - (4 if not first else 4):
                    mtype = self.buf[q]
                    msize = self._u(q + 1, 2)
                    mflags = self.buf[q + 3]
                    q += 4 + (2 if flags & 0x04 else 0)
                    if mtype == 0x10:
                        blocks.append((self._addr(q), self._addr(q) + self._len(q + self.so), False))
                    elif mtype != 0:
                        out.append((mtype, q, msize, mflags))
                    q += msize
            return out
        if self.buf[addr] != 1:
            raise H5Error(f"object header version {self.buf[addr]} at {addr} not supported")
        nmsg = self._u(addr + 2, 2)
        size = self._u(addr + 8, 4)
        blocks = [(addr + 16, addr + 16 + size)]
        while blocks and len(out) < nmsg + 64:
            start, end = blocks.pop(0)
            q = start
            while q + 8 <= end:
                mtype = self._u(q, 2)
                msize = self._u(q + 2, 2)
                mflags = self.buf[q + 4]
                q += 8
                if mtype == 0x10:
                    blocks.append((self._addr(q), self._addr(q) + self._len(q + self.so)))
                elif mtype != 0:
                    out.append((mtype, q, msize, mflags))
                q += msize
        return out

    # -- groups --------------------------------------------------------------------
    def _group_links(self, addr, scratch=None):
        links = {}
        msgs = self._messages(addr)
        for mtype, p, _, _ in msgs:
            if mtype == 0x11:  # symbol table: B-tree and local heap
                self._symbol_table(self._addr(p), self._addr(p + self.so), links)
            elif mtype == 0x06:  # link (compact storage)
                name, target = self._link(p)
                if target is not None:
                    links[name] = target
            elif mtype == 0x02:  # link info: dense storage
                heap = self._addr(p + 2 + (8 if self.buf[p + 1] & 1 else 0))
                if heap != UNDEF:
                    raise H5Error("groups in dense (fractal heap) storage are not supported")
        if not links and scratch and scratch[0] == 1:
            self._symbol_table(self._addr(scratch[1]), self._addr(scratch[1] + self.so), links)
        return links

    def _link(self, p):
        flags = self.buf[p + 1]
        q = p + 2
        ltype = 0
        if flags & 0x08:
            ltype = self.buf[q]
            q += 1
        if flags & 0x04:
            q += 8
        if flags & 0x10:
            q += 1
        nlen_size = 1 << (flags & 3)
        nlen = self._u(q, nlen_size)
        q += nlen_size
        name = self.buf[q:q + nlen].decode("utf-8")
        q += nlen
        return name, (self._addr(q) if ltype == 0 else None)

    def _symbol_table(self, btree, heap, links):
        if self.buf[heap:heap + 4] != b"HEAP":
            raise H5Error("bad local heap")
        data = self._addr(heap + 8 + 2 * self.sl)

        def name_at(off):
            s = data + off
            e = self.buf.index(b"\0", s)
            return self.buf[s:e].decode("utf-8")

        def walk(node):
            if self.buf[node:node + 4] == b"SNOD":
                n = self._u(node + 6, 2)
                q = node + 8
                for _ in range(n):
                    if self._u(q + 2 * self.so, 4) != 2:  # cache type 2: a soft link
                        links[name_at(self._addr(q))] = self._addr(q + self.so)
                    q += 2 * self.so + 24
                return
            if self.buf[node:node + 4] != b"TREE" or self.buf[node + 4] != 0:
                raise H5Error("bad group B-tree node")
            used = self._u(node + 6, 2)
            q = node + 8 + 2 * self.so + self.sl  # first key
            for _ in range(used):
                walk(self._addr(q))
                q += self.so + self.sl
        walk(btree)

    # -- datasets ------------------------------------------------------------------
    def _dataset(self, addr):
        shape, dtype, layout, filters, fill = None, None, None, [], None
        for mtype, p, size, _ in self._messages(addr):
            if mtype == 0x01:
                shape = self._dataspace(p)
            elif mtype == 0x03:
                dtype = self._datatype(p)
            elif mtype == 0x08:
                layout = self._layout(p, size)
            elif mtype == 0x0B:
                filters = self._filters(p)
        if shape is None or dtype is None or layout is None:
            raise H5Error(f"object at {addr} is not a dataset")
        return _Dataset(shape, dtype, layout, filters, fill)

    def _dataspace(self, p):
        ver, rank, flags = self.buf[p], self.buf[p + 1], self.buf[p + 2]
        if ver == 1:
            q = p + 8
        elif ver == 2:
            if self.buf[p + 3] == 2:
                return (0,)  # null dataspace
            q = p + 4
        else:
            raise H5Error(f"dataspace version {ver}")
        return tuple(self._len(q + i * self.sl) for i in range(rank))

    def _datatype(self, p):
        cls = self.buf[p] & 0x0F
        bits = self.buf[p + 1]
        size = self._u(p + 4, 4)
        order = ">" if bits & 1 else "<"
        if cls == 0:
            signed = bool(bits & 0x08)
            if size not in (1, 2, 4, 8):
                raise H5Error(f"integer size {size}")
            return np.dtype(f"{order}{'i' if signed else 'u'}{size}")
        if cls == 1:
            if size not in (2, 4, 8):
                raise H5Error(f"float size {size}")
            return np.dtype(f"{order}f{size}")
        raise H5Error(f"datatype class {cls} not supported")

    def _layout(self, p, size):
        ver = self.buf[p]
        if ver in (1, 2):
            rank, cls = self.buf[p + 1], self.buf[p + 2]
            q = p + 8
            addr = None
            if cls != 0:
                addr = self._addr(q)
                q += self.so
            dims = tuple(self._u(q + 4 * i, 4) for i in range(rank))
            if cls == 0:
                n = self._u(q + 4 * rank, 4)
                return ("compact", self.buf[q + 4 * rank + 4:q + 4 * rank + 4 + n])
            if cls == 1:
                return ("contiguous", addr, None)
            return ("chunked", addr, dims[:-1])
        if ver == 3:
            cls = self.buf[p + 1]
            if cls == 0:
                n = self._u(p + 2, 2)
                return ("compact", self.buf[p + 4:p + 4 + n])
            if cls == 1:
                return ("contiguous", self._addr(p + 2), self._len(p + 2 + self.so))
            if cls == 2:
                rank = self.buf[p + 2]
                addr = self._addr(p + 3)
                q = p + 3 + self.so
                dims = tuple(self._u(q + 4 * i, 4) for i in range(rank))
                return ("chunked", addr, dims[:-1])
        raise H5Error(f"data layout version {ver} not supported")

    def _filters(self, p):
        ver, n = self.buf[p], self.buf[p + 1]
        q = p + (8 if ver == 1 else 2)
        out = []
        for _ in range(n):
            fid = self._u(q, 2)
            if ver == 1 or fid >= 256:
                nlen = self._u(q + 2, 2)
                flags, nval = self._u(q + 4, 2), self._u(q + 6, 2)
                q += 8
            else:
                nlen = 0
                flags, nval = self._u(q + 2, 2), self._u(q + 4, 2)
                q += 6
            if ver == 1:
                nlen = (nlen + 7) // 8 * 8
            q += nlen
            vals = [self._u(q + 4 * i, 4) for i in range(nval)]
            q += 4 * nval
            if ver == 1 and nval % 2:
                q += 4
            out.append((fid, flags, vals))
        return out

    def _unfilter(self, raw, filters, mask, itemsize):
        for i, (fid, _, vals) in reversed(list(enumerate(filters))):
            if mask & (1 << i):
                continue
            if fid == 1:
                raw = zlib.decompress(raw)
            elif fid == 2:
                es = vals[0] if vals else itemsize
                a = np.frombuffer(raw, np.uint8)
                n = len(a) // es
                raw = a[:n * es].reshape(es, n).T.tobytes() + a[n * es:].tobytes()
            elif fid == 3:
                raw = raw[:-4]  # Fletcher32 checksum, not verified
            else:
                raise H5Error(f"filter {fid} not supported")
        return raw

    def _read(self, ds):
        count = int(np.prod(ds.shape)) if ds.shape else 1
        kind = ds.layout[0]
        if kind == "compact":
            return np.frombuffer(ds.layout[1], ds.dtype, count).reshape(ds.shape)
        if kind == "contiguous":
            addr = ds.layout[1]
            if addr == UNDEF:
                return np.zeros(ds.shape, ds.dtype)
            n = count * ds.dtype.itemsize
            if addr + n > len(self.buf):
                raise H5Error("dataset extends past the end of the file")
            return np.frombuffer(self.buf[addr:addr + n], ds.dtype, count).reshape(ds.shape)
        out = np.zeros(ds.shape, ds.dtype)
        btree, chunk = ds.layout[1], ds.layout[2]
        rank = len(chunk)
        if btree == UNDEF:
            return out

        def walk(node):
            if self.buf[node:node + 4] != b"TREE" or self.buf[node + 4] != 1:
                raise H5Error("bad chunk B-tree node")
            level, used = self.buf[node + 5], self._u(node + 6, 2)
            q = node + 8 + 2 * self.so
            ksize = 8 + 8 * (rank + 1)
            for _ in range(used):
                csize, mask = self._u(q, 4), self._u(q + 4, 4)
                offs = [self._u(q + 8 + 8 * d, 8) for d in range(rank)]
                child = self._addr(q + ksize)
                if level > 0:
                    walk(child)
                else:
                    raw = self._unfilter(self.buf[child:child + csize], ds.filters, mask,
                                         ds.dtype.itemsize)
                    block = np.frombuffer(raw, ds.dtype, int(np.prod(chunk))).reshape(chunk)
                    dst = tuple(slice(o, min(o + c, s)) for o, c, s in zip(offs, chunk, ds.shape))
                    src = tuple(slice(0, sl.stop - sl.start) for sl in dst)
                    out[dst] = block[src]
                q += ksize + self.so
        walk(btree)
        return out

    # -- public --------------------------------------------------------------------
    def _resolve(self, path):
        addr, scratch = self.root, self.root_scratch
        for part in [x for x in path.split("/") if x]:
            links = self._group_links(addr, scratch)
            if part not in links:
                raise KeyError(path)
            addr, scratch = links[part], None
        return addr

    def keys(self, group="/"):
        return sorted(self._group_links(self._resolve(group),
                                        self.root_scratch if group in ("/", "") else None))

    def read(self, path):
        """The dataset at `path` as a numpy array (native byte order); a
        scalar dataspace gives a 0-d array."""
        ds = self._dataset(self._resolve(path))
        a = self._read(ds)
        return a.astype(a.dtype.newbyteorder("="), copy=False)


# ---- minimal writer -------------------------------------------------------------

def write(path, datasets):
    """Writes {"/group/name": numpy array or scalar} as an HDF5 file: superblock
    version 0, 8-byte offsets and lengths, symbol-table groups, contiguous
    little-endian datasets (integers and floats)."""
    tree = {}
    for full, val in datasets.items():
        parts = [x for x in full.split("/") if x]
        node = tree
        for g in parts[:-1]:
            node = node.setdefault(g, {})
        node[parts[-1]] = np.asarray(val)
    out = bytearray()

    def alloc(n):
        off = len(out)
        out.extend(b"\0" * ((n + 7) // 8 * 8))
        return off

    def put(off, data):
        out[off:off + len(data)] = data

    sb = alloc(96)  # superblock v0 with the root symbol table entry

    def dtype_msg(a):
        dt = a.dtype.newbyteorder("<")
        if dt.kind in "iu":
            bits = 0x08 if dt.kind == "i" else 0
            body = struct.pack("<BBBBI", 0x10 | 0, bits, 0, 0, dt.itemsize)
            body += struct.pack("<HH", 0, 8 * dt.itemsize)
        elif dt.kind == "f":
            sizes = {2: (10, 5, 0, 10, 15), 4: (23, 8, 0, 23, 127), 8: (52, 11, 0, 52, 1023)}
            eloc, esize, mloc, msize, bias = sizes[dt.itemsize]
            body = struct.pack("<BBBBI", 0x10 | 1, 0x20, 8 * dt.itemsize - 1, 0, dt.itemsize)
            body += struct.pack("<HHBBBBI", 0, 8 * dt.itemsize, eloc, esize, mloc, msize, bias)
        else:
            raise H5Error(f"cannot write {a.dtype}")
        return body

    def message(mtype, body):
        body = body + b"\0" * ((-len(body)) % 8)
        return struct.pack("<HHB3x", mtype, len(body), 0) + body

    def dataset(a):
        a = np.asarray(a, a.dtype.newbyteorder("<"))  # keeps 0-d arrays (scalars) 0-d
        data = alloc(a.nbytes)
        put(data, a.tobytes())
        dims = a.shape
        space = struct.pack("<BBBB4x", 1, len(dims), 0, 0) + b"".join(
            struct.pack("<Q", d) for d in dims)
        layout = struct.pack("<BB", 3, 1) + struct.pack("<QQ", data, a.nbytes)
        msgs = message(0x01, space) + message(0x03, dtype_msg(a)) + message(0x08, layout)
        hdr = alloc(16 + len(msgs))
        put(hdr, struct.pack("<BBHII4x", 1, 0, 3, 1, len(msgs)) + msgs)
        return hdr

    def group(node):
        names = sorted(node)
        children = {k: (group(v)[0] if isinstance(v, dict) else dataset(v))
                    for k, v in node.items()}
        heap_data = bytearray(b"\0" * 8)  # offset 0: the empty name
        offs = {}
        for k in names:
            offs[k] = len(heap_data)
            heap_data += k.encode() + b"\0"
            heap_data += b"\0" * ((-len(heap_data)) % 8)
        hd = alloc(len(heap_data))
        put(hd, bytes(heap_data))
        heap = alloc(32)
        put(heap, b"HEAP" + struct.pack("<B3xQQQ", 0, len(heap_data), UNDEF, hd))
        snod = alloc(8 + 40 * max(len(names), 1))
        body = b"SNOD" + struct.pack("<BBH", 1, 0, len(names))
        for k in names:
            body += struct.pack("<QQII16x", offs[k], children[k], 0, 0)
        put(snod, body)
        btree = alloc(24 + 8 * 3)
        last = offs[names[-1]] if names else 0
        put(btree, b"TREE" + struct.pack("<BBHQQ", 0, 0, 1, UNDEF, UNDEF)
            + struct.pack("<QQQ", 0, snod, last))
        msgs = message(0x11, struct.pack("<QQ", btree, heap))
        hdr = alloc(16 + len(msgs))
        put(hdr, struct.pack("<BBHII4x", 1, 0, 1, 1, len(msgs)) + msgs)
        return hdr, btree, heap

    root, rbtree, rheap = group(tree)
    body = SIGNATURE + bytes([0, 0, 0, 0, 0, 8, 8, 0]) + struct.pack("<HHI", 4, 16, 0)
    body += struct.pack("<QQQQ", 0, UNDEF, len(out), UNDEF)
    body += struct.pack("<QQII", 0, root, 1, 0) + struct.pack("<QQ", rbtree, rheap)
    put(sb, body)
    put(sb + 40, struct.pack("<Q", len(out)))  # end of file address
    with open(path, "wb") as f:
        f.write(bytes(out))


# ---- the schema -----------------------------------------------------------------

def read_bursts(path):
    """(BurstData, SgDivSr) from a file in the reference's schema
    (src/loadHdf5.cpp:37-74; photon counts istop - istart + 1,
    src/loadHdf5.cpp:101-108)."""
    from .synth import BurstData
    f = H5File(path)
    g = "/sub_bursts_l/"
    start, stop = f.read(g + "start").astype(np.int64), f.read(g + "stop").astype(np.int64)
    istart, istop = f.read(g + "istart").astype(np.int64), f.read(g + "istop").astype(np.int64)
    times = f.read("times").astype(np.int64)
    mask_ad, mask_dd = f.read("mask_ad").astype(np.uint8), f.read("mask_dd").astype(np.uint8)
    T = f.read("T_burst_duration").astype(np.float32)
    sg = f.read("SgDivSr").astype(np.float32)
    clk_p = float(np.float32(f.read("clk_p")))
    bg_ad, bg_dd = float(np.float32(f.read("bg_ad_rate"))), float(np.float32(f.read("bg_dd_rate")))
    n = len(start)
    if not (len(stop) == len(istart) == len(istop) == len(T) == len(sg) == n):
        raise H5Error("per-burst datasets differ in length (src/main.cpp:121-123)")
    if not (len(mask_ad) == len(mask_dd) == len(times)):
        raise H5Error("per-photon datasets differ in length (src/main.cpp:121)")
    phcount = (istop - istart + 1).astype(np.int32)
    d = BurstData(start=start, stop=stop, istart=istart, phcount=phcount, times=times,
                  mask_ad=mask_ad, mask_dd=mask_dd, T=T, clk_p=clk_p, bg_ad=bg_ad, bg_dd=bg_dd)
    return d, sg


def write_bursts(path, d, SgDivSr):
    """Writes bursts in the reference's schema (the datasets read_bursts reads)."""
    g = "/sub_bursts_l/"
    write(path, {
        g + "start": np.asarray(d.start, np.int64), g + "stop": np.asarray(d.stop, np.int64),
        g + "istart": np.asarray(d.istart, np.int64), g + "istop": np.asarray(d.istop, np.int64),
        "times": np.asarray(d.times, np.int64),
        "mask_ad": np.asarray(d.mask_ad, np.uint8), "mask_dd": np.asarray(d.mask_dd, np.uint8),
        "T_burst_duration": np.asarray(d.T, np.float32),
        "SgDivSr": np.asarray(SgDivSr, np.float32),
        "clk_p": np.float32(d.clk_p), "bg_ad_rate": np.float32(d.bg_ad),
        "bg_dd_rate": np.float32(d.bg_dd),
    })
