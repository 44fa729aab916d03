"""Read-only HDF5 datasets for catalog ingestion, without h5py.

The reference reads box catalogs in HDF5 with h5py (``h5_reader``,
``catalog_processor.py:16-88``): named one-dimensional datasets, one per
column, the row count from the first requested column present.  h5py is not
part of this package, so this module implements the subset of the HDF5 file
format (HDF Group, "HDF5 File Format Specification", version 3.0) that such
files use:

* superblock versions 0-3;
* object headers version 1 and 2 (with continuation blocks);
* groups: symbol tables (version 1 B-tree of symbol nodes + local heap),
  compact link messages, and dense link storage (the link messages of the
  fractal heap, read from its direct blocks in order through the doubling
  table; groups that had links deleted may leave stale messages there);
* datasets: dataspace v1/v2, fixed-point and IEEE floating-point datatypes of
  either byte order, layouts version 3 (compact, contiguous, chunked with a
  version 1 B-tree) and version 4 (compact, contiguous, chunked with the
  single-chunk, implicit and fixed-array indexes);
* filters: deflate (zlib), shuffle, fletcher32 (checksum stripped).

Anything else (extensible-array or version 2 B-tree chunk indexes, variable
length or compound types, other filters, external storage) raises
``NotImplementedError`` naming the feature.  Checksums of version 2
structures are not verified.
"""
import zlib

import numpy as np

SIGNATURE = b"\x89HDF\r\n\x1a\n"
UNDEF = None  # undefined address


class File:
    """An HDF5 file opened for reading (the whole file in memory, or a
    numpy memory map of it)."""

    def __init__(self, path):
        self.path = path
        self.buf = np.memmap(path, dtype=np.uint8, mode="r")
        self.b = memoryview(self.buf)
        base = self._find_superblock()
        self._superblock(base)

    # ------------------------------------------------------------ low level
    def _bytes(self, off, size):
        return bytes(self.b[off:off + size])

    def _u(self, off, size):
        return int.from_bytes(self.b[off:off + size], "little")

    def _addr(self, off):
        v = self._u(off, self.so)
        return UNDEF if v == (1 << (8 * self.so)) - 1 else v + self.base

    def _find_superblock(self):
        off = 0
        n = len(self.buf)
        while off + 8 <= n:
            if self._bytes(off, 8) == SIGNATURE:
                return off
            off = 512 if off == 0 else off * 2
        raise ValueError(f"{self.path}: not an HDF5 file (no superblock signature)")

    def _superblock(self, off):
        ver = self.b[off + 8]
        if ver in (0, 1):
            self.so, self.sl = self.b[off + 13], self.b[off + 14]
            p = off + 24 + (4 if ver == 1 else 0)
            self.base = self._u(p, self.so)
            p += 4 * self.so  # base, free space, end of file, driver info
            # root group symbol table entry: name offset, object header address
            self.root = self._u(p + self.so, self.so) + self.base
        elif ver in (2, 3):
            self.so, self.sl = self.b[off + 9], self.b[off + 10]
            p = off + 12
            self.base = self._u(p, self.so)
            self.root = self._u(p + 3 * self.so, self.so) + self.base
        else:
            raise NotImplementedError(f"HDF5 superblock version {ver}")

    # ------------------------------------------------------ object headers
    def _messages(self, addr):
        """List of (type, data offset, size) of the object header at addr."""
        out = []
        if self._bytes(addr, 4) == b"OHDR":
            ver, flags = self.b[addr + 4], self.b[addr + 5]
            if ver != 2:
                raise NotImplementedError(f"object header version {ver}")
            p = addr + 6
            if flags & 0x20:
                p += 16
            if flags & 0x10:
                p += 4
            csize = 1 << (flags & 3)
            size0 = self._u(p, csize)
            p += csize
            blocks = [(p, size0)]
            track = bool(flags & 0x04)
            while blocks:
                start, size = blocks.pop(0)
                q, end = start, start + size
                while q + 4 <= end:
                    mtype, msize, mflags = self.b[q], self._u(q + 1, 2), self.b[q + 3]
                    q += 4 + (2 if track else 0)
                    if mtype == 0x10:  # continuation: OCHK block
                        caddr, clen = self._addr(q), self._u(q + self.so, self.sl)
                        if self._bytes(caddr, 4) != b"OCHK":
                            raise ValueError(f"{self.path}: bad object header continuation")
                        blocks.append((caddr + 4, clen - 8))
                    elif mtype != 0:
                        out.append((mtype, q, msize, mflags))
                    q += msize
            return out
        ver = self.b[addr]
        if ver != 1:
            raise NotImplementedError(f"object header version {ver}")
        hsize = self._u(addr + 8, 4)
        blocks = [(addr + 16, hsize)]
        while blocks:
            start, size = blocks.pop(0)
            q, end = start, start + size
            while q + 8 <= end:
                mtype, msize, mflags = self._u(q, 2), self._u(q + 2, 2), self.b[q + 4]
                q += 8
                if mtype == 0x10:  # continuation: a block of further messages
                    blocks.append((self._addr(q), self._u(q + self.so, self.sl)))
                elif mtype != 0:
                    out.append((mtype, q, msize, mflags))
                q += msize
        return out

    # --------------------------------------------------------------- groups
    def _links(self, addr):
        """{name: object header address} of the group at addr."""
        msgs = self._messages(addr)
        links = {}
        for mtype, p, size, _ in msgs:
            if mtype == 0x11:  # symbol table: v1 B-tree + local heap
                btree, heap = self._addr(p), self._addr(p + self.so)
                links.update(self._symbol_table(btree, heap))
            elif mtype == 0x06:
                name, target = self._link_message(p)
                if target is not None:
                    links[name] = target
            elif mtype == 0x02:  # link info: dense storage in a fractal heap
                flags = self.b[p + 1]
                q = p + 2 + (8 if flags & 1 else 0)
                fheap = self._addr(q)
                if fheap is not UNDEF:
                    for obj in self._fractal_heap_objects(fheap):
                        name, target = self._link_message_bytes(obj)
                        if target is not None:
                            links[name] = target
        return links

    def _link_message_bytes(self, data):
        return _LinkParser(self, data).parse()

    def _link_message(self, p):
        return _LinkParser(self, None, p).parse()

    def _local_heap(self, addr):
        if self._bytes(addr, 4) != b"HEAP":
            raise ValueError(f"{self.path}: bad local heap")
        return self._addr(addr + 8 + 2 * self.sl)

    def _symbol_table(self, btree, heap):
        data = self._local_heap(heap)
        links = {}
        stack = [btree]
        while stack:
            node = stack.pop()
            if self._bytes(node, 4) != b"TREE":
                raise ValueError(f"{self.path}: bad group B-tree node")
            level, used = self.b[node + 5], self._u(node + 6, 2)
            p = node + 8 + 2 * self.so
            for i in range(used):
                child = self._addr(p + self.sl + i * (self.sl + self.so))
                if level > 0:
                    stack.append(child)
                else:
                    links.update(self._symbol_node(child, data))
        return links

    def _symbol_node(self, addr, heap_data):
        if self._bytes(addr, 4) != b"SNOD":
            raise ValueError(f"{self.path}: bad symbol table node")
        n = self._u(addr + 6, 2)
        out = {}
        esize = 2 * self.so + 24
        for i in range(n):
            e = addr + 8 + i * esize
            name_off = self._u(e, self.so)
            obj = self._addr(e + self.so)
            s = heap_data + name_off
            end = s
            while self.b[end] != 0:
                end += 1
            out[self._bytes(s, end - s).decode()] = obj
        return out

    def _fractal_heap_objects(self, addr):
        """Every managed object of the fractal heap at addr (the link
        messages of a dense group), read block by block."""
        if self._bytes(addr, 4) != b"FRHP":
            raise ValueError(f"{self.path}: bad fractal heap header")
        p = addr + 5
        id_len, io_len = self._u(p, 2), self._u(p + 2, 2)
        flags = self.b[p + 4]
        p += 5 + 4  # max managed object size
        p += self.sl + self.so  # next huge id, huge B-tree
        p += self.sl + self.so  # free space amount, manager
        p += 4 * self.sl  # managed space, allocated, iterator offset, managed object count
        p += 2 * self.sl  # huge objects size / count
        p += 2 * self.sl  # tiny objects size / count
        width = self._u(p, 2)
        start_block = self._u(p + 2, self.sl)
        max_direct = self._u(p + 2 + self.sl, self.sl)
        max_heap_bits = self._u(p + 2 + 2 * self.sl, 2)
        start_rows = self._u(p + 4 + 2 * self.sl, 2)
        root = self._addr(p + 6 + 2 * self.sl)
        cur_rows = self._u(p + 6 + 2 * self.sl + self.so, 2)
        if io_len:
            raise NotImplementedError("fractal heap with I/O filters")
        checksum_direct = bool(flags & 2)
        off_bytes = (max_heap_bits + 7) // 8
        geom = dict(width=width, start=start_block, max_direct=max_direct, off_bytes=off_bytes,
                    cs=checksum_direct)
        blocks = []
        if root is not UNDEF:
            if cur_rows == 0:
                blocks.append((root, start_block))
            else:
                self._indirect_blocks(root, cur_rows, geom, blocks)
        objs = []
        for baddr, bsize in blocks:
            objs.extend(self._direct_block_objects(baddr, bsize, geom))
        return objs

    def _indirect_blocks(self, addr, nrows, geom, out):
        if self._bytes(addr, 4) != b"FHIB":
            raise ValueError(f"{self.path}: bad fractal heap indirect block")
        p = addr + 5 + self.so + geom["off_bytes"]
        size = geom["start"]
        for r in range(nrows):
            if r >= 2:
                size *= 2
            for _ in range(geom["width"]):
                child = self._addr(p)
                p += self.so
                if size <= geom["max_direct"]:
                    if child is not UNDEF:
                        out.append((child, size))
                else:
                    raise NotImplementedError("fractal heap with nested indirect blocks")

    def _direct_block_objects(self, addr, size, geom):
        """Link messages packed in a direct block: parsed one after the
        other from the block's start until the free space (a zero version
        byte)."""
        if self._bytes(addr, 4) != b"FHDB":
            raise ValueError(f"{self.path}: bad fractal heap direct block")
        p = addr + 5 + self.so + geom["off_bytes"] + (4 if geom["cs"] else 0)
        end = addr + size
        objs = []
        while p < end and self.b[p] == 1:
            parser = _LinkParser(self, None, p)
            name, target = parser.parse()
            objs.append(self._bytes(p, parser.end - p))
            p = parser.end
        return objs

    def _resolve(self, path):
        addr = self.root
        for part in [s for s in path.split("/") if s]:
            links = self._links(addr)
            if part not in links:
                raise KeyError(path)
            addr = links[part]
        return addr

    def keys(self):
        return list(self._links(self.root).keys())

    def __contains__(self, path):
        try:
            self._resolve(path)
            return True
        except KeyError:
            return False

    def __getitem__(self, path):
        return Dataset(self, self._resolve(path), path)

    def close(self):
        self.b = self.buf = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class _LinkParser:
    """Link message (type 0x06) at file offset p, or in the bytes data."""

    def __init__(self, f, data, p=0):
        self.f, self.data, self.p = f, data, p

    def _byte(self, q):
        return self.data[q] if self.data is not None else self.f.b[q]

    def _uint(self, q, n):
        if self.data is not None:
            return int.from_bytes(self.data[q:q + n], "little")
        return self.f._u(q, n)

    def parse(self):
        q = self.p
        if self._byte(q) != 1:
            raise NotImplementedError(f"link message version {self._byte(q)}")
        flags = self._byte(q + 1)
        q += 2
        ltype = 0
        if flags & 0x08:
            ltype = self._byte(q)
            q += 1
        if flags & 0x04:
            q += 8
        if flags & 0x10:
            q += 1
        nlen_size = 1 << (flags & 3)
        nlen = self._uint(q, nlen_size)
        q += nlen_size
        if self.data is not None:
            name = bytes(self.data[q:q + nlen]).decode()
        else:
            name = self.f._bytes(q, nlen).decode()
        q += nlen
        target = None
        if ltype == 0:
            raw = self._uint(q, self.f.so)
            target = None if raw == (1 << (8 * self.f.so)) - 1 else raw + self.f.base
            q += self.f.so
        else:  # soft / external links: not followed
            q += 2 + self._uint(q, 2)
        self.end = q
        return name, target


class Dataset:
    """A dataset: ``shape``, ``dtype`` and ``[...]`` / ``read()``."""

    def __init__(self, f, addr, name):
        self.f, self.name = f, name
        msgs = f._messages(addr)
        self.shape = self.dtype = self.layout = None
        self.filters = []
        for mtype, p, size, _ in msgs:
            if mtype == 0x01:
                self.shape = self._dataspace(p)
            elif mtype == 0x03:
                self.dtype = self._datatype(p)
            elif mtype == 0x08:
                self.layout = (p, size)
            elif mtype == 0x0B:
                self.filters = self._filters(p)
            elif mtype == 0x17:
                raise NotImplementedError(f"{name}: external data storage")
        if self.shape is None or self.dtype is None or self.layout is None:
            raise ValueError(f"{f.path}:{name} is not a dataset")

    def __len__(self):
        return self.shape[0] if self.shape else 1

    # ----------------------------------------------------------- messages
    def _dataspace(self, p):
        f = self.f
        ver, rank, flags = f.b[p], f.b[p + 1], f.b[p + 2]
        q = p + (8 if ver == 1 else 4)
        if ver == 2 and f.b[p + 3] == 2:
            return ()  # null dataspace
        return tuple(f._u(q + i * f.sl, f.sl) for i in range(rank))

    def _datatype(self, p):
        f = self.f
        cls = f.b[p] & 0x0F
        bits0 = f.b[p + 1]
        size = f._u(p + 4, 4)
        order = ">" if bits0 & 1 else "<"
        if cls == 0:  # fixed point
            signed = bool(bits0 & 0x08)
            if size not in (1, 2, 4, 8):
                raise NotImplementedError(f"{self.name}: {size}-byte integers")
            return np.dtype(f"{order}{'i' if signed else 'u'}{size}")
        if cls == 1:  # IEEE floating point (VAX order is not supported)
            if bits0 & 0x40:
                raise NotImplementedError(f"{self.name}: VAX floating-point byte order")
            if size not in (2, 4, 8):
                raise NotImplementedError(f"{self.name}: {size}-byte floats")
            return np.dtype(f"{order}f{size}")
        raise NotImplementedError(f"{self.name}: datatype class {cls}")

    def _filters(self, p):
        f = self.f
        ver, n = f.b[p], f.b[p + 1]
        q = p + (8 if ver == 1 else 2)
        out = []
        for _ in range(n):
            fid = f._u(q, 2)
            q += 2
            nlen = 0
            if ver == 1 or fid >= 256:
                nlen = f._u(q, 2)
                q += 2
            flags, nval = f._u(q, 2), f._u(q + 2, 2)
            q += 4
            if ver == 1:
                nlen = (nlen + 7) // 8 * 8
            q += nlen
            vals = [f._u(q + 4 * i, 4) for i in range(nval)]
            q += 4 * nval
            if ver == 1 and nval % 2:
                q += 4
            out.append((fid, flags, vals))
        return out

    # --------------------------------------------------------------- data
    def _unfilter(self, raw, mask):
        data = raw
        for i in range(len(self.filters) - 1, -1, -1):
            fid, _, vals = self.filters[i]
            if mask >> i & 1:
                continue
            if fid == 1:
                data = zlib.decompress(data)
            elif fid == 2:
                es = vals[0] if vals else self.dtype.itemsize
                a = np.frombuffer(data, dtype=np.uint8)
                n = len(a) // es
                body = a[:n * es].reshape(es, n).T.reshape(-1)
                data = body.tobytes() + a[n * es:].tobytes()
            elif fid == 3:
                data = data[:-4]
            else:
                raise NotImplementedError(f"{self.name}: HDF5 filter {fid}")
        return data

    def read(self):
        f = self.f
        p, _ = self.layout
        ver = f.b[p]
        count = int(np.prod(self.shape)) if self.shape else 1
        nbytes = count * self.dtype.itemsize
        cls = f.b[p + 1]
        if ver not in (3, 4):
            raise NotImplementedError(f"{self.name}: data layout version {ver}")
        if cls == 0:  # compact
            size = f._u(p + 2, 2)
            out = np.frombuffer(f._bytes(p + 4, size), dtype=self.dtype, count=count)
        elif cls == 1:  # contiguous
            addr = f._addr(p + 2)
            if addr is UNDEF:
                out = np.zeros(count, dtype=self.dtype)
            else:
                out = np.frombuffer(f._bytes(addr, nbytes), dtype=self.dtype, count=count)
        elif cls == 2:
            out = self._read_chunked(p, ver, count)
        else:
            raise NotImplementedError(f"{self.name}: layout class {cls}")
        return np.array(out, dtype=self.dtype.newbyteorder("=")).reshape(self.shape)

    def _chunk_list(self, p, ver, rank):
        """[(offsets, address, stored size, filter mask)] and chunk dims."""
        f = self.f
        if ver == 3:
            ndim = f.b[p + 2]
            btree = f._addr(p + 3)
            q = p + 3 + f.so
            cdims = [f._u(q + 4 * i, 4) for i in range(ndim)][:rank]
            return self._btree_chunks(btree, ndim), cdims
        flags, ndim, enc = f.b[p + 2], f.b[p + 3], f.b[p + 4]
        q = p + 5
        cdims = [f._u(q + enc * i, enc) for i in range(ndim)][:rank]
        q += enc * ndim
        itype = f.b[q]
        q += 1
        csize = int(np.prod(cdims)) * self.dtype.itemsize
        if itype == 1:  # single chunk
            size, mask = csize, 0
            if flags & 2:
                size, mask = f._u(q, f.sl), f._u(q + f.sl, 4)
                q += f.sl + 4
            return [((0,) * rank, f._addr(q), size, mask)], cdims
        if itype == 2:  # implicit: chunks stored in order, unfiltered
            addr = f._addr(q)
            grid = [-(-s // c) for s, c in zip(self.shape, cdims)]
            out = []
            for i, idx in enumerate(np.ndindex(*grid)):
                out.append((tuple(j * c for j, c in zip(idx, cdims)), addr + i * csize, csize, 0))
            return out, cdims
        if itype == 3:  # fixed array
            q += 1  # page bits
            return self._fixed_array_chunks(f._addr(q), cdims, bool(self.filters)), cdims
        raise NotImplementedError(f"{self.name}: chunk index type {itype} (extensible array / v2 B-tree)")

    def _btree_chunks(self, addr, ndim):
        f = self.f
        out = []
        stack = [addr]
        ksize = 8 + 8 * ndim
        while stack:
            node = stack.pop()
            if f._bytes(node, 4) != b"TREE":
                raise ValueError(f"{f.path}: bad chunk B-tree node")
            level, used = f.b[node + 5], f._u(node + 6, 2)
            q = node + 8 + 2 * f.so
            for i in range(used):
                k = q + i * (ksize + f.so)
                size, mask = f._u(k, 4), f._u(k + 4, 4)
                offs = tuple(f._u(k + 8 + 8 * j, 8) for j in range(ndim - 1))
                child = f._addr(k + ksize)
                if level > 0:
                    stack.append(child)
                else:
                    out.append((offs, child, size, mask))
        return out

    def _fixed_array_chunks(self, addr, cdims, filtered):
        f = self.f
        if f._bytes(addr, 4) != b"FAHD":
            raise ValueError(f"{f.path}: bad fixed array header")
        esize, pbits = f.b[addr + 6], f.b[addr + 7]
        n = f._u(addr + 8, f.sl)
        dblk = f._addr(addr + 8 + f.sl)
        if f._bytes(dblk, 4) != b"FADB":
            raise ValueError(f"{f.path}: bad fixed array data block")
        q = dblk + 6 + f.so
        npages = 0
        if n > (1 << pbits):
            npages = -(-n // (1 << pbits))
            q += (npages + 7) // 8 + 4  # page init bitmap + block checksum
        else:
            pass
        grid = [-(-s // c) for s, c in zip(self.shape, cdims)]
        csize = int(np.prod(cdims)) * self.dtype.itemsize
        out = []
        per_page = 1 << pbits
        for i, idx in enumerate(np.ndindex(*grid)):
            if i >= n:
                break
            if npages:
                page, j = divmod(i, per_page)
                e = q + page * (per_page * esize + 4) + j * esize
            else:
                e = q + i * esize
            a = f._addr(e)
            if filtered:
                size = f._u(e + f.so, esize - f.so - 4)
                mask = f._u(e + esize - 4, 4)
            else:
                size, mask = csize, 0
            if a is not UNDEF:
                out.append((tuple(j * c for j, c in zip(idx, cdims)), a, size, mask))
        return out

    def _read_chunked(self, p, ver, count):
        f = self.f
        rank = len(self.shape)
        chunks, cdims = self._chunk_list(p, ver, rank)
        out = np.zeros(self.shape, dtype=self.dtype)
        ccount = int(np.prod(cdims))
        for offs, addr, size, mask in chunks:
            raw = f._bytes(addr, size)
            if self.filters:
                raw = self._unfilter(raw, mask)
            block = np.frombuffer(raw, dtype=self.dtype, count=ccount).reshape(cdims)
            sl = tuple(slice(o, min(o + c, s)) for o, c, s in zip(offs, cdims, self.shape))
            out[sl] = block[tuple(slice(0, s.stop - s.start) for s in sl)]
        return out.reshape(-1)

    def __getitem__(self, key):
        return self.read()[key]


def read_columns(path, column_names):
    """``h5_reader`` (``catalog_processor.py:16-88``) for one process: a
    structured array with one float64 field per requested column, over one
    file or a list of files.  The row count of a file is the length of the
    first requested column it holds; requested columns a file lacks are
    left NaN (the reference leaves them uninitialised, ``np.empty``)."""
    files = [path] if isinstance(path, str) else list(path)
    dtype = [(c, "f8") for c in column_names]
    parts = []
    for fn in files:
        with File(fn) as f:
            present = [c for c in column_names if c in f]
            nrows = len(f[present[0]]) if present else 0
            part = np.full(nrows, np.nan, dtype=dtype)
            for c in present:
                part[c] = f[c].read()[:nrows]
            parts.append(part)
    if not parts:
        return np.zeros(0, dtype=dtype)
    return parts[0] if len(parts) == 1 else np.concatenate(parts)
