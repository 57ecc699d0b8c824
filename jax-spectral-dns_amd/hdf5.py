"""HDF5 field and trajectory files in the layout the reference writes.

The reference saves fields with h5py (field.py:144-151, field.py:718-736):
one group per file named str(time_step) holding one gzip-9 dataset named
after the field (a PhysicalField's carries an int64 attribute "time_step"),
and trajectories as one root dataset "trajectory" (equation.py:362-372);
it reads them back by group and dataset name (field.py:409-431,
field.py:1352-1373).  h5py is not importable here, so this module reads and
writes that subset of the HDF5 file format itself (NumPy + zlib), in the
structures libhdf5 writes by default (h5py's libver "earliest"):

  - superblock version 0, 8-byte offsets and lengths;
  - version 1 object headers (continuation blocks followed when reading);
  - groups as symbol tables: a version 1 B-tree of type 0 over symbol-table
    nodes (SNOD) whose link names live in a local heap, names in strcmp order;
  - datasets: simple dataspace (version 1, maximum sizes = sizes), fixed-point
    / IEEE floating-point / compound {r, i} (complex) datatypes, contiguous or
    chunked layout (version 3), chunks indexed by a version 1 B-tree of type 1
    and compressed by the deflate filter (zlib streams);
  - scalar / simple attributes (version 1 messages; versions 2 and 3 read).

Anything else (version 2 superblocks and object headers, link messages,
other filters) is reported as unsupported.  Files written here open in
h5py / libhdf5, and h5py's files open here; both directions are checked in
the tests against h5py itself where an interpreter with h5py is available.
Host code only; nothing here touches the GPU.
"""

from __future__ import annotations

import struct
import zlib
from typing import Any, Dict, List, Optional, Tuple, Union

import numpy as np

UNDEF = 0xFFFFFFFFFFFFFFFF
SIGNATURE = b"\x89HDF\r\n\x1a\n"
LEAF_K = 4  # symbol-table node capacity 2K = 8 entries
INTERNAL_K = 16  # group B-tree node capacity 2K = 32 children
CHUNK_K = 32  # chunk B-tree node capacity 2K = 64 children (version 0 superblocks)
HEAP_FREE_NULL = 1  # end of a local heap's free list

MSG_NIL, MSG_DATASPACE, MSG_DATATYPE, MSG_FILL_OLD, MSG_FILL = 0x0, 0x1, 0x3, 0x4, 0x5
MSG_LAYOUT, MSG_FILTERS, MSG_ATTRIBUTE, MSG_CONT, MSG_STAB, MSG_MTIME = 0x8, 0xB, 0xC, 0x10, 0x11, 0x12
FILTER_DEFLATE, FILTER_SHUFFLE = 1, 2


class HDF5Error(Exception):
    pass


class Dataset:
    """An array with its attributes (the value of a dataset entry in a tree)."""

    def __init__(self, data: np.ndarray, attrs: Optional[Dict[str, Any]] = None):
        self.data = np.asarray(data)
        self.attrs = dict(attrs or {})

    def __repr__(self) -> str:
        return "Dataset(shape=%s, dtype=%s, attrs=%s)" % (self.data.shape, self.data.dtype, self.attrs)


Tree = Dict[str, Union["Tree", Dataset]]


def _pad8(n: int) -> int:
    return (n + 7) & ~7


# ---------------------------------------------------------------------------
# datatypes
def _encode_dtype(dt: np.dtype) -> bytes:
    dt = np.dtype(dt)
    if dt.kind == "c":
        half = np.dtype("<f%d" % (dt.itemsize // 2))
        member = _encode_dtype(half)
        out = bytearray(struct.pack("<B3sI", 0x16, (2).to_bytes(3, "little"), dt.itemsize))
        for name, off in ((b"r", 0), (b"i", half.itemsize)):
            nm = name + b"\0"
            out += nm + b"\0" * (_pad8(len(nm)) - len(nm))
            out += struct.pack("<IB3sI4s16s", off, 0, b"\0" * 3, 0, b"\0" * 4, b"\0" * 16)
            out += member
        return bytes(out)
    if dt.kind == "f":
        if dt.itemsize == 8:
            bits = bytes([0x20, 63, 0])
            props = struct.pack("<HHBBBBI", 0, 64, 52, 11, 0, 52, 1023)
        elif dt.itemsize == 4:
            bits = bytes([0x20, 31, 0])
            props = struct.pack("<HHBBBBI", 0, 32, 23, 8, 0, 23, 127)
        else:
            raise HDF5Error("unsupported float size %d" % dt.itemsize)
        if dt.byteorder == ">":
            bits = bytes([bits[0] | 1]) + bits[1:]
        return struct.pack("<B3sI", 0x11, bits, dt.itemsize) + props
    if dt.kind in "iu":
        b0 = (0x08 if dt.kind == "i" else 0) | (1 if dt.byteorder == ">" else 0)
        return struct.pack("<B3sIHH", 0x10, bytes([b0, 0, 0]), dt.itemsize, 0, 8 * dt.itemsize)
    raise HDF5Error("unsupported dtype %s" % dt)


def _decode_dtype(b: bytes, p: int = 0) -> Tuple[np.dtype, int]:
    """NumPy dtype of the datatype message at b[p:], and its encoded length."""
    cv, bits, size = struct.unpack_from("<B3sI", b, p)
    cls, ver = cv & 0x0F, cv >> 4
    if cls == 0:  # fixed point
        order = ">" if bits[0] & 1 else "<"
        kind = "i" if bits[0] & 0x08 else "u"
        return np.dtype("%s%s%d" % (order, kind, size)), 12
    if cls == 1:  # floating point
        order = ">" if bits[0] & 1 else "<"
        if size not in (4, 8):
            raise HDF5Error("unsupported float size %d" % size)
        return np.dtype("%sf%d" % (order, size)), 20
    if cls == 6:  # compound
        nmem = int.from_bytes(bits[:2], "little")
        q = p + 8
        names, offsets, types = [], [], []
        for _ in range(nmem):
            end = b.index(b"\0", q)
            name = b[q:end].decode()
            if ver == 3:
                q = end + 1
                nb = max(1, (size.bit_length() + 7) // 8)
                off = int.from_bytes(b[q : q + nb], "little")
                q += nb
            else:
                q += _pad8(end - q + 1)
                off = struct.unpack_from("<I", b, q)[0]
                q += 4 + 1 + 3 + (4 + 4 + 16 if ver == 1 else 0)
            mt, ml = _decode_dtype(b, q)
            q += ml
            names.append(name)
            offsets.append(off)
            types.append(mt)
        dt = np.dtype({"names": names, "formats": types, "offsets": offsets, "itemsize": size})
        return dt, q - p
    raise HDF5Error("unsupported datatype class %d" % cls)


def _as_numpy(raw: np.ndarray) -> np.ndarray:
    """Compound {r, i} of equal floats -> complex; anything else unchanged."""
    dt = raw.dtype
    if dt.names and len(dt.names) == 2 and tuple(n.lower() for n in dt.names) in (("r", "i"), ("real", "imag")):
        fr, fi = dt.fields[dt.names[0]], dt.fields[dt.names[1]]
        if fr[0] == fi[0] and fr[0].kind == "f" and fr[1] == 0 and fi[1] == fr[0].itemsize:
            out = np.empty(raw.shape, dtype=np.dtype("c%d" % (2 * fr[0].itemsize)))
            out.real = raw[dt.names[0]]
            out.imag = raw[dt.names[1]]
            return out
    return raw


def _to_little(a: np.ndarray) -> np.ndarray:
    a = np.asarray(a)
    if not a.flags.c_contiguous:  # (np.ascontiguousarray would make a scalar 1-d)
        a = a.copy(order="C")
    if a.dtype.kind == "c":
        return a.astype("<c%d" % a.dtype.itemsize, copy=False)
    if a.dtype.kind in "fiu":
        return a.astype(a.dtype.newbyteorder("<"), copy=False)
    if a.dtype.kind == "b":
        return a.astype("<u1")
    raise HDF5Error("unsupported dtype %s" % a.dtype)


def _encode_space(shape: Tuple[int, ...]) -> bytes:
    if len(shape) == 0:
        return struct.pack("<BBBB4s", 1, 0, 0, 0, b"\0" * 4)
    out = struct.pack("<BBBB4s", 1, len(shape), 1, 0, b"\0" * 4)
    out += struct.pack("<%dQ" % len(shape), *shape) * 2  # sizes, then maximum sizes
    return out


def _decode_space(b: bytes, p: int) -> Tuple[int, ...]:
    ver, rank, flags = struct.unpack_from("<BBB", b, p)
    if ver == 1:
        q = p + 8
    elif ver == 2:
        if b[p + 3] == 2:  # null dataspace
            return (0,)
        q = p + 4
    else:
        raise HDF5Error("unsupported dataspace version %d" % ver)
    return tuple(struct.unpack_from("<%dQ" % rank, b, q)) if rank else ()


# ---------------------------------------------------------------------------
# writing
class _Writer:
    def __init__(self):
        self.buf = bytearray(96)  # superblock, filled in last

    def alloc(self, n: int) -> int:
        addr = _pad8(len(self.buf))
        self.buf += b"\0" * (addr + n - len(self.buf))
        return addr

    def put(self, data: bytes) -> int:
        addr = self.alloc(len(data))
        self.buf[addr : addr + len(data)] = data
        return addr

    def object_header(self, messages: List[Tuple[int, int, bytes]]) -> int:
        body = bytearray()
        for mtype, flags, data in messages:
            n = _pad8(len(data))
            body += struct.pack("<HHB3s", mtype, n, flags, b"\0" * 3) + data + b"\0" * (n - len(data))
        head = struct.pack("<BBHII4s", 1, 0, len(messages), 1, len(body), b"\0" * 4)
        return self.put(head + bytes(body))

    # -- groups ------------------------------------------------------------
    def group(self, tree: Tree) -> Tuple[int, int, int]:
        """Object header, B-tree and heap addresses of a group."""
        names = sorted(tree.keys(), key=lambda s: s.encode())
        entries = []
        for name in names:
            if not name or "/" in name:
                raise HDF5Error("invalid link name %r" % name)
            node = tree[name]
            if isinstance(node, Dataset):
                entries.append((name, self.dataset(node), None))
            elif isinstance(node, dict):
                entries.append((name,) + self._group_entry(node))
            else:
                entries.append((name, self.dataset(Dataset(node)), None))
        # local heap: "" at offset 0, then the names
        heap = bytearray(8)
        name_off = {}
        for name, _, _ in entries:
            nb = name.encode() + b"\0"
            name_off[name] = len(heap)
            heap += nb + b"\0" * (_pad8(len(nb)) - len(nb))
        data_addr = self.put(bytes(heap))
        heap_addr = self.put(struct.pack("<4sB3sQQQ", b"HEAP", 0, b"\0" * 3, len(heap), HEAP_FREE_NULL, data_addr))
        # symbol-table nodes of at most 2 LEAF_K entries (none: an empty B-tree)
        snods = []
        for i in range(0, len(entries), 2 * LEAF_K):
            chunk = entries[i : i + 2 * LEAF_K]
            body = bytearray(struct.pack("<4sBBH", b"SNOD", 1, 0, len(chunk)))
            for name, addr, cache in chunk:
                if cache is None:
                    body += struct.pack("<QQII16s", name_off[name], addr, 0, 0, b"\0" * 16)
                else:
                    body += struct.pack("<QQII", name_off[name], addr, 1, 0) + struct.pack("<QQ", *cache)
            body += b"\0" * (8 + 2 * LEAF_K * 40 - len(body))
            last = name_off[chunk[-1][0]] if chunk else 0
            snods.append((self.put(bytes(body)), last))
        btree = self._btree_group(snods)
        header = self.object_header([(MSG_STAB, 0, struct.pack("<QQ", btree, heap_addr))])
        return header, btree, heap_addr

    def _group_entry(self, tree: Tree) -> Tuple[int, Tuple[int, int]]:
        header, btree, heap = self.group(tree)
        return header, (btree, heap)

    def _btree_group(self, snods: List[Tuple[int, int]]) -> int:
        """B-tree of type 0 over symbol-table nodes; keys are heap offsets of
        the last name of each child (key 0: the empty name)."""
        level, nodes = 0, [(addr, last) for addr, last in snods]
        cap = 2 * INTERNAL_K
        if not nodes:
            body = struct.pack("<4sBBHQQ", b"TREE", 0, 0, 0, UNDEF, UNDEF)
            return self.put(body + b"\0" * ((2 * cap + 1) * 8 + cap * 8))
        while True:
            parents = []
            for i in range(0, len(nodes), cap):
                kids = nodes[i : i + cap]
                body = bytearray(struct.pack("<4sBBHQQ", b"TREE", 0, level, len(kids), UNDEF, UNDEF))
                body += struct.pack("<Q", 0)
                for addr, last in kids:
                    body += struct.pack("<QQ", addr, last)
                body += b"\0" * (24 + (2 * cap + 1) * 8 - len(body))
                parents.append((self.put(bytes(body)), kids[-1][1]))
            if len(parents) == 1:
                return parents[0][0]
            nodes, level = parents, level + 1

    # -- datasets ----------------------------------------------------------
    def dataset(self, ds: Dataset, compression: Optional[int] = None) -> int:
        level = self.compression if compression is None else compression
        a = _to_little(ds.data)
        shape = a.shape
        dtmsg = _encode_dtype(a.dtype)
        msgs = [(MSG_DATASPACE, 0, _encode_space(shape)), (MSG_DATATYPE, 1, dtmsg)]
        raw = a.tobytes()
        if level is None or a.size == 0 or len(shape) == 0:
            addr = self.put(raw) if raw else UNDEF
            msgs.append((MSG_FILL, 1, struct.pack("<BBBBI", 2, 2, 2, 1, 0)))
            msgs.append((MSG_LAYOUT, 0, struct.pack("<BBQQ", 3, 1, addr, len(raw))))
        else:
            chunk = _chunk_shape(shape, a.dtype.itemsize, self.chunk_bytes)
            btree = self._chunks(a, chunk, level)
            msgs.append((MSG_FILL, 1, struct.pack("<BBBBI", 2, 3, 0, 1, 0)))
            name = b"deflate\0"
            filt = struct.pack("<BB6s", 1, 1, b"\0" * 6) + struct.pack("<HHHH", FILTER_DEFLATE, len(name), 1, 1)
            filt += name + struct.pack("<I", int(level)) + b"\0" * 4
            msgs.append((MSG_FILTERS, 1, filt))
            dims = list(chunk) + [a.dtype.itemsize]
            msgs.append((MSG_LAYOUT, 0, struct.pack("<BBBQ", 3, 2, len(dims), btree) + struct.pack("<%dI" % len(dims), *dims)))
        for key in sorted(ds.attrs):
            msgs.append((MSG_ATTRIBUTE, 0, _encode_attribute(key, ds.attrs[key])))
        return self.object_header(msgs)

    def _chunks(self, a: np.ndarray, chunk: Tuple[int, ...], level: int) -> int:
        """Write every chunk (edge chunks padded with zeros to the full chunk
        shape) and the B-tree of type 1 indexing them in C order."""
        shape, rank = a.shape, a.ndim
        grid = [-(-shape[d] // chunk[d]) for d in range(rank)]
        leaves = []
        for idx in np.ndindex(*grid):
            off = tuple(i * c for i, c in zip(idx, chunk))
            block = np.zeros(chunk, dtype=a.dtype)
            sl = tuple(slice(o, min(o + c, s)) for o, c, s in zip(off, chunk, shape))
            block[tuple(slice(0, s.stop - s.start) for s in sl)] = a[sl]
            comp = zlib.compress(block.tobytes(), int(level))
            leaves.append((off, len(comp), self.put(comp)))
        last = leaves[-1][0]
        end = tuple((o // c + 1) * c for o, c in zip(last, chunk)) + (a.dtype.itemsize,)
        return self._btree_chunks(leaves, end, rank)

    def _btree_chunks(self, leaves, end, rank) -> int:
        ksize = 8 + 8 * (rank + 1)
        cap = 2 * CHUNK_K

        def key(size, off):
            return struct.pack("<II", size, 0) + struct.pack("<%dQ" % (rank + 1), *(tuple(off) + (0,)))

        level = 0
        nodes = [(key(n, off), addr) for off, n, addr in leaves]
        end_key = struct.pack("<II", 0, 0) + struct.pack("<%dQ" % (rank + 1), *end)
        while True:
            parents = []
            groups = [nodes[i : i + cap] for i in range(0, len(nodes), cap)]
            for gi, kids in enumerate(groups):
                right = groups[gi + 1][0][0] if gi + 1 < len(groups) else end_key
                body = bytearray(struct.pack("<4sBBHQQ", b"TREE", 1, level, len(kids), UNDEF, UNDEF))
                for k, addr in kids:
                    body += k + struct.pack("<Q", addr)
                body += right
                body += b"\0" * (24 + (cap + 1) * ksize + cap * 8 - len(body))
                parents.append((kids[0][0], self.put(bytes(body))))
            if len(parents) == 1:
                return parents[0][1]
            nodes, level = parents, level + 1

    def finish(self, root: Tuple[int, int, int]) -> bytes:
        header, btree, heap = root
        sb = SIGNATURE + struct.pack("<BBBBBBBB", 0, 0, 0, 0, 0, 8, 8, 0)
        sb += struct.pack("<HHI", LEAF_K, INTERNAL_K, 0)
        sb += struct.pack("<QQQQ", 0, UNDEF, len(self.buf), UNDEF)
        sb += struct.pack("<QQII", 0, header, 1, 0) + struct.pack("<QQ", btree, heap)
        assert len(sb) == 96
        self.buf[:96] = sb
        return bytes(self.buf)


def _chunk_shape(shape: Tuple[int, ...], itemsize: int, target: int = 1 << 20) -> Tuple[int, ...]:
    """Halve the longest axis (first among equals) until a chunk is at most
    target bytes; chunks stay far below the 4 GiB limit of a chunk."""
    c = list(shape)
    while int(np.prod(c)) * itemsize > target and max(c) > 1:
        d = int(np.argmax(c))
        c[d] = -(-c[d] // 2)
    return tuple(c)


def _encode_attribute(name: str, value: Any) -> bytes:
    a = _to_little(np.asarray(value))
    nb = name.encode() + b"\0"
    dt = _encode_dtype(a.dtype)
    sp = _encode_space(a.shape)
    out = struct.pack("<BBHHH", 1, 0, len(nb), len(dt), len(sp))
    for part in (nb, dt, sp):
        out += part + b"\0" * (_pad8(len(part)) - len(part))
    return out + a.tobytes()


def write(filename: str, tree: Tree, compression: Optional[int] = 9, chunk_bytes: int = 1 << 20) -> None:
    """Write a file whose root group holds tree: nested dicts are groups,
    Dataset (or array) values are datasets.  compression: deflate level of
    every dataset (None: contiguous, uncompressed); chunk_bytes: largest
    chunk of a compressed dataset."""
    w = _Writer()
    w.compression = compression
    w.chunk_bytes = int(chunk_bytes)
    root = w.group(tree)
    with open(filename, "wb") as f:
        f.write(w.finish(root))


# ---------------------------------------------------------------------------
# reading
class _Reader:
    def __init__(self, data: bytes):
        self.b = data
        if data[:8] != SIGNATURE:
            raise HDF5Error("not an HDF5 file (or a user block precedes the superblock)")
        ver = data[8]
        if ver not in (0, 1):
            raise HDF5Error("superblock version %d is not supported" % ver)
        if data[13] != 8 or data[14] != 8:
            raise HDF5Error("only 8-byte offsets and lengths are supported")
        p = 24 + (4 if ver == 1 else 0)
        self.base = struct.unpack_from("<Q", data, p)[0]
        self.root = struct.unpack_from("<Q", data, p + 32 + 8)[0]

    def u64(self, p: int) -> int:
        return struct.unpack_from("<Q", self.b, self.base + p)[0]

    def messages(self, addr: int) -> List[Tuple[int, bytes]]:
        b, p = self.b, self.base + addr
        ver = b[p]
        if ver != 1:
            raise HDF5Error("object header version %d is not supported" % ver)
        nmsg, _, size = struct.unpack_from("<HII", b, p + 2)
        blocks = [(p + 16, size)]
        out = []
        while blocks:
            q, n = blocks.pop(0)
            end = q + n
            while q + 8 <= end:
                mtype, msize, flags = struct.unpack_from("<HHB", b, q)
                data = b[q + 8 : q + 8 + msize]
                if flags & 0x02:
                    raise HDF5Error("shared object header messages are not supported")
                if mtype == MSG_CONT:
                    caddr, clen = struct.unpack_from("<QQ", data, 0)
                    blocks.append((self.base + caddr, clen))
                elif mtype != MSG_NIL:
                    out.append((mtype, data))
                q += 8 + msize
        return out

    def node(self, addr: int) -> Union[Tree, Dataset]:
        msgs = self.messages(addr)
        types = {t for t, _ in msgs}
        if MSG_STAB in types:
            data = next(d for t, d in msgs if t == MSG_STAB)
            btree, heap = struct.unpack_from("<QQ", data, 0)
            return {name: self.node(a) for name, a in self.group_entries(btree, heap)}
        if MSG_DATASPACE in types and MSG_LAYOUT in types:
            return self.dataset(msgs)
        if 0x2 in types or 0x6 in types:
            raise HDF5Error("groups stored as link messages (new-style groups) are not supported")
        raise HDF5Error("object at %#x is neither a group nor a dataset" % addr)

    def group_entries(self, btree: int, heap: int) -> List[Tuple[str, int]]:
        b = self.b
        hp = self.base + heap
        if b[hp : hp + 4] != b"HEAP":
            raise HDF5Error("bad local heap signature")
        heap_data = self.base + struct.unpack_from("<Q", b, hp + 24)[0]

        def name_at(off):
            s = heap_data + off
            return b[s : b.index(b"\0", s)].decode()

        out = []

        def walk(addr):
            p = self.base + addr
            sig = b[p : p + 4]
            if sig == b"SNOD":
                n = struct.unpack_from("<H", b, p + 6)[0]
                for i in range(n):
                    off, obj = struct.unpack_from("<QQ", b, p + 8 + 40 * i)
                    out.append((name_at(off), obj))
                return
            if sig != b"TREE" or b[p + 4] != 0:
                raise HDF5Error("bad group B-tree node")
            n = struct.unpack_from("<H", b, p + 6)[0]
            for i in range(n):
                walk(struct.unpack_from("<Q", b, p + 24 + 8 + 16 * i)[0])

        walk(btree)
        return out

    def dataset(self, msgs) -> Dataset:
        shape = dtype = layout = None
        filters: List[Tuple[int, List[int]]] = []
        attrs: Dict[str, Any] = {}
        for t, d in msgs:
            if t == MSG_DATASPACE:
                shape = _decode_space(d, 0)
            elif t == MSG_DATATYPE:
                dtype = _decode_dtype(d, 0)[0]
            elif t == MSG_LAYOUT:
                layout = d
            elif t == MSG_FILTERS:
                filters = _decode_filters(d)
            elif t == MSG_ATTRIBUTE:
                k, v = self.attribute(d)
                attrs[k] = v
        ver, cls = layout[0], layout[1]
        if ver != 3:
            raise HDF5Error("data layout version %d is not supported" % ver)
        n = int(np.prod(shape)) if shape else 1
        if cls == 0:  # compact
            size = struct.unpack_from("<H", layout, 2)[0]
            raw = np.frombuffer(layout[4 : 4 + size], dtype=dtype, count=n)
        elif cls == 1:  # contiguous
            addr, size = struct.unpack_from("<QQ", layout, 2)
            if addr == UNDEF:
                raw = np.zeros(n, dtype=dtype)
            else:
                raw = np.frombuffer(self.b, dtype=dtype, count=n, offset=self.base + addr)
        elif cls == 2:
            ndim = layout[2]
            btree = struct.unpack_from("<Q", layout, 3)[0]
            chunk = struct.unpack_from("<%dI" % ndim, layout, 11)[:-1]
            raw = self.chunked(btree, tuple(shape), chunk, dtype, filters)
        else:
            raise HDF5Error("data layout class %d is not supported" % cls)
        arr = np.array(raw).reshape(shape if shape else ())
        return Dataset(_as_numpy(arr), {k: _as_numpy(np.asarray(v)) if isinstance(v, np.ndarray) else v for k, v in attrs.items()})

    def chunked(self, btree, shape, chunk, dtype, filters) -> np.ndarray:
        out = np.zeros(shape, dtype=dtype)
        if btree == UNDEF:
            return out
        rank = len(shape)
        b = self.b
        ksize = 8 + 8 * (rank + 1)

        def walk(addr):
            p = self.base + addr
            if b[p : p + 4] != b"TREE" or b[p + 4] != 1:
                raise HDF5Error("bad chunk B-tree node")
            level, n = b[p + 5], struct.unpack_from("<H", b, p + 6)[0]
            q = p + 24
            for _ in range(n):
                size, mask = struct.unpack_from("<II", b, q)
                off = struct.unpack_from("<%dQ" % rank, b, q + 8)
                child = struct.unpack_from("<Q", b, q + ksize)[0]
                q += ksize + 8
                if level > 0:
                    walk(child)
                    continue
                data = b[self.base + child : self.base + child + size]
                for i, (fid, cd) in reversed(list(enumerate(filters))):
                    if mask & (1 << i):
                        continue
                    if fid == FILTER_DEFLATE:
                        data = zlib.decompress(data)
                    elif fid == FILTER_SHUFFLE:
                        s = dtype.itemsize
                        data = np.frombuffer(data, np.uint8).reshape(s, -1).T.tobytes()
                    else:
                        raise HDF5Error("filter %d is not supported" % fid)
                block = np.frombuffer(data, dtype=dtype, count=int(np.prod(chunk))).reshape(chunk)
                sl = tuple(slice(o, min(o + c, s)) for o, c, s in zip(off, chunk, shape))
                out[sl] = block[tuple(slice(0, x.stop - x.start) for x in sl)]

        walk(btree)
        return out

    def attribute(self, d: bytes) -> Tuple[str, Any]:
        ver = d[0]
        if ver == 1:
            nlen, tlen, slen = struct.unpack_from("<HHH", d, 2)
            q = 8
            name = d[q : q + nlen - 1].decode()
            q += _pad8(nlen)
            dt = _decode_dtype(d, q)[0]
            q += _pad8(tlen)
            shape = _decode_space(d, q)
            q += _pad8(slen)
        elif ver in (2, 3):
            nlen, tlen, slen = struct.unpack_from("<HHH", d, 2)
            q = 8 + (1 if ver == 3 else 0)
            name = d[q : q + nlen - 1].decode()
            q += nlen
            dt = _decode_dtype(d, q)[0]
            q += tlen
            shape = _decode_space(d, q)
            q += slen
        else:
            raise HDF5Error("attribute message version %d is not supported" % ver)
        n = int(np.prod(shape)) if shape else 1
        val = _as_numpy(np.frombuffer(d, dtype=dt, count=n, offset=q).reshape(shape))
        return name, (val[()] if val.ndim == 0 else val.copy())


def _decode_filters(d: bytes) -> List[Tuple[int, List[int]]]:
    ver, n = d[0], d[1]
    q = 8 if ver == 1 else 2
    out = []
    for _ in range(n):
        fid = struct.unpack_from("<H", d, q)[0]
        if ver == 1 or fid >= 256:
            nlen, flags, ncd = struct.unpack_from("<HHH", d, q + 2)
            q += 8
            q += _pad8(nlen) if ver == 1 else nlen
        else:
            flags, ncd = struct.unpack_from("<HH", d, q + 2)
            q += 6
        cd = list(struct.unpack_from("<%dI" % ncd, d, q))
        q += 4 * ncd
        if ver == 1 and ncd % 2:
            q += 4
        out.append((fid, cd))
    return out


def read(filename: str) -> Tree:
    """The whole file as nested dicts of Dataset (NumPy arrays + attrs)."""
    with open(filename, "rb") as f:
        r = _Reader(f.read())
    return r.node(r.root)


def names(tree: Tree) -> List[str]:
    """Link names of a group in h5py's iteration order (strcmp order)."""
    return sorted(tree.keys(), key=lambda s: s.encode())
