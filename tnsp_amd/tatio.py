"""TAT's binary tensor format (``Tensor::dump`` / ``Tensor::load``).

Byte layout, version 1 (TAT/include/TAT/miscellaneous/io.hpp:687-705, the
item writers at :199-209, 263-339, 456-472), little-endian:

* ``"TAT"`` then the version as uint16;
* names: uint64 count, then each name as uint64 length + bytes;
* edges: uint64 count, then per edge: the arrow as one byte (fermionic
  symmetries only), uint64 segment count and the segments as the raw memory
  of ``std::pair<Symmetry, uint64>``: the symmetry (a ``std::tuple`` of int32
  for U(1) and bool for Z(2) components, laid out last component first, as
  libstdc++ lays out tuples; one byte for no symmetry) at offset 0, padded to
  8, the dimension at offset 8 -- 16 bytes per segment;
* storage: uint64 count and the raw scalars (float64 for ``.D``, float32 for
  ``.S``).

Version 0 (no header) stores blocks ordered by their symmetry lists; loading
it moves each block to TAT's current order (``Core::_block_order_v0_to_v1``,
TAT/include/TAT/structure/core.hpp:202-242).  Padding bytes are written as
zeros and ignored when read.
"""
import struct

import numpy as np

from .shape import Edge

_SCALAR = {"float64": np.float64, "float32": np.float32}


def _sym_layout(kind):
    """(size of the symmetry, [(offset, is_z2)] per component in component
    order)."""
    comps = list(kind.comps)
    if not comps:
        return 1, []
    offsets = [None] * len(comps)
    off = 0
    for i in reversed(range(len(comps))):
        z2 = comps[i][0]
        size = 1 if z2 else 4
        off = (off + size - 1) // size * size
        offsets[i] = (off, z2)
        off += size
    align = 1 if all(z2 for z2, _ in comps) else 4
    return (off + align - 1) // align * align, offsets


def _segment_bytes(kind, q, d):
    size, offs = _sym_layout(kind)
    raw = bytearray(16)
    for (off, z2), v in zip(offs, q):
        if z2:
            raw[off] = 1 if v else 0
        else:
            struct.pack_into("<i", raw, off, int(v))
    struct.pack_into("<Q", raw, 8, int(d))
    return bytes(raw)


def _segment_value(kind, raw):
    _, offs = _sym_layout(kind)
    q = tuple(int(raw[off] != 0) if z2 else struct.unpack_from("<i", raw, off)[0] for off, z2 in offs)
    return q, struct.unpack_from("<Q", raw, 8)[0]


def serialize(kind, names, edges, storage, version=1):
    """Bytes of a tensor with ``names``, ``edges`` (``Edge`` objects) and
    ``storage`` (numpy array in TAT's storage order)."""
    out = bytearray()
    if version == 1:
        out += b"TAT" + struct.pack("<H", 1)
    elif version != 0:
        raise ValueError("TAT binary versions are 0 and 1")
    out += struct.pack("<Q", len(names))
    for n in names:
        b = n.encode()
        out += struct.pack("<Q", len(b)) + b
    out += struct.pack("<Q", len(edges))
    for e in edges:
        if kind.is_fermi:
            out += b"\x01" if e.arrow else b"\x00"
        out += struct.pack("<Q", len(e.segments))
        for q, d in e.segments:
            out += _segment_bytes(kind, q, d)
    arr = np.ascontiguousarray(storage)
    out += struct.pack("<Q", arr.size) + arr.tobytes()
    return bytes(out)


def parse(kind, data, dtype="float64"):
    """(names, edges, storage, version) of TAT tensor bytes.  Version 0
    storage is returned in its own block order: ``v0_to_v1`` reorders it
    (``Tensor.load`` does)."""
    data = bytes(data)
    pos = 0
    version = 0
    if data[:3] == b"TAT":
        version = struct.unpack_from("<H", data, 3)[0]
        pos = 5
        if version != 1:
            raise ValueError(f"unsupported TAT binary version {version}")

    def u64():
        nonlocal pos
        v = struct.unpack_from("<Q", data, pos)[0]
        pos += 8
        return v

    names = []
    for _ in range(u64()):
        n = u64()
        names.append(data[pos:pos + n].decode())
        pos += n
    edges = []
    for _ in range(u64()):
        arrow = False
        if kind.is_fermi:
            arrow = data[pos] != 0
            pos += 1
        segs = []
        for _ in range(u64()):
            segs.append(_segment_value(kind, data[pos:pos + 16]))
            pos += 16
        edges.append(Edge(kind, segs, arrow))
    count = u64()
    dt = np.dtype(_SCALAR[dtype])
    if pos + count * dt.itemsize > len(data):
        raise ValueError("truncated TAT tensor bytes")
    storage = np.frombuffer(data, dtype=dt, count=count, offset=pos).copy()
    return names, edges, storage, version


def padding_positions(kind, data):
    """Byte positions of the segment padding in tensor bytes (whatever the
    writer left there: TAT writes the pair's raw memory)."""
    data = bytes(data)
    pos = 5 if data[:3] == b"TAT" else 0
    size, _ = _sym_layout(kind)

    def u64(p):
        return struct.unpack_from("<Q", data, p)[0]

    n = u64(pos)
    pos += 8
    for _ in range(n):
        pos += 8 + u64(pos)
    out = []
    ne = u64(pos)
    pos += 8
    for _ in range(ne):
        pos += 1 if kind.is_fermi else 0
        ns = u64(pos)
        pos += 8
        for _ in range(ns):
            # an empty symmetry's one byte holds no value either
            out.extend(range(pos + (size if kind.length else 0), pos + 8))
            pos += 16
    return out


def v0_to_v1(blocks, storage):
    """Storage in the version 1 order from the version 0 order.  ``blocks``:
    (symmetries tuple, size) of every stored block in TAT's current order;
    version 0 kept the blocks sorted by their symmetry lists."""
    new_off, off = {}, 0
    for syms, size in blocks:
        new_off[syms] = (off, size)
        off += size
    out = np.empty_like(storage)
    old = 0
    for syms in sorted(new_off):
        o, size = new_off[syms]
        out[o:o + size] = storage[old:old + size]
        old += size
    return out


def v1_to_v0(blocks, storage):
    """The inverse of ``v0_to_v1`` (writes what an old TAT would have)."""
    pos, off = {}, 0
    for syms, size in blocks:
        pos[syms] = (off, size)
        off += size
    out = np.empty_like(storage)
    old = 0
    for syms in sorted(pos):
        o, size = pos[syms]
        out[old:old + size] = storage[o:o + size]
        old += size
    return out
