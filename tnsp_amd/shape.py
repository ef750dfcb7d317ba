"""Symmetries, edges and block layouts on the host side of the C ABI.

Mirrors the value types of the reference binding (PyTAT/PyTAT.hpp:128-330
for Symmetry / Edge): a symmetry is a tuple of ints (Z2 components as 0/1),
an edge is an ordered tuple of ``(symmetry, dimension)`` segments plus a fermi
arrow.  The block table itself lives in the native library
(``tns_shape``), built by the same rule as TAT's core
(TAT/include/TAT/structure/core.hpp:152-191).
"""
import ctypes

import numpy as np

from . import _lib

KIND_IDS = {
    "No": 0,
    "BoseZ2": 1,
    "BoseU1": 2,
    "FermiU1": 3,
    "FermiZ2": 4,
    "FermiU1BoseZ2": 5,
    "FermiU1BoseU1": 6,
    "FermiU1FermiU1": 7,
}
# per component: (is_z2, is_fermi)  (TAT/include/TAT/structure/symmetry.hpp:269-276)
COMPONENTS = {
    "No": (),
    "BoseZ2": ((True, False),),
    "BoseU1": ((False, False),),
    "FermiU1": ((False, True),),
    "FermiZ2": ((True, True),),
    "FermiU1BoseZ2": ((False, True), (True, False)),
    "FermiU1BoseU1": ((False, True), (False, False)),
    "FermiU1FermiU1": ((False, True), (False, True)),
}


class Kind:
    def __init__(self, name):
        self.name = name
        self.id = KIND_IDS[name]
        self.comps = COMPONENTS[name]
        self.length = len(self.comps)
        self.is_fermi = any(f for _, f in self.comps)

    def make(self, value):
        if value is None:
            vals = []
        elif isinstance(value, (tuple, list)):
            vals = list(value)
        else:
            vals = [value]
        if len(vals) > self.length:
            raise ValueError(f"too many components for {self.name} symmetry")
        vals = vals + [0] * (self.length - len(vals))
        return tuple(int(bool(v)) if z2 else int(v) for v, (z2, _) in zip(vals, self.comps))

    def zero(self):
        return (0,) * self.length

    def neg(self, q):
        return tuple(x if z2 else -x for x, (z2, _) in zip(q, self.comps))

    def add(self, a, b):
        return tuple((x ^ y) if z2 else (x + y) for x, y, (z2, _) in zip(a, b, self.comps))

    def parity(self, q):
        p = False
        for x, (_, fermi) in zip(q, self.comps):
            if fermi:
                p ^= (x % 2) != 0
        return p

    def display(self, q):
        """Scalar for one-component symmetries (bool for Z2), tuple otherwise."""
        if self.length == 0:
            return ()
        if self.length == 1:
            return bool(q[0]) if self.comps[0][0] else q[0]
        return tuple(bool(x) if z2 else x for x, (z2, _) in zip(q, self.comps))


KINDS = {n: Kind(n) for n in KIND_IDS}


class Edge:
    """PyTAT-style edge (PyTAT/PyTAT.hpp:190-330): segments + arrow."""

    __slots__ = ("kind", "segments", "arrow")

    def __init__(self, kind, segments, arrow=False):
        self.kind = kind
        self.segments = tuple((kind.make(q), int(d)) for q, d in segments)
        self.arrow = bool(arrow) if kind.is_fermi else False

    @classmethod
    def build(cls, kind, spec):
        if isinstance(spec, Edge):
            return spec
        if isinstance(spec, (int, np.integer)) and not isinstance(spec, bool):
            return cls(kind, [(kind.zero(), int(spec))])
        if isinstance(spec, tuple) and len(spec) == 2 and isinstance(spec[1], bool):
            segs = spec[0]
            if isinstance(segs, (int, np.integer)):
                return cls(kind, [(kind.zero(), int(segs))], spec[1])
            return cls(kind, _segments(kind, segs), spec[1])
        return cls(kind, _segments(kind, spec))

    def conjugate(self):
        return Edge(self.kind, [(self.kind.neg(q), d) for q, d in self.segments], not self.arrow)

    def dimension(self):
        return sum(d for _, d in self.segments)

    total_dimension = dimension

    def position_by_symmetry(self, q):
        q = self.kind.make(q)
        for i, (s, _) in enumerate(self.segments):
            if s == q:
                return i
        raise KeyError("no such symmetry in edge")

    def dimension_by_symmetry(self, q):
        return self.segments[self.position_by_symmetry(q)][1]

    def index_by_point(self, point):
        q, off = point
        q = self.kind.make(q)
        total = off
        for s, d in self.segments:
            if s == q:
                return total
            total += d
        raise KeyError("symmetry not found in edge")

    def coord_by_index(self, index):
        off = int(index)
        for i, (_, d) in enumerate(self.segments):
            if off < d:
                return i, off
            off -= d
        raise IndexError("index beyond edge dimension")

    def point_by_index(self, index):
        p, off = self.coord_by_index(index)
        return self.kind.display(self.segments[p][0]), off

    def __eq__(self, other):
        return isinstance(other, Edge) and self.segments == other.segments and self.arrow == other.arrow

    def __hash__(self):
        return hash((self.segments, self.arrow))

    def __repr__(self):
        segs = [(self.kind.display(q), d) for q, d in self.segments]
        return f"Edge({segs}, arrow={self.arrow})" if self.kind.is_fermi else f"Edge({segs})"


def _segments(kind, segs):
    out = []
    for item in segs:
        is_pair = isinstance(item, (tuple, list)) and len(item) == 2
        if kind.length >= 2:
            is_pair = is_pair and isinstance(item[0], (tuple, list))
        if is_pair:
            out.append((item[0], item[1]))
        else:
            out.append((item, 1))
    return out


class NameTable:
    """Interns edge names (any hashable) to the int32 ids of the C ABI."""

    def __init__(self):
        self.to_id = {}
        self.to_name = []

    def id(self, name):
        i = self.to_id.get(name)
        if i is None:
            i = len(self.to_name)
            self.to_id[name] = i
            self.to_name.append(name)
        return i

    def ids(self, names):
        return [self.id(n) for n in names]

    def name(self, i):
        return self.to_name[i]


NAMES = NameTable()


class Shape:
    """Owning wrapper of a ``tns_shape`` handle plus its Python view."""

    __slots__ = ("handle", "kind", "_names", "_edges", "size", "_blocks")

    def __init__(self, handle, kind=None, names=None, edges=None, lazy=False):
        self.handle = handle
        L = _lib.lib()
        k, rank, nseg_total, nblocks, size = (ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int64(),
                                              ctypes.c_int64())
        _lib.check(L.tns_shape_counts(handle, ctypes.byref(k), ctypes.byref(rank), ctypes.byref(nseg_total),
                                      ctypes.byref(nblocks), ctypes.byref(size)))
        self.size = size.value
        self._blocks = None
        if kind is None:
            kind = next(kk for kk in KINDS.values() if kk.id == k.value)
        self.kind = kind
        self._names = list(names) if names is not None else None
        self._edges = list(edges) if edges is not None else None
        if not lazy and (names is None or edges is None):
            self._export()

    @property
    def names(self):
        if self._names is None:
            self._export()
        return self._names

    @property
    def edges(self):
        if self._edges is None:
            self._export()
        return self._edges

    def _export(self):
        """Names and edges read back from the handle."""
        L = _lib.lib()
        kind = self.kind
        k, rank, nseg_total, nblocks, size = (ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int64(),
                                              ctypes.c_int64())
        _lib.check(L.tns_shape_counts(self.handle, ctypes.byref(k), ctypes.byref(rank), ctypes.byref(nseg_total),
                                      ctypes.byref(nblocks), ctypes.byref(size)))
        r, ts = rank.value, nseg_total.value
        nm = (ctypes.c_int32 * max(1, r))()
        ns = (ctypes.c_int32 * max(1, r))()
        sy = (ctypes.c_int32 * max(1, 2 * ts))()
        dm = (ctypes.c_int64 * max(1, ts))()
        ar = (ctypes.c_uint8 * max(1, r))()
        _lib.check(L.tns_shape_export(self.handle, nm, ns, sy, dm, ar))
        names, edges, s = [], [], 0
        for i in range(r):
            names.append(NAMES.name(nm[i]))
            segs = []
            for _ in range(ns[i]):
                q = tuple(sy[2 * s + c] for c in range(kind.length))
                segs.append((q, dm[s]))
                s += 1
            edges.append(Edge(kind, segs, bool(ar[i])))
        self._names = list(names)
        self._edges = list(edges)

    @classmethod
    def create(cls, kind, names, edges):
        edges = [Edge.build(kind, e) for e in edges]
        if len(names) != len(edges):
            raise ValueError("names and edges differ in length")
        nseg, syms, dims, arrows = [], [], [], []
        for e in edges:
            nseg.append(len(e.segments))
            for q, d in e.segments:
                qq = list(q) + [0] * (2 - len(q))
                syms.extend(qq)
                dims.append(d)
            arrows.append(int(e.arrow))
        h = ctypes.c_void_p()
        L = _lib.lib()
        _lib.check(L.tns_shape_create(kind.id, len(names), _lib.arr32(NAMES.ids(names)), _lib.arr32(nseg),
                                      _lib.arr32(syms), _lib.arr64(dims),
                                      (ctypes.c_uint8 * max(1, len(arrows)))(*arrows), ctypes.byref(h)))
        return cls(h, kind, names, edges)

    def rename(self, mapping):
        old = [n for n in mapping]
        for n in old:
            if n not in self.names:
                raise KeyError("Name missing in edge_rename")
        h = ctypes.c_void_p()
        _lib.check(_lib.lib().tns_shape_rename(self.handle, len(old), _lib.arr32(NAMES.ids(old)),
                                               _lib.arr32(NAMES.ids([mapping[n] for n in old])), ctypes.byref(h)))
        return Shape(h, self.kind, [mapping.get(n, n) for n in self.names], self.edges)

    def blocks(self):
        """[(positions tuple, offset, shape)] in storage order."""
        if self._blocks is None:
            L = _lib.lib()
            nb = ctypes.c_int64()
            _lib.check(L.tns_shape_counts(self.handle, None, None, None, ctypes.byref(nb), None))
            r = len(self.names)
            pos = (ctypes.c_int64 * max(1, nb.value * r))()
            off = (ctypes.c_int64 * max(1, nb.value))()
            _lib.check(L.tns_shape_blocks(self.handle, pos, off))
            out = []
            for b in range(nb.value):
                p = tuple(pos[b * r + i] for i in range(r))
                shp = tuple(self.edges[i].segments[p[i]][1] for i in range(r))
                out.append((p, off[b], shp))
            self._blocks = out
        return self._blocks

    def block_of(self, positions):
        for p, off, shp in self.blocks():
            if p == tuple(positions):
                return off, shp
        raise KeyError("no such block")

    def index(self, name):
        return self.names.index(name)

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and _lib._lib is not None:
            _lib._lib.tns_shape_free(h)
            self.handle = None
