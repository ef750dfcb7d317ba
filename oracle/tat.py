"""CPU oracle: TAT's symmetric block-sparse tensor core restated in numpy.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``): imported by tests,
``__graft_entry__.smoke()`` and the cpu_baseline leg of ``bench.py`` only.

Every function cites the reference source it restates (paths relative to
the TNSP repository root).  Semantics, not code, are followed: symmetries are
tuples of ints, edges are tuples of ``(symmetry, dim)`` segments plus a fermi
arrow, a tensor is a list of names, a list of edges and one flat storage array
holding every allowed block contiguously.
"""
import itertools

import numpy as np


# ----------------------------------------------------------------------------
# Symmetry  (TAT/include/TAT/structure/symmetry.hpp:95-239, aliases :269-276)
# ----------------------------------------------------------------------------
class SymmetryKind:
    """A symmetry group = tuple of components, each Z2 (xor) or U1 (+), each
    bose or fermi.  ``parity`` is the XOR over fermi components of oddness
    (symmetry.hpp:222-238)."""

    def __init__(self, name, comps):
        self.name = name
        self.comps = tuple(comps)  # (is_z2, is_fermi)
        self.length = len(self.comps)
        self.is_fermi = any(f for _, f in self.comps)

    def zero(self):
        return (0,) * self.length

    def make(self, value):
        if value is None:
            return self.zero()
        if isinstance(value, (tuple, list)):
            vals = list(value)
        else:
            vals = [value]
        vals = vals + [0] * (self.length - len(vals))
        if len(vals) != self.length:
            raise ValueError("bad symmetry value")
        return tuple(int(bool(v)) if z2 else int(v) for v, (z2, _) in zip(vals, self.comps))

    def add(self, a, b):
        return tuple((x ^ y) if z2 else (x + y) for x, y, (z2, _) in zip(a, b, self.comps))

    def neg(self, a):
        return tuple(x if z2 else -x for x, (z2, _) in zip(a, self.comps))

    def parity(self, a):
        p = False
        for x, (_, fermi) in zip(a, self.comps):
            if fermi:
                p ^= (x % 2) != 0
        return p

    def __repr__(self):
        return f"SymmetryKind({self.name})"


No = SymmetryKind("No", ())
BoseZ2 = SymmetryKind("BoseZ2", ((True, False),))
BoseU1 = SymmetryKind("BoseU1", ((False, False),))
FermiU1 = SymmetryKind("FermiU1", ((False, True),))
FermiZ2 = SymmetryKind("FermiZ2", ((True, True),))
FermiU1BoseZ2 = SymmetryKind("FermiU1BoseZ2", ((False, True), (True, False)))
FermiU1BoseU1 = SymmetryKind("FermiU1BoseU1", ((False, True), (False, False)))
FermiU1FermiU1 = SymmetryKind("FermiU1FermiU1", ((False, True), (False, True)))
KINDS = {k.name: k for k in (No, BoseZ2, BoseU1, FermiU1, FermiZ2, FermiU1BoseZ2, FermiU1BoseU1, FermiU1FermiU1)}


# ----------------------------------------------------------------------------
# Edge  (TAT/include/TAT/structure/edge.hpp:33-210 segments, :224-292 arrow)
# ----------------------------------------------------------------------------
class Edge:
    """Ordered segment list ``[(q, d)]`` plus a fermi arrow.  Bose edges have
    no arrow (edge.hpp:212-222: it always reads False)."""

    __slots__ = ("kind", "segments", "arrow")

    def __init__(self, kind, segments, arrow=False):
        self.kind = kind
        self.segments = tuple((kind.make(q), int(d)) for q, d in segments)
        self.arrow = bool(arrow) if kind.is_fermi else False

    @classmethod
    def build(cls, kind, spec):
        """Accept PyTAT-style edge specs: Edge, int (dimension), list of
        (sym, dim), list of syms, or (segments, arrow)."""
        if isinstance(spec, Edge):
            return spec
        if isinstance(spec, (int, np.integer)) and not isinstance(spec, bool):
            return cls(kind, [(kind.zero(), int(spec))])
        if isinstance(spec, tuple) and len(spec) == 2 and isinstance(spec[1], bool) and isinstance(spec[0], (list, tuple)):
            segs = spec[0]
            if isinstance(segs, (int, np.integer)):
                return cls(kind, [(kind.zero(), int(segs))], spec[1])
            return cls(kind, _normalize_segments(kind, segs), spec[1])
        return cls(kind, _normalize_segments(kind, spec))

    def conjugate(self):
        # edge.hpp:194-201, 289-291: negate charges, keep order, flip arrow
        return Edge(self.kind, [(self.kind.neg(q), d) for q, d in self.segments], not self.arrow)

    def total_dimension(self):
        return sum(d for _, d in self.segments)

    def position_by_symmetry(self, q):
        for i, (s, _) in enumerate(self.segments):
            if s == q:
                return i
        return None

    def dimension_by_symmetry(self, q):
        p = self.position_by_symmetry(q)
        return None if p is None else self.segments[p][1]

    def index_by_point(self, point):
        q, off = point
        total = off
        for s, d in self.segments:
            if s == q:
                return total
            total += d
        raise KeyError("symmetry not found in edge")

    def coord_by_index(self, index):
        off = index
        for i, (s, d) in enumerate(self.segments):
            if off < d:
                return i, off
            off -= d
        raise IndexError("index beyond edge dimension")

    def point_by_index(self, index):
        p, off = self.coord_by_index(index)
        return self.segments[p][0], off

    def __eq__(self, other):
        return isinstance(other, Edge) and self.segments == other.segments and self.arrow == other.arrow

    def __repr__(self):
        return f"Edge({list(self.segments)}, {self.arrow})"


class _Blocks:
    def __init__(self, t):
        self.t = t

    def _view(self, key):
        t = self.t
        items = list(key)
        if items and not isinstance(items[0], (tuple, list)):
            raise TypeError("blocks[...] takes (name, symmetry) pairs")
        given = [n for n, _ in items]
        mapping = {n: q for n, q in items}
        blk = t.block_by_symmetries(mapping) if t.rank else t.block(())
        return blk.transpose([t.names.index(n) for n in given]) if given else blk

    def __getitem__(self, key):
        return self._view(key)

    def __setitem__(self, key, value):
        self._view(key)[...] = value


def _normalize_segments(kind, segs):
    """Segments are ``(sym, dim)`` pairs or bare symmetries (dimension 1);
    multi-component symmetries must be given as tuples."""
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


def _row_major(shape):
    return itertools.product(*[range(s) for s in shape])


def _leadings(dims):
    lead = [1] * len(dims)
    for i in range(len(dims) - 2, -1, -1):
        lead[i] = lead[i + 1] * dims[i + 1]
    return lead


# ----------------------------------------------------------------------------
# Tensor  (structure/core.hpp:152-191 block layout, structure/tensor.hpp)
# ----------------------------------------------------------------------------
class Tensor:
    """Block-sparse tensor.  Blocks are the segment-position tuples whose
    charges sum to zero, enumerated row-major (last edge fastest) and stored
    contiguously, each block row-major (core.hpp:152-191)."""

    def __init__(self, kind, names, edges, dtype=np.float64, storage=None):
        self.kind = kind
        self.names = list(names)
        self.edges = [Edge.build(kind, e) for e in edges]
        if len(self.names) != len(self.edges):
            raise ValueError("names/edges length mismatch")
        if len(set(self.names)) != len(self.names):
            raise ValueError("duplicated names")
        self.dtype = np.dtype(dtype)
        self.block_table = {}  # positions -> (offset, shape)
        self.block_order = []
        offset = 0
        for pos in _row_major([len(e.segments) for e in self.edges]):
            total = kind.zero()
            for e, p in zip(self.edges, pos):
                total = kind.add(total, e.segments[p][0])
            if total == kind.zero():
                shape = tuple(e.segments[p][1] for e, p in zip(self.edges, pos))
                self.block_table[pos] = (offset, shape)
                self.block_order.append(pos)
                offset += int(np.prod(shape, dtype=np.int64))
        self.size = offset
        if storage is None:
            self.data = np.zeros(offset, dtype=self.dtype)
        else:
            storage = np.asarray(storage, dtype=self.dtype).reshape(-1)
            if storage.size != offset:
                raise ValueError("storage size mismatch")
            self.data = storage.copy()

    # -- construction helpers ------------------------------------------------
    @classmethod
    def scalar(cls, kind, number, names=(), symmetries=(), arrows=(), dtype=np.float64):
        """``Tensor(number, names, edge_symmetry, edge_arrow)``: one element,
        every edge ``[(q, 1)]`` (PyTAT.hpp:780-787, tensor.hpp one-element ctor)."""
        names = list(names)
        syms = list(symmetries) or [kind.zero()] * len(names)
        arrs = list(arrows) or [False] * len(names)
        edges = [Edge(kind, [(s, 1)], a) for s, a in zip(syms, arrs)]
        t = cls(kind, names, edges, dtype)
        if t.size != 1:
            raise ValueError("symmetries of one-element tensor must sum to zero")
        t.data[0] = number
        return t

    def copy(self):
        t = Tensor.__new__(Tensor)
        t.kind, t.names, t.edges, t.dtype = self.kind, list(self.names), list(self.edges), self.dtype
        t.block_table, t.block_order, t.size = self.block_table, self.block_order, self.size
        t.data = self.data.copy()
        return t

    def _shallow(self, names):
        t = Tensor.__new__(Tensor)
        t.kind, t.names, t.edges, t.dtype = self.kind, list(names), list(self.edges), self.dtype
        t.block_table, t.block_order, t.size = self.block_table, self.block_order, self.size
        t.data = self.data  # shared, like TAT's shared core
        return t

    @property
    def rank(self):
        return len(self.names)

    @property
    def storage(self):
        """Flat storage (numpy view); assignment writes in place."""
        return self.data

    @storage.setter
    def storage(self, value):
        self.data[...] = np.asarray(value, dtype=self.dtype).reshape(-1)

    @property
    def blocks(self):
        """``blocks[[(name, sym), ...]]`` -> block view with axes in the given
        name order (PyTAT.hpp:672-681)."""
        return _Blocks(self)

    def edge_by_name(self, name):
        return self.edges[self.names.index(name)]

    def zero_(self):
        self.data[:] = 0
        return self

    def range_(self, first=0, step=1):
        # tensor.hpp:584-593: fill in storage order
        self.data[:] = first + step * np.arange(self.size, dtype=self.dtype)
        return self

    def block(self, positions):
        off, shape = self.block_table[tuple(positions)]
        n = int(np.prod(shape, dtype=np.int64))
        return self.data[off:off + n].reshape(shape)

    def block_by_symmetries(self, mapping):
        """``blocks[{name: sym}]`` (PyTAT.hpp:672-681) -> numpy view in this
        tensor's edge order."""
        pos = []
        for n, e in zip(self.names, self.edges):
            q = self.kind.make(dict(mapping)[n])
            p = e.position_by_symmetry(q)
            if p is None:
                raise KeyError("no such symmetry on edge")
            pos.append(p)
        return self.block(pos)

    def _coords(self, mapping):
        pos, offs = [], []
        for n, e in zip(self.names, self.edges):
            v = mapping[n]
            if isinstance(v, tuple):
                q, off = v
                p = e.position_by_symmetry(self.kind.make(q))
                if p is None:
                    raise KeyError("no such symmetry on edge")
            else:
                p, off = e.coord_by_index(int(v))
            pos.append(p)
            offs.append(off)
        return tuple(pos), tuple(offs)

    def __getitem__(self, mapping):
        pos, offs = self._coords(mapping)
        return self.block(pos)[offs]

    def __setitem__(self, mapping, value):
        pos, offs = self._coords(mapping)
        self.block(pos)[offs] = value

    # -- norms / scalar arithmetic  (tensor.hpp:631-660, scalar.hpp) -----------
    def norm_max(self):
        return float(np.max(np.abs(self.data))) if self.size else 0.0

    def norm_2(self):
        return float(np.sqrt(np.sum(np.abs(self.data) ** 2)))

    def norm_sum(self):
        return float(np.sum(np.abs(self.data)))

    def norm_num(self):
        return float(self.size)

    def _zip(self, other, fn):
        t = self.copy()
        if isinstance(other, Tensor):
            # tensor.hpp:440-461: the other operand is transposed to our names
            if other.names != self.names and other.rank == self.rank:
                other = other.transpose(self.names)
            if other.size != self.size:
                raise ValueError("zip over tensors of different size")
            t.data = fn(self.data, other.data).astype(self.dtype)
        else:
            t.data = fn(self.data, other).astype(self.dtype)
        return t

    def __add__(self, o):
        return self._zip(o, np.add)

    def __sub__(self, o):
        return self._zip(o, np.subtract)

    def __mul__(self, o):
        return self._zip(o, np.multiply)

    def __truediv__(self, o):
        return self._zip(o, np.divide)

    __radd__ = __add__
    __rmul__ = __mul__

    def __float__(self):
        if self.size != 1:
            raise ValueError("not a one-element tensor")
        return float(self.data[0])

    # -- edge_rename / transpose  (tensor.hpp:760-797) ------------------------
    def edge_rename(self, mapping):
        """New names, SAME storage (no copy; PyTAT test_edge_rename.py:5-15)."""
        for k in mapping:
            if k not in self.names:
                raise KeyError("name missing in edge_rename")
        return self._shallow([mapping.get(n, n) for n in self.names])

    def transpose(self, target_names):
        if list(target_names) == self.names:
            return self._shallow(self.names)
        return self.edge_operator({}, set(), {}, list(target_names))

    def reverse_edge(self, reversed_names, apply_parity=False, parity_exclude_names=()):
        return self.edge_operator({}, set(reversed_names), {}, self.names, apply_parity,
                                  excl_rev_before=set(parity_exclude_names))

    def merge_edge(self, merge, apply_parity=False, parity_exclude_names_merge=(), parity_exclude_names_reverse=()):
        """split_and_merge.hpp:55-116: the merged edge goes where the last
        member of its group was; empty groups are appended."""
        target = []
        for name in reversed(self.names):
            found = False
            for new, olds in merge.items():
                if name in olds:
                    if name == olds[-1]:
                        target.append(new)
                    found = True
                    break
            if not found:
                target.append(name)
        for new, olds in merge.items():
            if len(olds) == 0:
                target.append(new)
        target.reverse()
        return self.edge_operator({}, set(), merge, target, apply_parity,
                                  excl_rev_after=set(parity_exclude_names_reverse),
                                  excl_merge=set(parity_exclude_names_merge))

    def split_edge(self, split, apply_parity=False, parity_exclude_names_split=()):
        """split_and_merge.hpp:28-52."""
        target = []
        for name in self.names:
            if name in split:
                target.extend(n for n, _ in split[name])
            else:
                target.append(name)
        return self.edge_operator(split, set(), {}, target, apply_parity, excl_split=set(parity_exclude_names_split))

    # -- the core: edge_operator  (implement/edge_operator.hpp:35-692) ---------
    def edge_operator(self, split_map, reversed_names, merge_map, new_names, apply_parity=False,
                      excl_split=(), excl_rev_before=(), excl_rev_after=(), excl_merge=(), cut=None):
        """Split -> reverse -> transpose -> reverse -> merge, with an optional
        per-segment cut applied first.  Sign of each destination block is the
        XOR of: reversed odd segments (edge_operator.hpp:497-520), inversions
        between odd edges in the transpose (:521-555, always applied), and
        ``(n_odd & 2) != 0`` per merged / split group (:591, :612), the
        reverse/merge/split parts gated by the apply_parity marks
        (:430-472)."""
        kind = self.kind
        fermi = kind.is_fermi
        excl_split, excl_rev_before = set(excl_split), set(excl_rev_before)
        excl_rev_after, excl_merge = set(excl_rev_after), set(excl_merge)
        cut = cut or {}

        # step 1: cut (:94-124)
        rank_bs = self.rank
        names_bs = self.names
        edges_bs = []
        for i, n in enumerate(names_bs):
            e = self.edges[i]
            if n in cut:
                segs = []
                for q, d in e.segments:
                    if q in cut[n]:
                        nd = cut[n][q]
                        if nd != 0:
                            segs.append((q, min(nd, d)))
                    else:
                        segs.append((q, d))
                edges_bs.append(Edge(kind, segs, e.arrow))
            else:
                edges_bs.append(e)

        # step 2: split (:126-218)
        names_as, split_flags, edges_as = [], [], []
        split_offsets = [None] * rank_bs
        for i in range(rank_bs):
            n = names_bs[i]
            if n in split_map:
                sub = split_map[n]
                sub_edges = []
                for sn, se in sub:
                    se = Edge.build(kind, se) if not isinstance(se, Edge) else se
                    names_as.append(sn)
                    split_flags.append(i)
                    ed = Edge(kind, se.segments, edges_bs[i].arrow)
                    edges_as.append(ed)
                    sub_edges.append(ed)
                table = {}
                bank = [0] * len(edges_bs[i].segments)
                for pos in _row_major([len(e.segments) for e in sub_edges]):
                    total, size, npar = kind.zero(), 1, 0
                    for e, p in zip(sub_edges, pos):
                        q, d = e.segments[p]
                        total = kind.add(total, q)
                        size *= d
                        npar += int(kind.parity(q))
                    position = edges_bs[i].position_by_symmetry(total)
                    if position is None:
                        table[pos] = None
                        continue
                    table[pos] = (position, bank[position], npar)
                    bank[position] += size
                split_offsets[i] = table
            else:
                names_as.append(n)
                split_flags.append(i)
                edges_as.append(edges_bs[i])
        rank_t = len(names_as)

        # step 3: reverse before transpose (:220-242)
        rev_before = [False] * rank_t
        edges_bt = list(edges_as)
        if fermi:
            for i in range(rank_t):
                if names_as[i] in reversed_names:
                    rev_before[i] = True
                    edges_bt[i] = Edge(kind, edges_as[i].segments, not edges_as[i].arrow)

        # step 4: names down to up, transpose plan (:244-303)
        names_am = list(new_names)
        rank_am = len(names_am)
        merge_flags, names_bm = [], []
        for j, mn in enumerate(names_am):
            if mn in merge_map:
                for old in merge_map[mn]:
                    names_bm.append(old)
                    merge_flags.append(j)
            else:
                names_bm.append(mn)
                merge_flags.append(j)
        if len(names_bm) != rank_t:
            raise ValueError("Tensor to transpose with different rank")
        idx_as = {n: i for i, n in enumerate(names_as)}
        d2s = []
        for n in names_bm:
            if n not in idx_as:
                raise KeyError("Tensor to transpose with incompatible name list")
            d2s.append(idx_as[n])
        edges_at = [edges_bt[d2s[i]] for i in range(rank_t)]

        # step 5: reverse before merge + merge plan (:305-413)
        edges_bm = list(edges_at)
        rev_after = [False] * rank_t
        result_edges = []
        merge_offsets = [None] * rank_am
        start = 0
        for j in range(rank_am):
            end = start
            while end < rank_t and merge_flags[end] == j:
                end += 1
            if merge_map and fermi:
                arrow = False if start == end else edges_bm[start].arrow
                for i in range(start, end):
                    if edges_bm[i].arrow != arrow:
                        edges_bm[i] = Edge(kind, edges_bm[i].segments, not edges_bm[i].arrow)
                        rev_after[i] = True
            else:
                arrow = edges_bm[start].arrow if end > start else False
            if merge_map and end != start + 1:
                sub_edges = edges_bm[start:end]
                table = {}
                merged = []  # [sym, dim]
                for pos in _row_major([len(e.segments) for e in sub_edges]):
                    total, size, npar = kind.zero(), 1, 0
                    for e, p in zip(sub_edges, pos):
                        q, d = e.segments[p]
                        total = kind.add(total, q)
                        size *= d
                        npar += int(kind.parity(q))
                    for k, (s, _) in enumerate(merged):
                        if s == total:
                            break
                    else:
                        k = len(merged)
                        merged.append([total, 0])
                    table[pos] = (k, merged[k][1], npar)
                    merged[k][1] += size
                result_edges.append(Edge(kind, [(s, d) for s, d in merged], arrow))
                merge_offsets[j] = table
            else:
                result_edges.append(Edge(kind, edges_bm[start].segments, edges_bm[start].arrow))
            start = end

        result = Tensor(kind, names_am, result_edges, self.dtype)

        # marks (:430-472): True => apply that half sign
        if apply_parity:
            mk_split = [n not in excl_split for n in names_bs]
            mk_rb = [n not in excl_rev_before for n in names_as]
            mk_ra = [n not in excl_rev_after for n in names_bm]
            mk_merge = [n not in excl_merge for n in names_am]
        else:
            mk_split = [n in excl_split for n in names_bs]
            mk_rb = [n in excl_rev_before for n in names_as]
            mk_ra = [n in excl_rev_after for n in names_bm]
            mk_merge = [n in excl_merge for n in names_am]

        # main loop (:474-690)
        rev_flag = [((rev_after[i] and mk_ra[i]) ^ (rev_before[d2s[i]] and mk_rb[d2s[i]])) for i in range(rank_t)]
        for pos_bm in _row_major([len(e.segments) for e in edges_bm]):
            total = kind.zero()
            par = False
            for i in range(rank_t):
                q = edges_bm[i].segments[pos_bm[i]][0]
                total = kind.add(total, q)
                if fermi and rev_flag[i] and kind.parity(q):
                    par ^= True
            if total != kind.zero():
                continue
            if fermi:
                odd = [kind.parity(edges_bm[i].segments[pos_bm[i]][0]) for i in range(rank_t)]
                for a in range(rank_t):
                    if not odd[a]:
                        continue
                    for b in range(a + 1, rank_t):
                        if odd[b] and d2s[a] > d2s[b]:
                            par ^= True
            pos_as = [0] * rank_t
            dims_bm = [0] * rank_t
            dims_as = [0] * rank_t
            for i in range(rank_t):
                p = pos_bm[i]
                pos_as[d2s[i]] = p
                d = edges_bm[i].segments[p][1]
                dims_bm[i] = d
                dims_as[d2s[i]] = d
            # merge side
            pos_am, off_am = [0] * rank_am, [0] * rank_am
            i = 0
            for j in range(rank_am):
                b = i
                while i < rank_t and merge_flags[i] == j:
                    i += 1
                if merge_offsets[j] is not None:
                    k, off, npar = merge_offsets[j][tuple(pos_bm[b:i])]
                    pos_am[j], off_am[j] = k, off
                    if (npar & 2) != 0 and mk_merge[j]:
                        par ^= True
                else:
                    pos_am[j], off_am[j] = pos_bm[b], 0
            # split side
            pos_bs, off_bs = [0] * rank_bs, [0] * rank_bs
            i = 0
            skip = False
            for j in range(rank_bs):
                b = i
                while i < rank_t and split_flags[i] == j:
                    i += 1
                if split_offsets[j] is not None:
                    entry = split_offsets[j][tuple(pos_as[b:i])]
                    if entry is None:
                        skip = True
                        break
                    k, off, npar = entry
                    pos_bs[j], off_bs[j] = k, off
                    if (npar & 2) != 0 and mk_split[j]:
                        par ^= True
                else:
                    pos_bs[j], off_bs[j] = pos_as[b], 0
            if skip:
                continue
            # positions in the un-cut source
            if cut:
                pos_src = []
                for j in range(rank_bs):
                    q = edges_bs[j].segments[pos_bs[j]][0]
                    pos_src.append(self.edges[j].position_by_symmetry(q))
            else:
                pos_src = pos_bs
            # leadings
            dst_dims_full = [result_edges[j].segments[pos_am[j]][1] for j in range(rank_am)]
            lead_am = _leadings(dst_dims_full)
            lead_bm = [0] * rank_t
            for i in range(rank_t - 1, -1, -1):
                if i != rank_t - 1 and merge_flags[i] == merge_flags[i + 1]:
                    lead_bm[i] = lead_bm[i + 1] * dims_bm[i + 1]
                else:
                    lead_bm[i] = lead_am[merge_flags[i]]
            src_dims_full = [self.edges[j].segments[pos_src[j]][1] for j in range(rank_bs)]
            lead_bs = _leadings(src_dims_full)
            lead_as = [0] * rank_t
            for i in range(rank_t - 1, -1, -1):
                if i != rank_t - 1 and split_flags[i] == split_flags[i + 1]:
                    lead_as[i] = lead_as[i + 1] * dims_as[i + 1]
                else:
                    lead_as[i] = lead_bs[split_flags[i]]
            if any(d == 0 for d in dims_bm):
                continue
            src_off = self.block_table[tuple(pos_src)][0] + sum(o * l for o, l in zip(off_bs, lead_bs))
            dst_off = result.block_table[tuple(pos_am)][0] + sum(o * l for o, l in zip(off_am, lead_am))
            # walk the destination order; dimension-one axes are dropped (as
            # mdspan_transform does, multidimension_span.hpp:322-357) so that
            # tensors with many one-dimensional legs stay within numpy's rank
            it = self.data.itemsize
            keep = [i for i in range(rank_t) if dims_bm[i] != 1]
            shape = [dims_bm[i] for i in keep]
            src_t = np.lib.stride_tricks.as_strided(self.data[src_off:], shape=shape,
                                                    strides=[lead_as[d2s[i]] * it for i in keep], writeable=False)
            dst_view = np.lib.stride_tricks.as_strided(result.data[dst_off:], shape=shape,
                                                       strides=[lead_bm[i] * it for i in keep])
            if par:
                dst_view[...] = -src_t
            else:
                dst_view[...] = src_t
        return result

    # -- contract  (implement/contract.hpp:307-620) -----------------------------
    def contract(self, other, pairs, fuse_names=()):
        if fuse_names:
            raise NotImplementedError("fuse is outside the restated path")
        return contract(self, other, pairs)

    # -- conjugate (implement/conjugate.hpp:32-120) ------------------------------
    def conjugate(self, trivial_metric=False):
        kind = self.kind
        if kind.length == 0:
            return self.copy()
        res = Tensor(kind, self.names, [e.conjugate() for e in self.edges], self.dtype)
        for pos in self.block_order:
            nodd, tm = 0, False
            if kind.is_fermi:
                for e, p in zip(self.edges, pos):
                    q = e.segments[p][0]
                    if kind.parity(q):
                        nodd += 1
                        tm ^= (e.arrow and trivial_metric)
            par = tm ^ ((nodd & 2) != 0)
            src = self.block(pos)
            res.block(pos)[...] = -np.conj(src) if par else np.conj(src)
        return res

    def svd(self, free_names_u, common_name_u, common_name_v, singular_name_u, singular_name_v, cut=-1):
        return svd(self, free_names_u, common_name_u, common_name_v, singular_name_u, singular_name_v, cut)

    def qr(self, direction, free_names, common_name_q, common_name_r):
        return qr(self, direction, free_names, common_name_q, common_name_r)

    def clear_symmetry(self):
        """Dense array in this tensor's name order, for No/Bose kinds
        (clear_symmetry.hpp); for fermi kinds the dense values are the raw
        block entries (sign conventions are the caller's business)."""
        dims = [e.total_dimension() for e in self.edges]
        out = np.zeros(dims, dtype=self.dtype)
        starts = []
        for e in self.edges:
            acc, s = 0, []
            for _, d in e.segments:
                s.append(acc)
                acc += d
            starts.append(s)
        for pos in self.block_order:
            sl = tuple(slice(starts[i][p], starts[i][p] + self.edges[i].segments[p][1]) for i, p in enumerate(pos))
            out[sl] = self.block(pos)
        return out

    def __repr__(self):
        return f"Tensor({self.kind.name}, {self.names}, {self.edges})"


# ----------------------------------------------------------------------------
# contract  (implement/contract.hpp:307-620)
# ----------------------------------------------------------------------------
def contract(t1, t2, pairs):
    kind = t1.kind
    fermi = kind.is_fermi
    pairs = set(pairs)
    c12 = {a: b for a, b in pairs}
    c21 = {b: a for a, b in pairs}
    for a, b in pairs:
        if a not in t1.names or b not in t2.names:
            raise KeyError("Name missing in contract")
        e1, e2 = t1.edge_by_name(a), t2.edge_by_name(b)
        if e1.conjugate() != e2 and not (not fermi and e1.conjugate().segments == e2.segments):
            raise ValueError("Incompatible edge segments in contract")
    for n1 in t1.names:
        if n1 in t2.names and n1 not in c12 and n1 not in c21:
            raise ValueError("Duplicated name in two contracting tensor but not fused or contracted")
    free1 = [n for n in t1.names if n not in c12]
    free2 = [n for n in t2.names if n not in c21]
    rev1, rev2, rev_result, common_rev1 = set(), set(), set(), set()
    split1, split2, names_result = [], [], []
    for i, n in enumerate(t1.names):
        e = t1.edges[i]
        if n not in c12:
            split1.append((n, Edge(kind, e.segments)))
            names_result.append(n)
            if fermi and e.arrow:
                rev1.add(n)
                rev_result.add(n)
        else:
            if fermi and not e.arrow:
                rev1.add(n)
                common_rev1.add(n)
    for i, n in enumerate(t2.names):
        e = t2.edges[i]
        if n not in c21:
            split2.append((n, Edge(kind, e.segments)))
            names_result.append(n)
            if fermi and e.arrow:
                rev2.add(n)
                rev_result.add(n)
        else:
            if fermi and e.arrow:
                rev2.add(n)

    common1, common2 = [], []

    def fit1():
        for n in t1.names:
            if n in c12:
                common1.append(n)
                common2.append(c12[n])

    def fit2():
        for n in t2.names:
            if n in c21:
                common2.append(n)
                common1.append(c21[n])

    # operand-order heuristic (:433-474)
    larger1 = t1.size > t2.size
    if len(free1) == 0:
        r1 = True
        fit2()
        r2 = (not common2) or common2[-1] == t2.names[-1]
    elif len(free2) == 0:
        r2 = True
        fit1()
        r1 = (not common1) or common1[-1] == t1.names[-1]
    elif larger1:
        if free1[-1] != t1.names[-1]:
            r1 = True
            fit1()
            r2 = (not common2) or common2[-1] == t2.names[-1]
        elif free2[-1] != t2.names[-1]:
            r2 = True
            fit2()
            r1 = (not common1) or common1[-1] == t1.names[-1]
        else:
            r1 = r2 = False
            fit1()
    else:
        if free2[-1] != t2.names[-1]:
            r2 = True
            fit2()
            r1 = (not common1) or common1[-1] == t1.names[-1]
        elif free1[-1] != t1.names[-1]:
            r1 = True
            fit1()
            r2 = (not common2) or common2[-1] == t2.names[-1]
        else:
            r1 = r2 = False
            fit2()

    C1, C2 = "\x00Contract_1", "\x00Contract_2"
    m1 = t1.edge_operator({}, rev1, {C1: free1, C2: common1}, [C1, C2] if r1 else [C2, C1], False,
                          excl_rev_before=common_rev1, excl_merge={C2})
    m2 = t2.edge_operator({}, rev2, {C2: free2, C1: common2}, [C2, C1] if r2 else [C1, C2], False)
    e0 = m1.edges[0 if r1 else 1]
    e1 = m2.edges[0 if r2 else 1]
    ecommon = m2.edges[1 if r2 else 0]
    prod = Tensor(kind, [C1, C2], [e0, e1], t1.dtype)
    for q, _ in e0.segments:
        p0 = e0.position_by_symmetry(q)
        p1 = e1.position_by_symmetry(kind.neg(q))
        if p1 is None:
            continue
        pc = ecommon.position_by_symmetry(q)
        C = prod.block((p0, p1))
        if pc is None:
            C[...] = 0
            continue
        A = m1.block((p0, pc) if r1 else (pc, p0))
        B = m2.block((p1, pc) if r2 else (pc, p1))
        A = A if r1 else A.T  # m x k
        B = B.T if r2 else B  # k x n
        alpha = -1.0 if (fermi and ((r1 ^ (not r2)) and kind.parity(q))) else 1.0
        if A.shape[1] == 0:
            C[...] = 0
        else:
            C[...] = alpha * (A @ B)
    return prod.edge_operator({C1: split1, C2: split2}, rev_result, {}, names_result, False)


# ----------------------------------------------------------------------------
# SVD  (implement/svd.hpp:260-538, truncation :430-481, S tensor :214-254)
# ----------------------------------------------------------------------------
def _cut_params(cut):
    """PyTAT.hpp:1040-1047: cut>=1 -> keep count, 0<cut<1 -> relative."""
    if cut is None or cut <= 0:
        return (1 << 62), 0.0
    if cut >= 1:
        return int(cut), 0.0
    return (1 << 62), float(cut)


def svd(t, free_names_u, common_name_u, common_name_v, singular_name_u, singular_name_v, cut=-1):
    kind = t.kind
    fermi = kind.is_fermi
    free_u = set(free_names_u)
    remain_cut, relative_cut = _cut_params(cut)
    put_v_right = (not t.names) or (t.names[-1] not in free_u)
    list_u, list_v, rev_in, rev_u, rev_v = [], [], set(), set(), set()
    res_u, res_v, fe_u, fe_v = [], [], [], []
    if put_v_right:
        res_v.append(common_name_v)
    else:
        res_u.append(common_name_u)
    for i, n in enumerate(t.names):
        e = t.edges[i]
        if n in free_u:
            list_u.append(n)
            res_u.append(n)
            fe_u.append((n, Edge(kind, e.segments)))
            if fermi and e.arrow:
                rev_u.add(n)
                rev_in.add(n)
        else:
            list_v.append(n)
            res_v.append(n)
            fe_v.append((n, Edge(kind, e.segments)))
            if fermi and e.arrow:
                rev_v.add(n)
                rev_in.add(n)
    if put_v_right:
        res_u.append(common_name_u)
    else:
        res_v.append(common_name_v)
    U_, V_ = "\x00SVD_U", "\x00SVD_V"
    merged = t.edge_operator({}, rev_in, {U_: list_u, V_: list_v}, [U_, V_] if put_v_right else [V_, U_], False)
    e0, e1 = merged.edges
    seg1, seg2 = [], []
    for q0, d0 in e0.segments:
        q1 = kind.neg(q0)
        d1 = e1.dimension_by_symmetry(q1)
        if d1 is not None:
            k = min(d0, d1)
            seg1.append((q1, k))
            seg2.append((q0, k))
    t1 = Tensor(kind, [U_, common_name_u] if put_v_right else [V_, common_name_v], [e0, Edge(kind, seg1)], t.dtype)
    t2 = Tensor(kind, [common_name_v, V_] if put_v_right else [common_name_u, U_], [Edge(kind, seg2), e1], t.dtype)
    ecommon2 = t2.edges[0]
    result_s = []  # [(sym, values)]
    for q, _ in e0.segments:
        p0 = e0.position_by_symmetry(q)
        p1 = e1.position_by_symmetry(kind.neg(q))
        if p1 is None:
            continue
        pc = ecommon2.position_by_symmetry(q)
        M = merged.block((p0, p1))
        m, n = M.shape
        k = min(m, n)
        s = np.zeros(k)
        if m * n != 0:
            u, s, vt = np.linalg.svd(M, full_matrices=False)
            t1.block((p0, pc))[...] = u
            t2.block((pc, p1))[...] = vt
        result_s.append((kind.neg(q) if put_v_right else q, np.array(s)))
    # truncation (:430-481)
    total_dim = sum(len(s) for _, s in result_s)
    smax = max([float(x) for _, s in result_s for x in s], default=0.0)
    threshold = relative_cut * smax
    rc = min(remain_cut, total_dim)
    remain = {q: 0 for q, _ in result_s}
    for _ in range(rc):
        best_q, best = None, 0.0
        for q, s in result_s:
            r = remain[q]
            if r != len(s) and s[r] > best:
                best, best_q = s[r], q
        if best > threshold:
            remain[best_q] += 1
        else:
            break
    result_s = [(q, s[:remain[q]]) for q, s in result_s if remain[q] != 0]
    remain_u = {q: remain[q] for q in remain}
    remain_v = {kind.neg(q): remain[q] for q in remain}
    tu, tv = (t1, t2) if put_v_right else (t2, t1)
    if fermi:
        rev_u.add(common_name_u)
    u = tu.edge_operator({U_: fe_u}, rev_u, {}, res_u, False, cut={common_name_u: remain_u})
    v = tv.edge_operator({V_: fe_v}, rev_v, {}, res_v, False, cut={common_name_v: remain_v})
    # S tensor (:214-254): edges (-q, false) / (q, true)
    s_t = Tensor(kind, [singular_name_u, singular_name_v],
                 [Edge(kind, [(kind.neg(q), len(s)) for q, s in result_s], False),
                  Edge(kind, [(q, len(s)) for q, s in result_s], True)], t.dtype)
    for q, s in result_s:
        p0 = s_t.edges[0].position_by_symmetry(kind.neg(q))
        p1 = s_t.edges[1].position_by_symmetry(q)
        blk = s_t.block((p0, p1))
        sign = -1.0 if (fermi and not put_v_right and kind.parity(q)) else 1.0
        blk[np.arange(len(s)), np.arange(len(s))] = sign * s
    return u, s_t, v


# ----------------------------------------------------------------------------
# QR  (implement/qr.hpp:310-508)
# ----------------------------------------------------------------------------
def qr(t, direction, free_names, common_name_q, common_name_r):
    kind = t.kind
    fermi = kind.is_fermi
    free = set(free_names)
    if direction in ("r", "R"):
        use_r = True
    elif direction in ("q", "Q"):
        use_r = False
    else:
        raise ValueError("Invalid direction in QR")
    use_qr = (not t.names) or ((t.names[-1] in free) == use_r)
    l1, l2, rev_in, rev1, rev2 = [], [], set(), set(), set()
    res1, res2, fe1, fe2 = [], [], [], []
    res2.append(common_name_r if use_qr else common_name_q)
    for i, n in enumerate(t.names):
        e = t.edges[i]
        if ((n in free) == use_r) == use_qr:
            l2.append(n)
            res2.append(n)
            fe2.append((n, Edge(kind, e.segments)))
            if fermi and e.arrow:
                rev2.add(n)
                rev_in.add(n)
        else:
            l1.append(n)
            res1.append(n)
            fe1.append((n, Edge(kind, e.segments)))
            if fermi and e.arrow:
                rev1.add(n)
                rev_in.add(n)
    res1.append(common_name_q if use_qr else common_name_r)
    Q1, Q2 = "\x00QR_1", "\x00QR_2"
    merged = t.edge_operator({}, rev_in, {Q1: l1, Q2: l2}, [Q1, Q2], False)
    e0, e1 = merged.edges
    seg1, seg2 = [], []
    for q0, d0 in e0.segments:
        q1 = kind.neg(q0)
        d1 = e1.dimension_by_symmetry(q1)
        if d1 is not None:
            k = min(d0, d1)
            seg1.append((q1, k))
            seg2.append((q0, k))
    t1 = Tensor(kind, [Q1, common_name_q if use_qr else common_name_r], [e0, Edge(kind, seg1)], t.dtype)
    t2 = Tensor(kind, [common_name_r if use_qr else common_name_q, Q2], [Edge(kind, seg2), e1], t.dtype)
    for q, _ in e0.segments:
        p0 = e0.position_by_symmetry(q)
        p1 = e1.position_by_symmetry(kind.neg(q))
        if p1 is None:
            continue
        pc = t2.edges[0].position_by_symmetry(q)
        M = merged.block((p0, p1))
        m, n = M.shape
        if m * n == 0:
            continue
        if use_qr:
            # M (m x n) = T1 (m x k, orthonormal cols) . T2 (k x n)
            qm, rm = np.linalg.qr(M, mode="reduced")
            t1.block((p0, pc))[...] = qm
            t2.block((pc, p1))[...] = rm
        else:
            # M = T1 (m x k) . T2 (k x n, orthonormal rows)  (LQ)
            qm, rm = np.linalg.qr(M.T, mode="reduced")
            t1.block((p0, pc))[...] = rm.T
            t2.block((pc, p1))[...] = qm.T
    if fermi:
        (rev1 if use_qr else rev2).add(common_name_q)
    n1 = t1.edge_operator({Q1: fe1}, rev1, {}, res1, False)
    n2 = t2.edge_operator({Q2: fe2}, rev2, {}, res2, False,
                          excl_rev_before=set() if use_qr else {common_name_q})
    return (n1, n2) if use_qr else (n2, n1)
