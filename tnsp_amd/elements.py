"""The sparse element table of a bond Hamiltonian, which drives the samplers.

tetragono's ``tensor_element`` (tetragono/tetragono/tensor_element.py:20-111,
paths relative to the TNSP repository root) turns an operator tensor with
edges ``I0 .. I{b-1}``, ``O0 .. O{b-1}`` into the table

    {in points: {out points: element}}

over its non-zero elements, where the in points are the conjugated
symmetries of the ``I`` edges and the out points those of the ``O`` edges.
``SweepSampling._single_term`` (sampling_lattice/sampling.py:117-143) reads
it: the current occupations of the bond select the in points, one target is
drawn uniformly among the out points, and the acceptance carries the factor
``n / n'`` of the numbers of targets from the configuration and from the
target.

The samplers here hold a configuration as one occupation per site, so the
table is keyed by occupations: the first component of each point's
symmetry (the particle number for FermiU1, the parity for FermiZ2), with
every physical segment of dimension one.  Targets keep tetragono's order:
blocks in lexicographic order of their segment positions (last edge
fastest), elements in row-major order inside a block.
"""
import itertools

import numpy as np


def element_table(kind, names, edges, blocks):
    """The table of an operator given as host data.  ``kind``: a
    ``shape.Kind``; ``names``: edge names; ``edges``: ``shape.Edge`` list;
    ``blocks``: ``[(segment positions, ndarray)]`` of its blocks (any
    order).  Values are Python floats."""
    rank = len(names)
    if rank % 2:
        raise ValueError("an operator has as many I edges as O edges")
    body = rank // 2
    where = {n: i for i, n in enumerate(names)}
    for i in range(body):
        if f"I{i}" not in where or f"O{i}" not in where:
            raise ValueError(f"operator edges must be I0..I{body - 1} and O0..O{body - 1}, got {names}")
    table = {}
    for pos, block in sorted(blocks, key=lambda pb: tuple(pb[0])):
        syms = [edges[i].segments[p][0] for i, p in enumerate(pos)]
        block = np.asarray(block)
        for idx in itertools.product(*(range(d) for d in block.shape)):
            value = float(block[idx])
            if value == 0:
                continue
            if any(idx[where[f"{c}{i}"]] != 0 for c in "IO" for i in range(body)):
                raise ValueError("the samplers need physical segments of dimension one (one point per occupation)")
            key_in = tuple(int(kind.neg(syms[where[f"I{i}"]])[0]) for i in range(body))
            key_out = tuple(int(syms[where[f"O{i}"]][0]) for i in range(body))
            table.setdefault(key_in, {})[key_out] = value
    return table


def tensor_element(tensor):
    """``tensor_element`` of a drop-in ``Tensor`` (its storage is read back
    to the host once)."""
    shape = tensor._shape
    data = tensor._data.detach().cpu().numpy()
    blocks = []
    for pos, off, shp in shape.blocks():
        n = int(np.prod(shp, dtype=np.int64))
        blocks.append((pos, data[off:off + n].reshape(shp)))
    return element_table(shape.kind, tensor.names, tensor.edges, blocks)


def tv_table(t=1.0, V=1.0):
    """The table's structure for the tV bond Hamiltonian H = -t CC + V NN
    (tetraku/tetraku/models/tV/__init__.py:39-45): a hop between a filled
    and an empty site when t != 0, the diagonal of a doubly occupied bond
    when V != 0.  The values are the bare couplings (the fermion signs of
    the hop elements depend on the operator's edge order and are applied by
    the observer)."""
    table = {}
    if t != 0:
        table[(0, 1)] = {(1, 0): -float(t)}
        table[(1, 0)] = {(0, 1): -float(t)}
    if V != 0:
        table[(1, 1)] = {(1, 1): float(V)}
    return table


class Moves:
    """A two-site table prepared for lockstep sampling: for every in key its
    targets, their count n, and for every target the count n' of its own
    targets."""

    def __init__(self, table):
        for k in table:
            if len(k) != 2:
                raise ValueError("the samplers take two-site (bond) operators")
        self.table = table
        self.targets = {k: list(v) for k, v in table.items() if v}
        for k, outs in self.targets.items():
            for o in outs:
                if o not in self.targets:
                    raise ValueError(f"target {o} of {k} has no elements of its own (n' = 0)")
        self.multi = any(len(v) > 1 for v in self.targets.values())

    def draw(self, na, nb, v):
        """Targets for chains with occupations (na, nb) on the bond: returns
        (active, ta, tb, n, n'), ``active`` marking chains whose occupations
        have elements; ``v`` (one uniform per chain, used only where a key
        has several targets) picks the target as tetragono's uniform_int
        does."""
        B = len(na)
        active = np.zeros(B, dtype=bool)
        ta, tb = na.copy(), nb.copy()
        n = np.ones(B)
        ns = np.ones(B)
        for k, outs in self.targets.items():
            sel = np.nonzero((na == k[0]) & (nb == k[1]))[0]
            if len(sel) == 0:
                continue
            active[sel] = True
            m = len(outs)
            j = np.zeros(len(sel), dtype=np.int64) if m == 1 else np.minimum((v[sel] * m).astype(np.int64), m - 1)
            arr = np.array(outs, dtype=na.dtype)
            ta[sel], tb[sel] = arr[j, 0], arr[j, 1]
            n[sel] = m
            ns[sel] = np.array([len(self.targets[o]) for o in outs], dtype=np.float64)[j]
        return active, ta, tb, n, ns
