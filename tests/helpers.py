"""Shared test helpers working on both the oracle facade and the GPU package."""
import itertools

import numpy as np


def storage(t):
    return np.asarray(t.storage, dtype=np.float64)


def to_dense(t):
    """Dense array in ``t.names`` order, blocks placed at their segment
    offsets (TAT's clear_symmetry layout for bose tensors)."""
    edges = t.edges
    dims = [sum(d for _, d in e.segments) for e in edges]
    out = np.zeros(dims)
    starts = []
    for e in edges:
        acc, s = 0, []
        for _, d in e.segments:
            s.append(acc)
            acc += d
        starts.append(s)
    names = t.names
    for pos in itertools.product(*[range(len(e.segments)) for e in edges]):
        syms = [edges[i].segments[p][0] for i, p in enumerate(pos)]
        total = np.zeros(len(syms[0]) if syms and len(syms[0]) else 0, dtype=np.int64)
        ok = True
        if syms and len(syms[0]):
            kinds = getattr(t, "model", None)
            comps = (kinds.kind.comps if kinds is not None and hasattr(kinds, "kind") else t.kind.comps)
            acc = [0] * len(comps)
            for q in syms:
                for c, (z2, _) in enumerate(comps):
                    acc[c] = (acc[c] ^ int(q[c])) if z2 else acc[c] + int(q[c])
            ok = all(a == 0 for a in acc)
        if not ok:
            continue
        if any(edges[i].segments[p][1] == 0 for i, p in enumerate(pos)):
            continue
        blk = np.asarray(t.blocks[[(names[i], syms[i]) for i in range(len(names))]], dtype=np.float64)
        sl = tuple(slice(starts[i][p], starts[i][p] + edges[i].segments[p][1]) for i, p in enumerate(pos))
        out[sl] = blk
    return out


def check_unitary(tensor, name, name_prime, fermi, tol=1e-6):
    """PyTAT/tests/test_svd.py:5-16 restated: T^dagger T over every edge but
    ``name`` is the identity on (name, name') -- up to signs for fermions."""
    pairs = {(n, n) for n in tensor.names if n != name}
    conjugated = tensor.conjugate(True).edge_rename({name: name_prime})
    product = tensor.contract(conjugated, pairs).transpose([name, name_prime])
    d = to_dense(product)
    if fermi:
        d = np.abs(d)
    assert d.shape[0] == d.shape[1]
    assert np.max(np.abs(d - np.eye(d.shape[0]))) < tol


def c1_pair(tat, kind_name, seed):
    """Config-1 tensors A[i,j,k,l], B[k,l,m,n] (every edge dimension 8 as
    {-1:2, 0:4, +1:2}; FermiZ2: {even:4, odd:4}); B's k, l are the
    conjugates of A's.  ``tat`` is the oracle module; fixtures record the
    seed and the storage digest so the generator is pinned too."""
    K = tat.KINDS[kind_name]
    rng = np.random.default_rng(seed)
    e = [(0, 4), (1, 4)] if kind_name == "FermiZ2" else [(-1, 2), (0, 4), (1, 2)]
    A = tat.Tensor(K, ["i", "j", "k", "l"], [(e, False), (e, True), (e, False), (e, True)])
    A.data[:] = rng.standard_normal(A.size)
    B = tat.Tensor(K, ["k", "l", "m", "n"], [A.edges[2].conjugate(), A.edges[3].conjugate(), (e, True), (e, False)])
    B.data[:] = rng.standard_normal(B.size)
    return A, B


def digest(data):
    import hashlib

    return hashlib.sha256(np.ascontiguousarray(np.asarray(data), dtype=np.float64).tobytes()).hexdigest()


def load_golden():
    import json
    import os

    here = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    ops = json.load(open(os.path.join(here, "c1_ops.json")))
    arrays = dict(np.load(os.path.join(here, "c1_ops.npz"), allow_pickle=False))
    amps = json.load(open(os.path.join(here, "c2_amplitudes.json")))
    return ops, arrays, amps


def load_golden_8x8():
    import json
    import os

    here = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    return json.load(open(os.path.join(here, "peps8x8_amplitudes.json")))


def load_golden_zero():
    import json
    import os

    here = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    return json.load(open(os.path.join(here, "zero_amplitudes.json")))


def to_gpu(t, dtype="float64"):
    """Copy an oracle tensor (structure + storage) into a GPU tensor."""
    from tnsp_amd import TAT

    mod = getattr(TAT, t.kind.name)
    g = (mod.D if dtype == "float64" else mod.S).Tensor(t.names, [(list(e.segments), e.arrow) for e in t.edges])
    g.storage = t.data
    return g


def assert_structure(t, ref):
    """names, segments and arrows equal to a fixture entry (bit-exact)."""
    assert list(t.names) == ref["names"]
    segs = [[[list(q), int(d)] for q, d in e.segments] for e in t.edges]
    assert segs == ref["edges"]
    assert [bool(e.arrow) for e in t.edges] == ref["arrows"]
