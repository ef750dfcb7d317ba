"""GPU parity of the batched operations (tnsp_amd.batch): n tensors with one
name list but different block structures go through one planned launch per
kernel kind; every item must equal the oracle's single-tensor result to the
bar of the single-tensor tests (structure exact, values 1e-12 / 1e-10)."""
import numpy as np
import pytest

from helpers import storage
from test_gpu_ops import rel, same_structure, to_gpu

pytestmark = pytest.mark.gpu


def _edge(kind, rng, lo=1, hi=4):
    charges = [0, 1] if kind == "FermiZ2" else [-1, 0, 1]
    return [(c, int(rng.integers(lo, hi + 1))) for c in charges]


def items(kind, n, seed=3):
    """n pairs A[i,j,k,l], B[k,l,m,n] whose segment dimensions differ per
    item (item 0 is all ones: the smallest sectors)."""
    from oracle import tat

    K = tat.KINDS[kind]
    rng = np.random.default_rng(seed)
    out = []
    for it in range(n):
        hi = 1 if it == 0 else 5
        e = [_edge(kind, rng, 1, hi) for _ in range(6)]
        A = tat.Tensor(K, ["i", "j", "k", "l"], [(e[0], False), (e[1], True), (e[2], False), (e[3], True)])
        A.data[:] = rng.standard_normal(A.size)
        B = tat.Tensor(K, ["k", "l", "m", "n"], [A.edges[2].conjugate(), A.edges[3].conjugate(), (e[4], True),
                                                 (e[5], False)])
        B.data[:] = rng.standard_normal(B.size)
        out.append((A, B))
    return out


KINDS = ["FermiU1", "FermiZ2", "BoseU1"]


@pytest.mark.parametrize("kind", KINDS)
def test_batch_contract(kind):
    from tnsp_amd import batch

    pairs = items(kind, 7)
    ga = batch.TensorBatch.from_tensors([to_gpu(a) for a, _ in pairs])
    gb = batch.TensorBatch.from_tensors([to_gpu(b) for _, b in pairs])
    # ("z", "z") names an edge on neither side: dropped, as safe_contract does
    res = batch.contract(ga, gb, [("k", "k"), ("l", "l"), ("z", "z")])
    assert len(res) == len(pairs)
    for i, (a, b) in enumerate(pairs):
        o = a.contract(b, {("k", "k"), ("l", "l")})
        g = res.tensor(i)
        same_structure(g, o)
        assert rel(storage(g), o.data) < 1e-12, i


@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("direction,free", [("r", ["m", "n"]), ("q", ["i"])])
def test_batch_qr(kind, direction, free):
    from tnsp_amd import batch

    pairs = items(kind, 6, seed=5)
    cs = [a.contract(b, {("k", "k"), ("l", "l")}) for a, b in pairs]
    q, r = batch.qr(batch.TensorBatch.from_tensors([to_gpu(c) for c in cs]), direction, free, "cq", "cr")
    for i, c in enumerate(cs):
        oq, orr = c.qr(direction, set(free), "cq", "cr")
        same_structure(q.tensor(i), oq)
        same_structure(r.tensor(i), orr)
        back = q.tensor(i).contract(r.tensor(i), {("cq", "cr")}).transpose(c.names)
        assert rel(storage(back), c.data) < 1e-10, i


@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("cut", [-1, 6])
def test_batch_svd(kind, cut):
    from tnsp_amd import batch

    pairs = items(kind, 6, seed=9)
    cs = [a.contract(b, {("k", "k"), ("l", "l")}) for a, b in pairs]
    u, s, v = batch.svd(batch.TensorBatch.from_tensors([to_gpu(c) for c in cs]), ["i", "m"], "cu", "cv", "su", "sv",
                        cut)
    for i, c in enumerate(cs):
        ou, os_, ov = c.svd({"i", "m"}, "cu", "cv", "su", "sv", cut)
        for g, o in ((u.tensor(i), ou), (s.tensor(i), os_), (v.tensor(i), ov)):
            same_structure(g, o)
        assert rel(storage(s.tensor(i)), os_.data) < 1e-10, i
        r0 = ou.contract(os_, {("cu", "su")}).contract(ov, {("sv", "cv")})
        r1 = u.tensor(i).contract(s.tensor(i), {("cu", "su")}).contract(v.tensor(i), {("sv", "cv")})
        same_structure(r1, r0)
        assert rel(storage(r1), r0.data) < 1e-10, i


def test_batch_rename_take_select_values():
    from tnsp_amd import batch

    pairs = items("FermiU1", 4, seed=11)
    ga = batch.TensorBatch.from_tensors([to_gpu(a) for a, _ in pairs])
    rn = ga.rename({"i": "x", "absent": "y"})
    assert rn.names == ["x", "j", "k", "l"]
    for i, (a, _) in enumerate(pairs):
        same_structure(rn.tensor(i), a.edge_rename({"i": "x"}))
        assert np.array_equal(storage(rn.tensor(i)), a.data)
    sub = ga.take([3, 1])
    assert np.array_equal(storage(sub.tensor(0)), pairs[3][0].data)
    assert np.array_equal(storage(sub.tensor(1)), pairs[1][0].data)
    # select: item i is choices[which[i]]
    sel = batch.TensorBatch.select([ga.tensor(0), ga.tensor(2)], [1, 0, 1])
    for i, w in enumerate([2, 0, 2]):
        assert np.array_equal(storage(sel.tensor(i)), pairs[w][0].data)
    # full contraction of each A with its conjugate: one-element tensors
    ca = batch.TensorBatch.from_tensors([to_gpu(a.conjugate()) for a, _ in pairs])
    full = batch.contract(ga, ca, [(n, n) for n in "ijkl"])
    want = [a.contract(a.conjugate(), {(n, n) for n in "ijkl"}).data[0] for a, _ in pairs]
    assert np.allclose(full.values(), want, rtol=1e-12, atol=0)


@pytest.mark.parametrize("kind", KINDS)
def test_batch_conjugate_and_scale(kind):
    from tnsp_amd import batch

    pairs = items(kind, 5, seed=13)
    ga = batch.TensorBatch.from_tensors([to_gpu(a) for a, _ in pairs])
    for tm in (False, True):
        cj = batch.conjugate(ga, trivial_metric=tm)
        for i, (a, _) in enumerate(pairs):
            o = a.conjugate(tm)
            same_structure(cj.tensor(i), o)
            assert np.array_equal(storage(cj.tensor(i)), o.data), (i, tm)
    f = np.arange(1, len(pairs) + 1, dtype=np.float64) * 0.5
    sc = batch.scale(ga, f)
    for i, (a, _) in enumerate(pairs):
        assert np.array_equal(storage(sc.tensor(i)), a.data * f[i])
