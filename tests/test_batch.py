"""Host side of the batched operations: plans made for many items on host
threads (tns_plan_*_many) equal the plans made one at a time — same output
structures, workspace sizes and flop counts — and planning errors of any
item surface as an error of the whole call."""
import numpy as np
import pytest

KINDS = ["FermiU1", "FermiZ2", "BoseU1"]


def shapes(kind, n, seed=0):
    from tnsp_amd.shape import KINDS as K, Shape

    rng = np.random.default_rng(seed)
    charges = [0, 1] if kind == "FermiZ2" else [-1, 0, 1]
    sa, sb = [], []
    for _ in range(n):
        e = [[((c,), int(rng.integers(1, 6))) for c in charges] for _ in range(6)]
        a = Shape.create(K[kind], ["i", "j", "k", "l"], [(e[0], False), (e[1], True), (e[2], False), (e[3], True)])
        conj = [[((-c if kind != "FermiZ2" else c),), d] for (c,), d in e[2]]
        conj3 = [[((-c if kind != "FermiZ2" else c),), d] for (c,), d in e[3]]
        b = Shape.create(K[kind], ["k", "l", "m", "n"], [([tuple(x) for x in conj], True),
                                                         ([tuple(x) for x in conj3], False), (e[4], True),
                                                         (e[5], False)])
        sa.append(a)
        sb.append(b)
    return sa, sb


def same_plan(handle, q, nout):
    """Plan ``handle`` (owned by a batch) against the single plan ``q``."""
    import ctypes

    from tnsp_amd import _lib
    from tnsp_amd.plan import DTYPE_CODE
    from tnsp_amd.shape import Shape

    L = _lib.lib()
    h = ctypes.c_void_p(int(handle))
    for w in range(nout):
        o = ctypes.c_void_p()
        _lib.check(L.tns_plan_output_shape(h, w, ctypes.byref(o)))
        x, y = Shape(o, q.kind), q.output(w)
        assert x.names == y.names
        assert x.edges == y.edges
        assert x.size == y.size
    for dt in ("float64", "float32"):
        b = ctypes.c_int64()
        _lib.check(L.tns_plan_workspace(h, DTYPE_CODE[dt], ctypes.byref(b)))
        assert b.value == q.workspace(dt)
    f = ctypes.c_int64()
    _lib.check(L.tns_plan_flops(h, ctypes.byref(f)))
    assert f.value == q.flops()


def handles(shapes):
    return [s.handle.value for s in shapes]


@pytest.mark.parametrize("kind", KINDS)
def test_contract_many_equals_single(kind):
    from tnsp_amd import batch, plan

    sa, sb = shapes(kind, 9)
    pairs = [("k", "k"), ("l", "l")]
    many = batch.plan_contract(handles(sa), handles(sb), pairs)
    for a, b, p in zip(sa, sb, many.arr):
        same_plan(p, plan.contract(a, b, pairs), 1)


@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("direction", ["r", "q"])
def test_qr_many_equals_single(kind, direction):
    from tnsp_amd import batch, plan

    sa, _ = shapes(kind, 7, seed=2)
    many = batch.plan_qr(handles(sa), direction, ["i", "k"], "cq", "cr")
    for a, p in zip(sa, many.arr):
        same_plan(p, plan.qr(a, direction, ["i", "k"], "cq", "cr"), 2)


@pytest.mark.parametrize("kind", KINDS)
def test_svd_many_equals_single(kind):
    from tnsp_amd import batch, plan

    sa, _ = shapes(kind, 7, seed=4)
    many = batch.plan_svd(handles(sa), ["j"], "cu", "cv", "su", "sv")
    for a, p in zip(sa, many.arr):
        # output structures depend on the kept singular values: known after the
        # factor phase on the device
        same_plan(p, plan.svd(a, ["j"], "cu", "cv", "su", "sv"), 0)


def test_many_reports_item_error():
    from tnsp_amd import batch

    sa, sb = shapes("BoseU1", 5, seed=6)
    # item 3 contracts two edges of equal arrow: not conjugate, rejected
    sb[3] = sa[3].rename({"i": "m", "j": "n"})
    with pytest.raises(RuntimeError):
        batch.plan_contract(handles(sa), handles(sb), [("k", "k"), ("l", "l")])
