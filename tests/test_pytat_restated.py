"""The reference binding's own tests (PyTAT/tests/*.py), restated for the
operations on the hot path, run against the oracle facade (CPU) and against
the GPU package (``-m gpu``).  Each test names the reference test it follows.
"""
import itertools

import numpy as np
import pytest

from helpers import check_unitary, storage, to_dense


def Fedge(*segs):
    return (list(segs), False)


def Tedge(*segs):
    return (list(segs), True)


# ---------------------------------------------------------------- creation
def test_create_symmetry_basic(TAT):
    """test_create_symmetry_tensor.py:5-28"""
    a = TAT.BoseZ2.D.Tensor(["Left", "Right", "Up"],
                            [[(True, 3), (False, 1)], [(True, 1), (False, 2)], [(True, 2), (False, 3)]]).range_()
    assert a.names == ["Left", "Right", "Up"]
    assert a.rank == 3
    assert storage(a).size == 1 * 2 * 3 + 1 * 1 * 2 + 3 * 2 * 2 + 3 * 1 * 3
    assert np.asarray(a.blocks[[("Left", True), ("Right", False), ("Up", True)]]).shape == (3, 2, 2)
    assert np.asarray(a.blocks[[("Left", False), ("Right", True), ("Up", True)]]).shape == (1, 1, 2)
    assert a[{"Left": (True, 2), "Right": (False, 0), "Up": (True, 1)}] == 3 * 1 * 3 + 9
    assert a[{"Left": 2, "Right": 1, "Up": 1}] == 3 * 1 * 3 + 9
    assert a[{"Left": (False, 0), "Right": (False, 1), "Up": (False, 2)}] == 3 * 1 * 3 + 3 * 2 * 2 + 1 * 1 * 2 + 5
    assert a[{"Left": 3, "Right": 2, "Up": 4}] == 3 * 1 * 3 + 3 * 2 * 2 + 1 * 1 * 2 + 5


def test_create_fermi_basic(TAT):
    """test_create_fermi_tensor.py:19-41"""
    a = TAT.FermiZ2.D.Tensor(["Left", "Right", "Up"], [
        Tedge((True, 3), (False, 1)),
        Fedge((True, 1), (False, 2)),
        Tedge((True, 2), (False, 3)),
    ]).range_()
    assert [e.arrow for e in a.edges] == [True, False, True]
    assert np.asarray(a.blocks[("Left", True), ("Right", False), ("Up", True)]).shape == (3, 2, 2)
    assert np.asarray(a.blocks[("Left", False), ("Up", True), ("Right", True)]).shape == (1, 2, 1)
    assert a[{"Left": (True, 2), "Right": (False, 0), "Up": (True, 1)}] == 3 * 1 * 3 + 9
    assert a[{"Left": (False, 0), "Right": (False, 1), "Up": (False, 2)}] == 3 * 1 * 3 + 3 * 2 * 2 + 1 * 1 * 2 + 5


def test_create_rank0_and_empty(TAT):
    """test_create_fermi_tensor.py:44-80"""
    a = TAT.FermiZ2.D.Tensor([], []).range_(2333)
    assert a.names == []
    assert list(storage(a)) == [2333]
    assert float(a) == 2333
    b = TAT.FermiU1.D.Tensor(["Left", "Right", "Up"], [
        Fedge((0, 0)),
        Tedge((-1, 1), (0, 2), (1, 3)),
        Tedge((-1, 2), (0, 3), (1, 1)),
    ]).zero_()
    assert storage(b).size == 0
    assert np.asarray(b.blocks[("Left", 0), ("Right", +1), ("Up", -1)]).shape == (0, 3, 2)
    c = TAT.FermiU1.D.Tensor(["Left", "Right", "Up"], [
        Fedge(),
        Fedge((-1, 1), (0, 2), (1, 3)),
        Tedge((-1, 2), (0, 3), (1, 1)),
    ]).zero_()
    assert storage(c).size == 0


def test_edge_rename_shares_storage(TAT):
    """test_edge_rename.py:5-15: renaming does not copy the storage."""
    t1 = TAT.BoseZ2.D.Tensor(["Left", "Right", "Phy"], [
        [(False, 1), (True, 2)], [(False, 3), (True, 4)], [(False, 5), (True, 6)]]).range_()
    t2 = t1.edge_rename({"Left": "Up"})
    assert t1.names == ["Left", "Right", "Phy"]
    assert t2.names == ["Up", "Right", "Phy"]
    t1.storage = np.arange(storage(t1).size) * 2.0
    assert np.array_equal(storage(t2), np.arange(storage(t1).size) * 2.0)


# ---------------------------------------------------------------- transpose
def test_transpose_no_symmetry(TAT):
    """test_transpose.py:5-25"""
    a = TAT.No.D.Tensor(["Left", "Right"], [2, 3]).range_()
    b = a.transpose(["Right", "Left"])
    assert list(storage(a)) == [0, 1, 2, 3, 4, 5]
    assert list(storage(b)) == [0, 3, 1, 4, 2, 5]
    a = TAT.No.D.Tensor(["i", "j", "k"], [2, 3, 4]).range_()
    da = to_dense(a)
    for order in itertools.permutations("ijk"):
        b = a.transpose(list(order))
        assert np.array_equal(to_dense(b), da.transpose(["ijk".index(c) for c in order]))


def test_transpose_parity_high_dimension(TAT):
    """test_transpose.py:90-140: sign = inversions between odd edges."""
    edge = TAT.FermiZ2.Edge([(False, 2), (True, 2)])
    a = TAT.FermiZ2.D.Tensor(["i", "j", "k", "l", "m", "n"], [edge] * 6).range_()
    b = a.transpose(["l", "j", "i", "n", "k", "m"])
    da, db = to_dense(a), to_dense(b)
    order = "ljinkm"
    for idx in itertools.product(range(4), repeat=6):
        p = [(v & 2) != 0 for v in idx]
        if p[0] ^ p[1] ^ p[2] ^ p[3] ^ p[4] ^ p[5]:
            continue
        pi, pj, pk, pl, pm, pn = p
        parity = (pl and (pi ^ pj ^ pk)) ^ (pj and pi) ^ (pn and (pk ^ pm))
        va = da[idx]
        vb = db[tuple(idx["ijklmn".index(c)] for c in order)]
        assert vb == (-va if parity else va)


def test_transpose_z2_high_dimension(TAT):
    """test_transpose.py:55-88"""
    edge = TAT.BoseZ2.Edge([(False, 2), (True, 2)])
    a = TAT.BoseZ2.D.Tensor(["i", "j", "k", "l", "m", "n"], [edge] * 6).range_()
    b = a.transpose(["l", "j", "i", "n", "k", "m"])
    assert np.array_equal(to_dense(b), to_dense(a).transpose(["ijklmn".index(c) for c in "ljinkm"]))


# ---------------------------------------------------------------- merge/split
def test_split_merge_no_symmetry(TAT):
    """test_split_and_merge.py:5-16"""
    a = TAT.No.D.Tensor(["Left", "Right"], [2, 3]).range_()
    b = a.merge_edge({"Merged": ["Left", "Right"]})
    assert np.array_equal(storage(b), storage(a))
    assert (b.split_edge({"Merged": [("Left", 2), ("Right", 3)]}) - a).norm_max() == 0
    c = a.merge_edge({"Merged": ["Right", "Left"]})
    assert np.array_equal(storage(c), storage(a.transpose(["Right", "Left"])))
    assert (c.split_edge({"Merged": [("Right", 3), ("Left", 2)]}) - a).norm_max() == 0


def test_split_merge_u1_basic(TAT):
    """test_split_and_merge.py:31-39"""
    a = TAT.BoseU1.D.Tensor(["i", "j"], [[-1, 0, +1], [-1, 0, +1]]).range_()
    d = TAT.BoseU1.D.Tensor(["m"], [[(-2, 1), (-1, 2), (0, 3), (+1, 2), (+2, 1)]]).range_()
    b = a.merge_edge({"m": ["i", "j"]})
    assert (d - b).norm_max() == 0
    c = b.split_edge({"m": [("i", [-1, 0, +1]), ("j", [-1, 0, +1])]})
    assert (c - a).norm_max() == 0


def test_split_merge_fermi_round_trip(TAT):
    """test_split_and_merge.py:53-63"""
    edge = [(-1, 2), (0, 2), (+1, 2)]
    a = TAT.FermiU1.D.Tensor(["1", "2", "3", "4", "5"], [edge] * 5).range_()
    for i in range(5):
        for j in range(i, 5):
            for p in (False, True):
                names = a.names[i:j]
                b = a.merge_edge({"m": names}, p)
                c = b.split_edge({"m": [(n, edge) for n in names]}, p)
                assert (c - a).norm_max() == 0


def test_split_merge_fermi_sign_compare_u1(TAT):
    """test_split_and_merge.py:66-101: merge sign is (n_odd & 2)."""
    edge = [(-1, 1), (0, 1), (+1, 1)]
    a_u1 = TAT.BoseU1.D.Tensor(["1", "2", "3", "4", "5"], [edge] * 5).range_()
    a_f = TAT.FermiU1.D.Tensor(["1", "2", "3", "4", "5"], [edge] * 5).range_()
    for i in range(5):
        for j in range(i, 5):
            names = a_u1.names[i:j]
            b_u1 = storage(a_u1.merge_edge({"m": names}))
            b_f = storage(a_f.merge_edge({"m": names}, True))
            for s in itertools.product([-1, 0, 1], repeat=5):
                if sum(s) != 0:
                    continue
                item = a_u1[{str(x + 1): (s[x], 0) for x in range(5)}]
                assert item in b_u1
                count = sum(s[x] != 0 for x in range(i, j))
                assert ((-item) if (count & 2) else item) in b_f


# ---------------------------------------------------------------- contract
def test_contract_no_symmetry_examples(TAT):
    """test_contract.py:5-36 (exact integers)"""
    a = TAT.No.D.Tensor(["A", "B"], [2, 2]).range_()
    b = TAT.No.D.Tensor(["C", "D"], [2, 2]).range_()
    for pair, names, expect in [
        (("A", "C"), ["B", "D"], [4, 6, 6, 10]),
        (("A", "D"), ["B", "C"], [2, 6, 3, 11]),
        (("B", "C"), ["A", "D"], [2, 3, 6, 11]),
        (("B", "D"), ["A", "C"], [1, 3, 3, 13]),
    ]:
        c = a.contract(b, {pair})
        assert c.names == names
        assert list(storage(c)) == expect
    x = TAT.No.D.Tensor(["A", "B", "C", "D"], [1, 2, 3, 4]).range_()
    y = TAT.No.D.Tensor(["E", "F", "G", "H"], [3, 1, 2, 4]).range_()
    r = x.contract(y, {("B", "G"), ("D", "H")})
    assert r.names == ["A", "C", "E", "F"]
    assert list(storage(r)) == [316, 796, 1276, 428, 1164, 1900, 540, 1532, 2524]


def test_contract_u1_equals_dense(TAT):
    """test_contract.py:39-69: symmetric contraction == dense contraction."""
    e1 = [(-1, 2), (0, 2), (+1, 2)]
    e2 = [(+1, 2), (0, 2), (-1, 2)]
    a = TAT.BoseU1.D.Tensor(["a", "b", "c", "d"], [e1, e2, e1, e2]).range_()
    b = TAT.BoseU1.D.Tensor(["e", "f", "g", "h"], [e1, e2, e1, e2]).range_()
    da, db = to_dense(a), to_dense(b)
    for plan in [{("a", "f"), ("b", "e")}, {("a", "h"), ("c", "f")}, {("b", "g"), ("d", "e")},
                 {("c", "h"), ("d", "g")}, {("a", "f"), ("d", "g")}]:
        c = a.contract(b, plan)
        ia = [("abcd".index(x)) for x, _ in sorted(plan)]
        ib = [("efgh".index(y)) for _, y in sorted(plan)]
        expect = np.tensordot(da, db, axes=(ia, ib))
        assert np.array_equal(to_dense(c), expect)


def test_contract_fermi_operand_order(TAT):
    """test_contract.py:72-113: swapping operands gives the same tensor, and
    fermi storage equals U1 storage for this arrow pattern."""
    fe1 = [(-1, 2), (0, 2), (+1, 2)]
    fe2 = [(+1, 2), (0, 2), (-1, 2)]
    fa = TAT.FermiU1.D.Tensor(["a", "b", "c", "d"], [(fe1, True), fe2, (fe1, True), (fe2, True)]).range_()
    fb = TAT.FermiU1.D.Tensor(["e", "f", "g", "h"], [fe1, fe2, (fe1, True), fe2]).range_()
    fc = fa.contract(fb, {("d", "e"), ("c", "f")})
    fd = fb.contract(fa, {("e", "d"), ("f", "c")})
    assert (fc - fd).norm_max() == 0
    ua = TAT.BoseU1.D.Tensor(["a", "b", "c", "d"], [fe1, fe2, fe1, fe2]).range_()
    ub = TAT.BoseU1.D.Tensor(["e", "f", "g", "h"], [fe1, fe2, fe1, fe2]).range_()
    uc = ua.contract(ub, {("d", "e"), ("c", "f")})
    assert np.array_equal(storage(fc), storage(uc))


def test_contract_with_split_and_merge(TAT):
    """test_contract.py:116-128"""
    edge1 = ([(-1, 2), (0, 2), (+1, 2)], False)
    edge2 = ([(+1, 2), (0, 2), (-1, 2)], True)
    a = TAT.FermiU1.D.Tensor(["a", "b", "c", "d"], [edge1, edge2, edge1, edge2]).range_()
    b = TAT.FermiU1.D.Tensor(["e", "f", "g", "h"], [edge1, edge2, edge1, edge2]).range_()
    c = a.contract(b, {("a", "f"), ("b", "g"), ("c", "h")})
    am = a.merge_edge({"m": ["b", "a"]}, False)
    bm = b.merge_edge({"m": ["g", "f"]}, True)
    cm = am.contract(bm, {("m", "m"), ("c", "h")})
    assert (c - cm).norm_max() == 0


def test_contract_with_reverse(TAT):
    """test_contract.py:131-161"""
    a = TAT.FermiZ2.D.Tensor(["i", "j"], [([(False, 2), (True, 2)], False), ([(False, 2), (True, 2)], True)]).range_()
    b = TAT.FermiZ2.D.Tensor(["i", "j"], [([(False, 2), (True, 2)], False),
                                          ([(False, 2), (True, 2)], True)]).range_().transpose(["j", "i"])
    c = a.contract(b, {("j", "i")})
    cr = a.reverse_edge({"j"}, False).contract(b.reverse_edge({"i"}, True), {("j", "i")})
    assert (c - cr).norm_max() == 0
    edge1 = ([(-1, 2), (0, 2), (+1, 2)], False)
    edge2 = ([(+1, 2), (0, 2), (-1, 2)], True)
    a = TAT.FermiU1.D.Tensor(["a", "b", "c", "d"], [edge1, edge2, edge1, edge2]).range_()
    b = TAT.FermiU1.D.Tensor(["e", "f", "g", "h"], [edge1, edge2, edge1, edge2]).range_()
    c = a.contract(b, {("a", "f"), ("b", "g"), ("c", "h")})
    cr = a.reverse_edge({"b", "a"}, False).contract(b.reverse_edge({"g", "f"}, True),
                                                      {("a", "f"), ("b", "g"), ("c", "h")})
    assert (c - cr).norm_max() == 0


def test_contract_corner_cases(TAT):
    """test_contract.py:179-214: k=0 blocks and unmatched sectors."""
    a = TAT.No.D.Tensor(["A", "B"], [2, 0]).range_()
    b = TAT.No.D.Tensor(["C", "D"], [0, 2]).range_()
    c = a.contract(b, {("B", "C")})
    assert storage(c).size == 4 and c.norm_max() == 0
    a = TAT.BoseZ2.D.Tensor(["A", "B"], [[(True, 2)], [(False, 0)]]).range_()
    b = TAT.BoseZ2.D.Tensor(["C", "D"], [[(False, 0)], [(False, 2)]]).range_()
    assert storage(a.contract(b, {("B", "C")})).size == 0
    a = TAT.BoseZ2.D.Tensor(["A", "B"], [[(False, 2)], [(True, 0)]]).range_()
    b = TAT.BoseZ2.D.Tensor(["C", "D"], [[(True, 0)], [(False, 2)]]).range_()
    c = a.contract(b, {("B", "C")})
    assert storage(c).size == 4 and c.norm_max() == 0


# ---------------------------------------------------------------- svd / qr
SVD_EDGES = [
    ([(-1, 1), (0, 1), (-2, 1)], True),
    ([(0, 1), (1, 2)], False),
    ([(0, 2), (1, 2)], False),
    ([(-2, 2), (-1, 1), (0, 2)], True),
]


@pytest.mark.parametrize("sym,cut,free", [("No", -1, {"C", "A"}), ("No", 2, {"C", "A"}), ("BoseU1", -1, {"C", "A"}),
                                          ("BoseU1", 7, {"C", "A"}), ("FermiU1", -1, {"C", "A"}),
                                          ("FermiU1", 8, {"B", "D"})])
def test_svd_reconstruct_and_unitary(TAT, sym, cut, free):
    """test_svd.py:19-97"""
    mod = getattr(TAT, sym)
    if sym == "No":
        a = mod.D.Tensor(["A", "B", "C", "D"], [2, 3, 4, 5] if cut < 0 else [5, 4, 3, 2]).range_()
    else:
        a = mod.D.Tensor(["A", "B", "C", "D"], SVD_EDGES).range_()
    u, s, v = a.svd(free, "E", "F", "U", "V", cut)
    check_unitary(u, "E", "E'", sym.startswith("Fermi"))
    check_unitary(v, "F", "F'", sym.startswith("Fermi"))
    b = v.contract(s, {("F", "V")}).contract(u, {("U", "E")})
    assert (a - b).norm_max() < 1e-6


def test_svd_cut_too_small(TAT):
    """test_svd.py:100-129"""
    a = TAT.No.D.Tensor(["A", "B"], [2, 2]).zero_()
    a[{"A": 0, "B": 0}] = 1
    u, s, v = a.svd({"B"}, "E", "F", "U", "V", 8)
    b = v.contract(s, {("F", "V")}).contract(u, {("U", "E")})
    assert (a - b).norm_max() < 1e-6
    assert storage(s).size == 1
    a = TAT.FermiU1.D.Tensor(["A", "B"], [[(0, 1), (+1, 1)], [(-1, 1), (0, 1)]]).range_(0, 1)
    u, s, v = a.svd({"B"}, "E", "F", "U", "V", 8)
    b = v.contract(s, {("F", "V")}).contract(u, {("U", "E")})
    assert (a - b).norm_max() < 1e-6
    assert storage(s).size == 1


@pytest.mark.parametrize("sym", ["No", "BoseU1", "FermiU1"])
@pytest.mark.parametrize("direction,free", [("Q", {"A", "C"}), ("R", {"B", "D"}), ("q", {"D"}), ("r", {"A"})])
def test_qr_reconstruct_and_unitary(TAT, sym, direction, free):
    """test_qr.py: Q R == A and Q is an isometry."""
    mod = getattr(TAT, sym)
    if sym == "No":
        a = mod.D.Tensor(["A", "B", "C", "D"], [2, 3, 4, 5]).range_()
    else:
        a = mod.D.Tensor(["A", "B", "C", "D"], SVD_EDGES).range_()
    q, r = a.qr(direction, free, "QR", "RQ")
    check_unitary(q, "QR", "QR'", sym.startswith("Fermi"))
    b = q.contract(r, {("QR", "RQ")})
    assert (a - b).norm_max() < 1e-6


def test_conjugate_twice_is_identity(TAT):
    """test_conjugate.py: conjugating twice returns the tensor."""
    a = TAT.FermiU1.D.Tensor(["A", "B", "C", "D"], SVD_EDGES).range_()
    for tm in (False, True):
        b = a.conjugate(tm).conjugate(tm)
        assert (a - b).norm_max() == 0
