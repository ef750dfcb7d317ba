"""GPU parity of the block-sparse operations against the CPU oracle.

Same seeded inputs on both sides.  Bar: block structure, edge segments,
arrows and fermion signs bit-exact; permutes (transpose / merge / split /
reverse / conjugate) bit-exact; GEMM / QR / SVD values within 1e-10 relative
in float64 (1e-4 in float32); QR / SVD compared on gauge-invariant products.
"""
import itertools

import numpy as np
import pytest

from helpers import storage

pytestmark = pytest.mark.gpu

E3 = [(-1, 2), (0, 4), (1, 2)]


def pair_c1(kind, seed=1):
    """Config-1 tensors: A[i,j,k,l], B[k,l,m,n], every edge dimension 8 as
    charges {-1:2, 0:4, +1:2}; B's k, l are the conjugates of A's."""
    from oracle import tat

    K = tat.KINDS[kind]
    rng = np.random.default_rng(seed)
    if kind.startswith("Fermi") and kind.endswith("Z2"):
        e = [(0, 4), (1, 4)]
    else:
        e = E3
    A = tat.Tensor(K, ["i", "j", "k", "l"], [(e, False), (e, True), (e, False), (e, True)])
    A.data[:] = rng.standard_normal(A.size)
    B = tat.Tensor(K, ["k", "l", "m", "n"], [A.edges[2].conjugate(), A.edges[3].conjugate(), (e, True), (e, False)])
    B.data[:] = rng.standard_normal(B.size)
    return A, B


def to_gpu(t, dtype="float64"):
    from tnsp_amd import TAT

    mod = getattr(TAT, t.kind.name)
    g = (mod.D if dtype == "float64" else mod.S).Tensor(t.names, [(list(e.segments), e.arrow) for e in t.edges])
    g.storage = t.data
    return g


def same_structure(g, o):
    assert g.names == o.names
    assert [e.segments for e in g.edges] == [e.segments for e in o.edges]
    assert [e.arrow for e in g.edges] == [e.arrow for e in o.edges]


def rel(a, b):
    b = np.asarray(b)
    return np.max(np.abs(np.asarray(a) - b)) / max(np.max(np.abs(b)), 1e-300) if b.size else 0.0


KINDS = ["FermiU1", "FermiZ2", "BoseU1"]


@pytest.mark.parametrize("kind", KINDS)
def test_transpose_all_permutations_bit_exact(kind):
    A, _ = pair_c1(kind)
    gA = to_gpu(A)
    for perm in itertools.permutations("ijkl"):
        o = A.transpose(list(perm))
        g = gA.transpose(list(perm))
        same_structure(g, o)
        assert np.array_equal(storage(g), o.data), perm


@pytest.mark.parametrize("kind", KINDS)
def test_merge_split_reverse_bit_exact(kind):
    A, _ = pair_c1(kind)
    gA = to_gpu(A)
    for names in (["j", "i"], ["k", "l", "i"], ["l"]):
        for p in (False, True):
            o = A.merge_edge({"m": names}, p)
            g = gA.merge_edge({"m": names}, p)
            same_structure(g, o)
            assert np.array_equal(storage(g), o.data)
            seg = {n: A.edge_by_name(n) for n in names}
            os_ = o.split_edge({"m": [(n, list(seg[n].segments)) for n in names]}, p)
            gs = g.split_edge({"m": [(n, list(seg[n].segments)) for n in names]}, p)
            same_structure(gs, os_)
            assert np.array_equal(storage(gs), os_.data)
    for rev in ({"i"}, {"j", "k"}):
        for p in (False, True):
            o = A.reverse_edge(rev, p)
            g = gA.reverse_edge(rev, p)
            same_structure(g, o)
            assert np.array_equal(storage(g), o.data)
    for tm in (False, True):
        o = A.conjugate(tm)
        g = gA.conjugate(tm)
        same_structure(g, o)
        assert np.array_equal(storage(g), o.data)


@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("dtype", ["float64", "float32"])
def test_contract_matches_oracle(kind, dtype):
    A, B = pair_c1(kind)
    tol = 1e-12 if dtype == "float64" else 1e-5
    gA, gB = to_gpu(A, dtype), to_gpu(B, dtype)
    o = A.contract(B, {("k", "k"), ("l", "l")})
    g = gA.contract(gB, {("k", "k"), ("l", "l")})
    same_structure(g, o)
    assert rel(storage(g), o.data) < tol
    o2 = B.contract(A, {("k", "k"), ("l", "l")})
    g2 = gB.contract(gA, {("k", "k"), ("l", "l")})
    same_structure(g2, o2)
    assert rel(storage(g2), o2.data) < tol
    # three common edges, conjugate partner with one edge renamed
    A2 = A.conjugate().edge_rename({"l": "l2"})
    gA2 = gA.conjugate().edge_rename({"l": "l2"})
    o3 = A.contract(A2, {("i", "i"), ("j", "j"), ("k", "k")})
    g3 = gA.contract(gA2, {("i", "i"), ("j", "j"), ("k", "k")})
    same_structure(g3, o3)
    assert rel(storage(g3), o3.data) < tol


@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("free,cut", [({"i", "j"}, 16), ({"m", "n"}, 16), ({"i", "m"}, -1), ({"j"}, 5)])
def test_svd_matches_oracle(kind, free, cut):
    A, B = pair_c1(kind)
    C = A.contract(B, {("k", "k"), ("l", "l")})
    gC = to_gpu(C)
    u, s, v = C.svd(free, "cu", "cv", "su", "sv", cut)
    gu, gs, gv = gC.svd(free, "cu", "cv", "su", "sv", cut)
    for g, o in ((gu, u), (gs, s), (gv, v)):
        same_structure(g, o)
    assert rel(storage(gs), s.data) < 1e-10
    r0 = u.contract(s, {("cu", "su")}).contract(v, {("sv", "cv")})
    r1 = gu.contract(gs, {("cu", "su")}).contract(gv, {("sv", "cv")})
    same_structure(r1, r0)
    assert rel(storage(r1), r0.data) < 1e-10


@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("direction,free", [("r", {"m", "n"}), ("q", {"m", "n"}), ("r", {"i"}), ("q", {"i", "j"})])
def test_qr_matches_oracle(kind, direction, free):
    A, B = pair_c1(kind)
    C = A.contract(B, {("k", "k"), ("l", "l")})
    gC = to_gpu(C)
    q, r = C.qr(direction, free, "cq", "cr")
    gq, gr = gC.qr(direction, free, "cq", "cr")
    same_structure(gq, q)
    same_structure(gr, r)
    back = gq.contract(gr, {("cq", "cr")}).transpose(C.names)
    assert rel(storage(back), C.data) < 1e-10


@pytest.mark.parametrize("kind", ["FermiU1", "BoseU1"])
@pytest.mark.parametrize("direction", ["r", "q"])
def test_qr_large_blocks(kind, direction):
    """Sectors too large for LDS (tall 1100 x 150 and wide) take the blocked
    Householder path: Q R reconstructs the input and Q is orthonormal."""
    from oracle import tat

    K = tat.KINDS[kind]
    rng = np.random.default_rng(7)
    a = tat.Tensor(K, ["a", "b", "c"], [([(0, 700), (1, 400)], False), ([(0, 3), (-1, 2)], True),
                                        ([(0, 50), (1, 100), (-1, 30)], False)])
    a.data[:] = rng.standard_normal(a.size)
    g = to_gpu(a)
    q, r = a.qr(direction, {"c"}, "cq", "cr")
    gq, gr = g.qr(direction, {"c"}, "cq", "cr")
    same_structure(gq, q)
    same_structure(gr, r)
    back = gq.contract(gr, {("cq", "cr")}).transpose(a.names)
    assert rel(storage(back), a.data) < 1e-12
    # Q^dagger Q is the identity on the common edge (up to fermion signs)
    keep = [n for n in q.names if n != "cq"]
    prod = gq.conjugate().edge_rename({"cq": "cq'"}).contract(gq, {(n, n) for n in keep})
    ident = q.conjugate().edge_rename({"cq": "cq'"}).contract(q, {(n, n) for n in keep})
    same_structure(prod, ident)
    np.testing.assert_allclose(np.abs(storage(prod)), np.abs(ident.data), atol=1e-12)


@pytest.mark.parametrize("dtype", ["float64", "float32"])
def test_svd_register_sectors(dtype):
    """Sectors of 100-190 rows take the register-resident Jacobi kernel
    (per-op SVDs route like the engine's): singular values and U S V match
    the oracle, 1e-10 in float64 and 1e-4 in float32, truncated and not."""
    from oracle import tat

    K = tat.KINDS["BoseU1"]
    rng = np.random.default_rng(8)
    a = tat.Tensor(K, ["a", "b"], [([(0, 150), (1, 120), (-1, 190)], False), ([(0, 130), (-1, 100), (1, 170)], False)])
    a.data[:] = rng.standard_normal(a.size)
    g = to_gpu(a, dtype)
    tol = 1e-10 if dtype == "float64" else 1e-4
    for cut in (-1, 150):
        u, s, v = a.svd({"a"}, "cu", "cv", "su", "sv", cut)
        gu, gs, gv = g.svd({"a"}, "cu", "cv", "su", "sv", cut)
        for x, o in ((gu, u), (gs, s), (gv, v)):
            same_structure(x, o)
        assert rel(storage(gs), s.data) < tol
        r0 = u.contract(s, {("cu", "su")}).contract(v, {("sv", "cv")})
        r1 = gu.contract(gs, {("cu", "su")}).contract(gv, {("sv", "cv")})
        assert rel(storage(r1), r0.data) < tol


@pytest.mark.parametrize("dtype", ["float64", "float32"])
def test_contract_every_gemm_tile(dtype):
    """Sectors of 3 .. 130 rows / columns exercise the 16x16, 32x32 and 64x64
    GEMM tiles (size classes <= 16, <= 48, larger)."""
    from oracle import tat

    K = tat.KINDS["FermiU1"]
    rng = np.random.default_rng(11)
    seg = [(-1, 3), (0, 40), (1, 130)]
    A = tat.Tensor(K, ["i", "k"], [(seg, False), (seg, True)])
    A.data[:] = rng.standard_normal(A.size)
    B = tat.Tensor(K, ["k", "j"], [A.edges[1].conjugate(), (seg, True)])
    B.data[:] = rng.standard_normal(B.size)
    C = A.contract(B, {("k", "k")})
    gC = to_gpu(A, dtype).contract(to_gpu(B, dtype), {("k", "k")})
    same_structure(gC, C)
    assert rel(storage(gC), C.data) < (1e-12 if dtype == "float64" else 1e-5)


@pytest.mark.gpu
def test_svd_not_converged_raises():
    """With the Jacobi sweep limit forced down to one sweep, the random C1
    sectors do not converge: the kernels mark them and the SVD raises (as
    TAT does when gesvd reports info > 0); with the limit restored the same
    SVD succeeds.  Both the LDS kernel (small sectors) and the per-op path
    are exercised."""
    from oracle import tat
    from helpers import c1_pair, to_gpu
    from tnsp_amd import _lib

    A, B = c1_pair(tat, "FermiU1", 11)
    C = to_gpu(A).contract(to_gpu(B), {("k", "k"), ("l", "l")})
    L = _lib.lib()
    assert L.tns_set_svd_max_sweeps(1) == 0
    try:
        with pytest.raises(RuntimeError, match="converge"):
            C.svd({"i", "j"}, "cu", "cv", "su", "sv", 16)
    finally:
        assert L.tns_set_svd_max_sweeps(60) == 0
    u, s, v = C.svd({"i", "j"}, "cu", "cv", "su", "sv", 16)
    assert np.all(np.isfinite(np.asarray(s.storage)))


def test_transpose_large_dims_bit_exact():
    """Tasks far larger than one workgroup chunk, dimensions beyond 2^10 and
    odd sizes: every permutation equals numpy's transpose bit for bit."""
    from tnsp_amd import TAT

    dims = [301, 7, 1103]
    a = TAT.No.D.Tensor(["a", "b", "c"], dims)
    x = np.random.default_rng(5).standard_normal(int(np.prod(dims)))
    a.storage = x
    x = x.reshape(dims)
    for perm in itertools.permutations(range(3)):
        names = [["a", "b", "c"][p] for p in perm]
        assert np.array_equal(storage(a.transpose(names)), np.transpose(x, perm).reshape(-1)), perm


@pytest.mark.parametrize("dtype,tol", [("float64", 1e-12), ("float32", 2e-5)])
@pytest.mark.parametrize("rank_deficient", [False, True])
@pytest.mark.parametrize("cols", [177, 96, 20])
@pytest.mark.parametrize("order", ["ab", "ba"])
def test_qr_tall_beyond_lds(dtype, tol, rank_deficient, cols, order):
    """Tall blocks beyond LDS (1200 rows) take the blocked Householder
    kernel, stored either way round (order "ba" stores the block transposed,
    which TAT factors as an LQ of the wide matrix), with two equal columns
    (rank-deficient) or not: Q R = A, Q^T Q = 1 and R is upper triangular."""
    from tnsp_amd import TAT

    rng = np.random.default_rng(cols)
    x = rng.standard_normal((1200, cols))
    if rank_deficient:
        x[:, cols // 2] = x[:, 1]
    mod = TAT.No.D if dtype == "float64" else TAT.No.S
    if order == "ab":
        a = mod.Tensor(["a", "b"], [1200, cols])
        a.storage = x.reshape(-1)
    else:
        a = mod.Tensor(["b", "a"], [cols, 1200])
        a.storage = x.T.reshape(-1)
    q, r = a.qr("r", {"b"}, "Q", "R")
    qm = np.asarray(storage(q.transpose(["a", "Q"])), dtype=np.float64).reshape(1200, cols)
    rm = np.asarray(storage(r.transpose(["R", "b"])), dtype=np.float64).reshape(cols, cols)
    assert np.abs(qm @ rm - x).max() < tol * np.abs(x).max() * 10
    assert np.abs(qm.T @ qm - np.eye(cols)).max() < tol * 10
    assert np.array_equal(np.tril(rm, -1), np.zeros_like(rm))


@pytest.mark.parametrize("scale", [1.0, 1e-17, 1e-150])
@pytest.mark.parametrize("rows,cols,rank", [(4, 9, 3), (9, 4, 3), (40, 130, 35), (150, 170, 120), (300, 200, 150)])
def test_svd_structurally_rank_deficient(scale, rows, cols, rank):
    """Blocks whose columns span fewer dimensions than their count (zero
    columns of M, as the 8x8 PEPS produces at some configurations) and blocks
    at 1e-150: every Jacobi kernel (LDS, register, global) converges, the
    singular values equal LAPACK's to 1e-13 of the largest and U S V
    rebuilds the block.  The null-space columns used to keep rotating into
    the underflow range until the sweep limit."""
    from tnsp_amd import TAT

    rng = np.random.default_rng(rows + cols)
    x = np.zeros((rows, cols))
    x[:, :rank] = rng.standard_normal((rows, rank)) * scale
    a = TAT.No.D.Tensor(["a", "b"], [rows, cols])
    a.storage = x.reshape(-1)
    u, s, v = a.svd({"a"}, "U", "V", "SU", "SV")
    # exact zeros are dropped, as TAT's rule keeps only values above zero
    k = s.edges[0].dimension()
    assert rank <= k <= min(rows, cols)
    um = np.asarray(storage(u.transpose(["a", "U"]))).reshape(rows, k)
    sm = np.asarray(storage(s.transpose(["SU", "SV"]))).reshape(k, k)
    vm = np.asarray(storage(v.transpose(["V", "b"]))).reshape(k, cols)
    ref = np.linalg.svd(x, compute_uv=False)
    got = np.sort(np.abs(np.diag(sm)))[::-1]
    assert np.abs(got[:rank] - ref[:rank]).max() <= 1e-13 * ref[0]
    assert np.all(got[rank:] <= 1e-13 * ref[0])
    assert np.abs(um @ sm @ vm - x).max() <= 1e-12 * np.abs(x).max()


@pytest.mark.parametrize("rows,cols", [(192, 36), (36, 192), (120, 40), (100, 8), (64, 2), (90, 44)])
@pytest.mark.parametrize("graded", [False, True])
def test_svd_tall_lds_tier(rows, cols, graded):
    """Tall and wide blocks of the LDS tier (working set at most 76 KiB, the
    shapes that carry most of C4's LDS-tier Jacobi work), where the tracked
    column norms see the longest columns: singular values equal LAPACK's to
    1e-13 of the largest, U and V are orthonormal (V is accumulated here, not
    recovered) and U S V rebuilds the block, for random and graded
    (1 .. 1e-12) spectra."""
    from tnsp_amd import TAT

    rng = np.random.default_rng(rows * 1000 + cols)
    x = rng.standard_normal((rows, cols))
    if graded:
        u0, s0, v0 = np.linalg.svd(x, full_matrices=False)
        x = (u0 * (s0 * np.logspace(0, -12, s0.size))) @ v0
    a = TAT.No.D.Tensor(["a", "b"], [rows, cols])
    a.storage = x.reshape(-1)
    u, s, v = a.svd({"a"}, "U", "V", "SU", "SV")
    k = s.edges[0].dimension()
    assert k == min(rows, cols)
    um = np.asarray(storage(u.transpose(["a", "U"]))).reshape(rows, k)
    sm = np.asarray(storage(s.transpose(["SU", "SV"]))).reshape(k, k)
    vm = np.asarray(storage(v.transpose(["V", "b"]))).reshape(k, cols)
    ref = np.linalg.svd(x, compute_uv=False)
    got = np.sort(np.abs(np.diag(sm)))[::-1]
    assert np.abs(got - ref).max() <= 1e-13 * ref[0]
    assert np.abs(um.T @ um - np.eye(k)).max() <= 1e-12
    assert np.abs(vm @ vm.T - np.eye(k)).max() <= 1e-12
    assert np.abs(um @ sm @ vm - x).max() <= 1e-12 * np.abs(x).max()
