"""tns_plan_run_noalloc: the allocation-free run of a plan (workspace and task
records in the library's static device array) against tns_plan_run with a
caller workspace, through the C ABI.  Same kernels, so results are bit-exact.
Two calls in a row, the second on smaller tensors; every plan kind it accepts
(edge operator, conjugate, contract, QR with sectors beyond LDS)."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _tensor(kind, names, edges, seed):
    from tnsp_amd import TAT

    t = getattr(TAT, kind).D.Tensor(names, edges)
    t.storage = np.random.default_rng(seed).standard_normal(t._shape.size)
    return t


def _both(plan, inputs, n_out):
    """Outputs of plan.run (caller workspace) and plan.run_noalloc."""
    from tnsp_amd.tensor import _stream

    outs = []
    for mode in ("workspace", "noalloc"):
        bufs = [torch.full((plan.output(i).size,), float("nan"), dtype=torch.float64, device="cuda")
                for i in range(n_out)]
        ptrs = [b.data_ptr() for b in bufs]
        if mode == "workspace":
            ws = torch.empty(max(1, plan.workspace("float64")), dtype=torch.uint8, device="cuda")
            plan.run("float64", [x.data_ptr() for x in inputs], ptrs, ws.data_ptr(), _stream())
        else:
            plan.run_noalloc("float64", [x.data_ptr() for x in inputs], ptrs, _stream())
        torch.cuda.synchronize()
        outs.append([b.cpu().numpy() for b in bufs])
    return outs


def _check(plan, inputs, n_out):
    ref, got = _both(plan, inputs, n_out)
    for r, g in zip(ref, got):
        assert not np.isnan(r).any()
        assert np.array_equal(r, g)


@pytest.mark.parametrize("dims", [(6, 3, 3), (2, 1, 1)])
def test_noalloc_matches_workspace_run(dims):
    """First call on larger tensors, then the same ops on smaller ones: each
    noalloc result equals the workspace run bit for bit."""
    from tnsp_amd import plan as P

    a, b, c = dims
    seg = [(0, a), (1, b), (2, c)]
    cseg = [(0, a), (-1, b), (-2, c)]
    A = _tensor("FermiU1", ["i", "j", "k", "l"], [(seg, False), (seg, True), (seg, False), (seg, True)], 1)
    B = _tensor("FermiU1", ["k", "l", "m", "n"], [(cseg, True), (cseg, False), (seg, True), (seg, False)], 2)
    _check(P.edge_operator(A._shape, new_names=["l", "j", "i", "k"]), [A._data], 1)
    _check(P.conjugate(A._shape), [A._data], 1)
    _check(P.contract(A._shape, B._shape, [("k", "k"), ("l", "l")]), [A._data, B._data], 1)
    _check(P.qr(A._shape, "r", ["k", "l"], "Q", "R"), [A._data], 2)


def test_noalloc_qr_beyond_lds():
    """A tall sector (global-scratch QR) through the static array, then a
    small one: the second call must not see the first call's scratch."""
    from tnsp_amd import plan as P

    for rows in (48, 5):
        T = _tensor("No", ["a", "b", "c"], [rows, 40, 30], rows)
        _check(P.qr(T._shape, "r", ["c"], "Q", "R"), [T._data], 2)
