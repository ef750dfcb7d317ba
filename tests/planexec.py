"""Test-side numpy executor for exported plan task lists.

Runs the task records of a ``tnsp_amd.plan.Plan`` on host arrays so the CPU
test-suite can check the native planners (block matching, offsets, signs)
against the oracle without a GPU.  Decompositions use numpy's LAPACK; the
gauge they pick may differ from the GPU kernels, so tests compare gauge
invariant quantities for qr/svd.
"""
import numpy as np
from numpy.lib.stride_tricks import as_strided


def _view(buf, off, dims, strides):
    it = buf.itemsize
    return as_strided(buf[off:], shape=tuple(int(d) for d in dims), strides=tuple(int(s) * it for s in strides))


def run_perm(records, bufs):
    for r in records:
        src_buf, dst_buf, src, dst, rank, neg, n = (int(x) for x in r[:7])
        dims, ss, ds = r[7:7 + rank], r[15:15 + rank], r[23:23 + rank]
        sv = _view(bufs[src_buf], src, dims, ss)
        dv = _view(bufs[dst_buf], dst, dims, ds)
        dv[...] = -sv if neg else sv


def run_fill(records, bufs):
    for buf, off, n in records:
        bufs[int(buf)][int(off):int(off) + int(n)] = 0


def run_gemm(records, bufs):
    for r in records:
        a_buf, b_buf, c_buf, a, b, c, m, n, k, lda, ldb, ldc, ta, tb, alpha = (int(x) for x in r)
        A = _view(bufs[a_buf], a, (k, m), (lda, 1)).T if ta else _view(bufs[a_buf], a, (m, k), (lda, 1))
        B = _view(bufs[b_buf], b, (n, k), (ldb, 1)).T if tb else _view(bufs[b_buf], b, (k, n), (ldb, 1))
        C = _view(bufs[c_buf], c, (m, n), (ldc, 1))
        C[...] = alpha * (A @ B)


def run_qr(records, bufs):
    for in_buf, f1_buf, f2_buf, i, f1, f2, m, n, lq in (tuple(int(x) for x in r) for r in records):
        M = _view(bufs[in_buf], i, (m, n), (n, 1))
        k = min(m, n)
        if lq:
            q, r = np.linalg.qr(M.T)
            F1, F2 = r.T, q.T
        else:
            F1, F2 = np.linalg.qr(M)
        _view(bufs[f1_buf], f1, (m, k), (k, 1))[...] = F1
        _view(bufs[f2_buf], f2, (k, n), (n, 1))[...] = F2


def run_svd(records, bufs):
    for in_buf, u_buf, s_buf, v_buf, i, u, s, v, m, n in (tuple(int(x) for x in r) for r in records):
        M = _view(bufs[in_buf], i, (m, n), (n, 1))
        k = min(m, n)
        U, S, VT = np.linalg.svd(M, full_matrices=False)
        _view(bufs[u_buf], u, (m, k), (k, 1))[...] = U
        bufs[s_buf][s:s + k] = S
        _view(bufs[v_buf], v, (k, n), (n, 1))[...] = VT


def _alloc(sizes, inputs, dtype):
    bufs = []
    it = iter(inputs)
    for sz in sizes:
        bufs.append(np.array(next(it), dtype=dtype) if sz < 0 else np.zeros(max(sz, 0), dtype=dtype))
    return bufs


def execute(plan, inputs, dtype=np.float64, cut=(-1, 0.0)):
    """Run a plan on host arrays; returns the list of output storages."""
    if plan.type in ("edge_operator", "conjugate"):
        bufs = _alloc(plan.buffers(), inputs, dtype)
        run_perm(plan.stage(0), bufs)
        return [bufs[1]]
    if plan.type == "contract":
        bufs = _alloc(plan.buffers(), inputs, dtype)
        run_perm(plan.stage(0), bufs)
        run_perm(plan.stage(1), bufs)
        run_fill(plan.stage(3), bufs)
        run_gemm(plan.stage(2), bufs)
        run_perm(plan.stage(4), bufs)
        return [bufs[5]]
    if plan.type == "qr":
        bufs = _alloc(plan.buffers(), inputs, dtype)
        run_perm(plan.stage(0), bufs)
        run_qr(plan.stage(1), bufs)
        run_perm(plan.stage(2), bufs)
        run_perm(plan.stage(3), bufs)
        return [bufs[4], bufs[5]]
    if plan.type == "svd":
        sizes = plan.buffers()
        bufs = _alloc(sizes[:4], inputs, dtype)
        bufs.append(np.zeros(max(sizes[4], 0), dtype=np.float64))
        run_perm(plan.stage(0), bufs)
        run_svd(plan.stage(1), bufs)
        plan.svd_truncate_host(bufs[4], cut[0], cut[1])
        sizes = plan.buffers()
        for sz in sizes[5:]:
            bufs.append(np.zeros(sz, dtype=dtype))
        run_fill(plan.stage(4), bufs)
        src = bufs[4]
        for r in plan.stage(5):  # diag copies read the float64 singular values
            _, dst_buf, s0, d0, rank, neg, n = (int(x) for x in r[:7])
            step = int(r[23])
            vals = src[s0:s0 + n].astype(dtype)
            bufs[dst_buf][d0:d0 + n * step:step] = -vals if neg else vals
        run_perm(plan.stage(2), bufs)
        run_perm(plan.stage(3), bufs)
        return [bufs[5], bufs[6], bufs[7]]
    raise ValueError(plan.type)
