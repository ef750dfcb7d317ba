"""Many tensors through one op.

A :class:`TensorBatch` holds n tensors that share names, symmetry model and
dtype but differ in block structure and data (the same tensor of n sampled
configurations).  ``contract`` / ``qr`` / ``svd`` plan all n on host threads
and run them as one launch per kernel kind (``tns_plan_*_many``,
``tns_plans_run``), which is how the environment-based sweep of
``tnsp_amd.environment`` runs many chains at once.  Semantics per tensor are
those of the drop-in ``Tensor`` methods (PyTAT ``contract`` / ``qr`` /
``svd``, PyTAT/PyTAT.hpp:988-1074).

A batch is flat arrays, not per-item Python objects: native shape handles,
device pointers and element counts, plus the torch buffers the pointers live
in.  Outputs of one op share one buffer; ``tensor(i)`` views one item as a
drop-in ``Tensor``.
"""
import ctypes
import os

import numpy as np

from . import _lib
from .plan import DTYPE_CODE, cut_params
from .shape import NAMES, Shape
from .tensor import Tensor, _stream, _tdtype, _torch, device

_ES = {"float64": 8, "float32": 4}


def _threads():
    return int(os.environ.get("TNS_THREADS", min(16, os.cpu_count() or 1)))


def _vpp(arr):
    return arr.ctypes.data_as(_lib.c_vpp)


def _i64p(arr):
    return arr.ctypes.data_as(_lib.c_i64p)


class _Owned:
    """Native handles (shapes or plans) released together."""
    __slots__ = ("arr", "free")

    def __init__(self, arr, free):
        self.arr = arr
        self.free = free

    def __del__(self):
        if _lib._lib is not None and len(self.arr):
            getattr(_lib._lib, self.free)(len(self.arr), _vpp(self.arr))


class TensorBatch:
    """n tensors of one name list.  ``handles`` (uint64: native shape
    handles), ``ptrs`` (uint64: device addresses), ``sizes`` (int64:
    elements), ``bufs`` / ``buf_idx`` / ``offs``: the torch buffer and
    element offset of each item; ``keep``: owners of the handles."""
    __slots__ = ("model", "dtype", "names", "handles", "ptrs", "sizes", "bufs", "buf_idx", "offs", "keep")

    def __init__(self, model, dtype, names, handles, sizes, bufs, buf_idx, offs, keep):
        self.model = model
        self.dtype = dtype
        self.names = list(names)
        self.handles = np.ascontiguousarray(handles, dtype=np.uint64)
        self.sizes = np.ascontiguousarray(sizes, dtype=np.int64)
        self.bufs = bufs
        self.buf_idx = np.ascontiguousarray(buf_idx, dtype=np.int64)
        self.offs = np.ascontiguousarray(offs, dtype=np.int64)
        self.keep = keep
        base = np.array([b.data_ptr() for b in bufs], dtype=np.uint64)
        self.ptrs = base[self.buf_idx] + self.offs.astype(np.uint64) * np.uint64(_ES[dtype])

    # -- construction ---------------------------------------------------------------
    @classmethod
    def from_tensors(cls, tensors):
        t0 = tensors[0]
        for t in tensors:
            if t.names != t0.names or t.model is not t0.model or t.dtype != t0.dtype:
                raise ValueError("a tensor batch needs one name list, model and dtype")
        n = len(tensors)
        return cls(t0.model, t0.dtype, t0.names, [t._shape.handle.value for t in tensors],
                   [t._shape.size for t in tensors], [t._data for t in tensors], np.arange(n), np.zeros(n),
                   tuple(tensors))

    @classmethod
    def select(cls, choices, which):
        """Item i is ``choices[which[i]]`` (tensors of one name list): the
        same site tensor at the occupation each configuration holds."""
        t0 = choices[0]
        which = np.asarray(which, dtype=np.int64)
        h = np.array([t._shape.handle.value for t in choices], dtype=np.uint64)
        sz = np.array([t._shape.size for t in choices], dtype=np.int64)
        return cls(t0.model, t0.dtype, t0.names, h[which], sz[which], [t._data for t in choices], which,
                   np.zeros(len(which)), tuple(choices))

    @classmethod
    def repeat(cls, tensor, n):
        return cls.select([tensor], np.zeros(n, dtype=np.int64))

    @classmethod
    def concat(cls, batches):
        """The items of several batches of one name list, in order."""
        b0 = batches[0]
        bufs, idx, keep = [], [], []
        for b in batches:
            if b.names != b0.names:
                raise ValueError("concat needs one name list")
            idx.append(b.buf_idx + len(bufs))
            bufs.extend(b.bufs)
            keep.append(b.keep)
        return cls(b0.model, b0.dtype, b0.names, np.concatenate([b.handles for b in batches]),
                   np.concatenate([b.sizes for b in batches]), bufs, np.concatenate(idx),
                   np.concatenate([b.offs for b in batches]), tuple(keep))

    def __len__(self):
        return len(self.handles)

    def tensor(self, i):
        """Item i as a drop-in ``Tensor`` (shape copied, storage shared)."""
        h = ctypes.c_void_p()
        _lib.check(_lib.lib().tns_shape_rename(ctypes.c_void_p(int(self.handles[i])), 0, _lib.arr32([]),
                                               _lib.arr32([]), ctypes.byref(h)))
        o, n = int(self.offs[i]), int(self.sizes[i])
        return Tensor._wrap(self.model, self.dtype, Shape(h, self.model.kind), self.bufs[self.buf_idx[i]][o:o + n])

    def take(self, idx):
        idx = np.asarray(idx, dtype=np.int64)
        return TensorBatch(self.model, self.dtype, self.names, self.handles[idx], self.sizes[idx], self.bufs,
                           self.buf_idx[idx], self.offs[idx], self.keep)

    def rename(self, mapping):
        """safe_rename (tetragono/tetragono/utility.py:353-358): names absent
        from the batch are ignored; storage is shared."""
        m = {k: v for k, v in mapping.items() if k in self.names}
        if not m:
            return self
        old = list(m)
        out = np.zeros(len(self), dtype=np.uint64)
        _lib.check(_lib.lib().tns_shape_rename_many(len(self), _vpp(self.handles), len(old),
                                                    _lib.arr32(NAMES.ids(old)),
                                                    _lib.arr32(NAMES.ids([m[k] for k in old])), _vpp(out)))
        return TensorBatch(self.model, self.dtype, [m.get(n, n) for n in self.names], out, self.sizes, self.bufs,
                           self.buf_idx, self.offs, (_Owned(out, "tns_shapes_free"), self.keep))

    def values(self):
        """First storage element of every item (0 for an empty tensor): the
        value of one-element tensors."""
        out = np.zeros(len(self))
        for b in np.unique(self.buf_idx):
            sel = np.nonzero((self.buf_idx == b) & (self.sizes > 0))[0]
            if len(sel):
                idx = _torch().as_tensor(self.offs[sel], device=self.bufs[b].device)
                out[sel] = self.bufs[b][idx].double().cpu().numpy()
        return out


# -- plans ------------------------------------------------------------------------------
def _plans(n, fn, *args):
    out = np.zeros(n, dtype=np.uint64)
    _lib.check(fn(n, *args, _threads(), _vpp(out)))
    return _Owned(out, "tns_plans_free")


def plan_contract(a_handles, b_handles, pairs):
    """One contract plan per item pair, planned on host threads."""
    a_handles = np.ascontiguousarray(a_handles, dtype=np.uint64)
    b_handles = np.ascontiguousarray(b_handles, dtype=np.uint64)
    if len(a_handles) != len(b_handles):
        raise ValueError("batches differ in length")
    flat = [i for x, y in pairs for i in NAMES.ids([x, y])]
    return _plans(len(a_handles), _lib.lib().tns_plan_contract_many, _vpp(a_handles), _vpp(b_handles), len(pairs),
                  _lib.arr32(flat))


def plan_qr(handles, direction, free_names, cq, cr):
    handles = np.ascontiguousarray(handles, dtype=np.uint64)
    return _plans(len(handles), _lib.lib().tns_plan_qr_many, _vpp(handles), direction.encode()[:1], len(free_names),
                  _lib.arr32(NAMES.ids(free_names)), NAMES.id(cq), NAMES.id(cr))


def plan_svd(handles, free_u, cu, cv, su, sv):
    handles = np.ascontiguousarray(handles, dtype=np.uint64)
    return _plans(len(handles), _lib.lib().tns_plan_svd_many, _vpp(handles), len(free_u), _lib.arr32(NAMES.ids(free_u)),
                  NAMES.id(cu), NAMES.id(cv), NAMES.id(su), NAMES.id(sv))


def _names_of(handle, kind):
    h = ctypes.c_void_p()
    _lib.check(_lib.lib().tns_shape_rename(ctypes.c_void_p(int(handle)), 0, _lib.arr32([]), _lib.arr32([]),
                                           ctypes.byref(h)))
    return Shape(h, kind).names


def _outputs(plans, which, model, dtype):
    """Output batches ``which`` of a plan list: one device buffer per output,
    items packed back to back."""
    torch = _torch()
    n = len(plans.arr)
    L = _lib.lib()
    res = []
    for w in which:
        hs = np.zeros(n, dtype=np.uint64)
        sizes = np.zeros(n, dtype=np.int64)
        _lib.check(L.tns_plans_output_shapes(n, _vpp(plans.arr), w, _vpp(hs), _i64p(sizes)))
        owner = _Owned(hs, "tns_shapes_free")
        offs = np.zeros(n, dtype=np.int64)
        np.cumsum(sizes[:-1], out=offs[1:])
        buf = torch.empty(max(1, int(sizes.sum())), dtype=_tdtype(dtype), device=device())
        res.append(TensorBatch(model, dtype, _names_of(hs[0], model.kind), hs, sizes, [buf], np.zeros(n), offs,
                               owner))
    return res


def _workspaces(plans, dtype):
    n = len(plans.arr)
    sizes = np.zeros(n, dtype=np.int64)
    _lib.check(_lib.lib().tns_plans_workspace(n, _vpp(plans.arr), DTYPE_CODE[dtype], _i64p(sizes)))
    padded = (sizes + 255) // 256 * 256
    offs = np.zeros(n, dtype=np.uint64)
    np.cumsum(padded[:-1], out=offs[1:])
    buf = _torch().empty(max(256, int(padded.sum())), dtype=_torch().uint8, device=device())
    return buf, np.ascontiguousarray(offs + np.uint64(buf.data_ptr()))


def _run(plans, dtype, inputs, outputs, workspaces):
    _lib.check(_lib.lib().tns_plans_run(len(plans.arr), _vpp(plans.arr), DTYPE_CODE[dtype],
                                        _vpp(np.ascontiguousarray(inputs, dtype=np.uint64)),
                                        _vpp(np.ascontiguousarray(outputs, dtype=np.uint64)), _vpp(workspaces),
                                        ctypes.c_void_p(_stream())))


def _interleave(*cols):
    return np.stack(cols, axis=1).ravel()


# -- ops ----------------------------------------------------------------------------------
def contract(a, b, pairs):
    """Item-wise ``a[i].contract(b[i], pairs)``; pairs whose names are absent
    on either side are dropped first (safe_contract,
    tetragono/tetragono/utility.py:340-350)."""
    if len(a) != len(b):
        raise ValueError("batches differ in length")
    pr = [(x, y) for x, y in pairs if x in a.names and y in b.names]
    plans = plan_contract(a.handles, b.handles, pr)
    (res,) = _outputs(plans, [0], a.model, a.dtype)
    ws, wps = _workspaces(plans, a.dtype)
    _run(plans, a.dtype, _interleave(a.ptrs, b.ptrs), res.ptrs, wps)
    return res


def qr(t, direction, free_names, cq, cr):
    plans = plan_qr(t.handles, direction, [x for x in free_names if x in t.names], cq, cr)
    q, r = _outputs(plans, [0, 1], t.model, t.dtype)
    ws, wps = _workspaces(plans, t.dtype)
    _run(plans, t.dtype, t.ptrs, _interleave(q.ptrs, r.ptrs), wps)
    return q, r


def conjugate(t, trivial_metric=False):
    """Item-wise ``conjugate`` (TAT/include/TAT/implement/conjugate.hpp:32-120)."""
    if t.model.kind.length == 0:
        return t
    n = len(t)
    plans = _plans(n, _lib.lib().tns_plan_conjugate_many, _vpp(np.ascontiguousarray(t.handles)),
                   int(bool(trivial_metric)))
    (res,) = _outputs(plans, [0], t.model, t.dtype)
    ws, wps = _workspaces(plans, t.dtype)
    _run(plans, t.dtype, t.ptrs, res.ptrs, wps)
    return res


def scale(t, factors):
    """Item i times ``factors[i]`` (a new batch; the scalar product of
    PyTAT.hpp:707-718 per item)."""
    torch = _torch()
    n = len(t)
    offs = np.zeros(n, dtype=np.int64)
    np.cumsum(t.sizes[:-1], out=offs[1:])
    total = int(t.sizes.sum())
    buf = torch.empty(max(1, total), dtype=_tdtype(t.dtype), device=device())
    if total:
        f = np.repeat(np.asarray(factors, dtype=np.float64), t.sizes)
        src = torch.cat([t.bufs[b][o:o + s] for b, o, s in zip(t.buf_idx, t.offs, t.sizes) if s > 0])
        buf[:total] = src * torch.as_tensor(f, dtype=src.dtype, device=src.device)
    return TensorBatch(t.model, t.dtype, t.names, t.handles, t.sizes, [buf], np.zeros(n), offs, t.keep)


def svd(t, free_u, cu, cv, su, sv, cut=-1):
    plans = plan_svd(t.handles, [x for x in free_u if x in t.names], cu, cv, su, sv)
    ws, wps = _workspaces(plans, t.dtype)
    remain, relative = cut_params(cut)
    n = len(t)
    _lib.check(_lib.lib().tns_svd_factor_many(n, _vpp(plans.arr), DTYPE_CODE[t.dtype],
                                              _vpp(np.ascontiguousarray(t.ptrs)), _vpp(wps), remain, relative,
                                              _threads(), ctypes.c_void_p(_stream())))
    u, s, v = _outputs(plans, [0, 1, 2], t.model, t.dtype)
    _run(plans, t.dtype, t.ptrs, _interleave(u.ptrs, s.ptrs, v.ptrs), wps)
    return u, s, v
