"""Many tensors through one op.

A :class:`TensorBatch` holds n tensors that share names, symmetry model and
dtype but differ in block structure and data (the same tensor of n sampled
configurations).  ``contract`` / ``qr`` / ``svd`` plan all n on host threads
and run them as one launch per kernel kind (``tns_plan_*_many``,
``tns_plans_run``), which is how the environment-based amplitude updates of
``tnsp_amd.auxiliaries`` run for many chains at once.  Semantics per tensor
are those of the drop-in ``Tensor`` methods (PyTAT ``contract`` / ``qr`` /
``svd``, PyTAT/PyTAT.hpp:988-1074).
"""
import ctypes
import os

import numpy as np

from . import _lib
from .plan import DTYPE_CODE, Plan, cut_params
from .shape import NAMES, Shape
from .tensor import Tensor, _stream, _tdtype, _torch, device


def _threads():
    return int(os.environ.get("TNS_THREADS", min(16, os.cpu_count() or 1)))


def _vp_array(values):
    return (ctypes.c_void_p * max(1, len(values)))(*[int(v) for v in values])


class TensorBatch:
    __slots__ = ("model", "dtype", "names", "shapes", "data")

    def __init__(self, model, dtype, names, shapes, data):
        self.model = model
        self.dtype = dtype
        self.names = list(names)
        self.shapes = shapes
        self.data = data

    @classmethod
    def from_tensors(cls, tensors):
        t0 = tensors[0]
        for t in tensors:
            if t.names != t0.names or t.model is not t0.model or t.dtype != t0.dtype:
                raise ValueError("a tensor batch needs one name list, model and dtype")
        return cls(t0.model, t0.dtype, t0.names, [t._shape for t in tensors], [t._data for t in tensors])

    @classmethod
    def repeat(cls, tensor, n):
        return cls(tensor.model, tensor.dtype, tensor.names, [tensor._shape] * n, [tensor._data] * n)

    def __len__(self):
        return len(self.shapes)

    def tensor(self, i):
        return Tensor._wrap(self.model, self.dtype, self.shapes[i], self.data[i])

    def take(self, idx):
        return TensorBatch(self.model, self.dtype, self.names, [self.shapes[i] for i in idx],
                           [self.data[i] for i in idx])

    def put(self, idx, other):
        """This batch with items ``idx`` replaced by ``other``'s (same names)."""
        if other.names != self.names:
            raise ValueError("put needs the same names")
        shapes, data = list(self.shapes), list(self.data)
        for j, i in enumerate(idx):
            shapes[i] = other.shapes[j]
            data[i] = other.data[j]
        return TensorBatch(self.model, self.dtype, self.names, shapes, data)

    def rename(self, mapping):
        """safe_rename (tetragono/tetragono/utility.py:353-358): names absent
        from the batch are ignored; storage is shared."""
        m = {k: v for k, v in mapping.items() if k in self.names}
        if not m:
            return self
        old = list(m)
        o_ids = _lib.arr32(NAMES.ids(old))
        n_ids = _lib.arr32(NAMES.ids([m[k] for k in old]))
        L = _lib.lib()
        shapes = []
        for s in self.shapes:
            h = ctypes.c_void_p()
            _lib.check(L.tns_shape_rename(s.handle, len(old), o_ids, n_ids, ctypes.byref(h)))
            shapes.append(Shape(h, self.model.kind, lazy=True))
        return TensorBatch(self.model, self.dtype, [m.get(n, n) for n in self.names], shapes, self.data)

    def values(self):
        """First storage element of every item (0 for an empty tensor): the
        value of one-element tensors."""
        torch = _torch()
        out = np.zeros(len(self))
        nz = [i for i, d in enumerate(self.data) if d.numel() > 0]
        if nz:
            out[nz] = torch.stack([self.data[i][0] for i in nz]).double().cpu().numpy()
        return out


def _outputs(plans, which, model, dtype):
    """Output batches ``which`` of a list of plans: one device buffer, one
    view per item."""
    torch = _torch()
    res = []
    for w in which:
        hs = []
        for p in plans:
            h = ctypes.c_void_p()
            _lib.check(_lib.lib().tns_plan_output_shape(p.handle, w, ctypes.byref(h)))
            hs.append(h)
        shapes = [Shape(h, model.kind, lazy=True) for h in hs]
        names = shapes[0].names
        sizes = [s.size for s in shapes]
        buf = torch.empty(max(1, sum(sizes)), dtype=_tdtype(dtype), device=device())
        data, off = [], 0
        for n in sizes:
            data.append(buf[off:off + n])
            off += n
        res.append(TensorBatch(model, dtype, names, shapes, data))
    return res


def _workspaces(plans, dtype):
    torch = _torch()
    sizes = [p.workspace(dtype) for p in plans]
    offs, tot = [], 0
    for s in sizes:
        offs.append(tot)
        tot += (s + 255) // 256 * 256
    buf = torch.empty(max(256, tot), dtype=torch.uint8, device=device())
    base = buf.data_ptr()
    return buf, [base + o for o in offs]


def _run(plans, dtype, inputs, outputs, workspaces):
    _lib.check(_lib.lib().tns_plans_run(len(plans), _vp_array([p.handle.value for p in plans]), DTYPE_CODE[dtype],
                                        _vp_array(inputs), _vp_array(outputs), _vp_array(workspaces),
                                        ctypes.c_void_p(_stream())))


def _plans(handles, type_, kind):
    return [Plan(ctypes.c_void_p(h), type_, kind) for h in handles]


def _handles(shapes):
    return _vp_array([s.handle.value for s in shapes])


# -- host planning (no device needed) -------------------------------------------
def plan_contract(shapes_a, shapes_b, pairs):
    """One contract plan per item pair, planned on host threads."""
    n = len(shapes_a)
    if len(shapes_b) != n:
        raise ValueError("batches differ in length")
    flat = [i for x, y in pairs for i in NAMES.ids([x, y])]
    out = (ctypes.c_void_p * max(1, n))()
    _lib.check(_lib.lib().tns_plan_contract_many(n, _handles(shapes_a), _handles(shapes_b), len(pairs),
                                                 _lib.arr32(flat), _threads(), out))
    return _plans(out[:n], "contract", shapes_a[0].kind)


def plan_qr(shapes, direction, free_names, cq, cr):
    n = len(shapes)
    out = (ctypes.c_void_p * max(1, n))()
    _lib.check(_lib.lib().tns_plan_qr_many(n, _handles(shapes), direction.encode()[:1], len(free_names),
                                           _lib.arr32(NAMES.ids(free_names)), NAMES.id(cq), NAMES.id(cr),
                                           _threads(), out))
    return _plans(out[:n], "qr", shapes[0].kind)


def plan_svd(shapes, free_u, cu, cv, su, sv):
    n = len(shapes)
    out = (ctypes.c_void_p * max(1, n))()
    _lib.check(_lib.lib().tns_plan_svd_many(n, _handles(shapes), len(free_u), _lib.arr32(NAMES.ids(free_u)),
                                            NAMES.id(cu), NAMES.id(cv), NAMES.id(su), NAMES.id(sv), _threads(), out))
    return _plans(out[:n], "svd", shapes[0].kind)


def contract(a, b, pairs):
    """Item-wise ``a[i].contract(b[i], pairs)``; pairs whose names are absent
    on either side are dropped first (safe_contract,
    tetragono/tetragono/utility.py:340-350)."""
    n = len(a)
    pr = [(x, y) for x, y in pairs if x in a.names and y in b.names]
    plans = plan_contract(a.shapes, b.shapes, pr)
    (res,) = _outputs(plans, [0], a.model, a.dtype)
    ws, wps = _workspaces(plans, a.dtype)
    ins = [p for i in range(n) for p in (a.data[i].data_ptr(), b.data[i].data_ptr())]
    _run(plans, a.dtype, ins, [d.data_ptr() for d in res.data], wps)
    return res


def qr(t, direction, free_names, cq, cr):
    n = len(t)
    plans = plan_qr(t.shapes, direction, [x for x in free_names if x in t.names], cq, cr)
    q, r = _outputs(plans, [0, 1], t.model, t.dtype)
    ws, wps = _workspaces(plans, t.dtype)
    outs = [p for i in range(n) for p in (q.data[i].data_ptr(), r.data[i].data_ptr())]
    _run(plans, t.dtype, [d.data_ptr() for d in t.data], outs, wps)
    return q, r


def svd(t, free_u, cu, cv, su, sv, cut=-1):
    n = len(t)
    plans = plan_svd(t.shapes, [x for x in free_u if x in t.names], cu, cv, su, sv)
    ws, wps = _workspaces(plans, t.dtype)
    remain, relative = cut_params(cut)
    _lib.check(_lib.lib().tns_svd_factor_many(n, _vp_array([p.handle.value for p in plans]), DTYPE_CODE[t.dtype],
                                              _vp_array([d.data_ptr() for d in t.data]), _vp_array(wps), remain,
                                              relative, _threads(), ctypes.c_void_p(_stream())))
    u, s, v = _outputs(plans, [0, 1, 2], t.model, t.dtype)
    outs = [p for i in range(n) for p in (u.data[i].data_ptr(), s.data[i].data_ptr(), v.data[i].data_ptr())]
    _run(plans, t.dtype, [d.data_ptr() for d in t.data], outs, wps)
    return u, s, v
