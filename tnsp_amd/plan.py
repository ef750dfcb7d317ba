"""Plan objects over the C ABI: build (host planners), query, run (GPU).

Each ``Plan`` owns a ``tns_plan`` handle.  Building a plan only touches the
host; running it needs device buffers (torch tensors) and a HIP stream.
"""
import ctypes

import numpy as np

from . import _lib
from .shape import NAMES, Edge, Shape

DTYPE_CODE = {"float64": 0, "float32": 1}

PERM_FIELDS = 31
RECORD = {"perm": 31, "gemm": 15, "fill": 3, "qr": 9, "svd": 10}
# stage kinds per plan type (capi.cpp header)
STAGES = {
    "edge_operator": ["perm"],
    "conjugate": ["perm"],
    "contract": ["perm", "perm", "gemm", "fill", "perm"],
    "svd": ["perm", "svd", "perm", "perm", "fill", "perm"],
    "qr": ["perm", "qr", "perm", "perm"],
}


class Plan:
    __slots__ = ("handle", "type", "kind")

    def __init__(self, handle, type_, kind):
        self.handle = handle
        self.type = type_
        self.kind = kind

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and _lib._lib is not None:
            _lib._lib.tns_plan_free(h)
            self.handle = None

    # -- queries ------------------------------------------------------------
    def num_outputs(self):
        n = ctypes.c_int()
        _lib.check(_lib.lib().tns_plan_num_outputs(self.handle, ctypes.byref(n)))
        return n.value

    def output(self, which):
        h = ctypes.c_void_p()
        _lib.check(_lib.lib().tns_plan_output_shape(self.handle, which, ctypes.byref(h)))
        return Shape(h, self.kind)

    def workspace(self, dtype):
        b = ctypes.c_int64()
        _lib.check(_lib.lib().tns_plan_workspace(self.handle, DTYPE_CODE[dtype], ctypes.byref(b)))
        return b.value

    def flops(self):
        f = ctypes.c_int64()
        _lib.check(_lib.lib().tns_plan_flops(self.handle, ctypes.byref(f)))
        return f.value

    def buffers(self):
        n = ctypes.c_int()
        sizes = (ctypes.c_int64 * 16)()
        _lib.check(_lib.lib().tns_plan_buffers(self.handle, sizes, ctypes.byref(n)))
        return [sizes[i] for i in range(n.value)]

    def stage(self, index):
        """Task records of a stage as an int64 array (rows = tasks)."""
        n = ctypes.c_int64()
        L = _lib.lib()
        _lib.check(L.tns_plan_stage_count(self.handle, index, ctypes.byref(n)))
        width = RECORD[STAGES[self.type][index]]
        out = np.zeros((n.value, width), dtype=np.int64)
        if n.value:
            _lib.check(L.tns_plan_stage_export(self.handle, index, out.ctypes.data_as(_lib.c_i64p)))
        return out

    def svd_sectors(self):
        n = ctypes.c_int()
        L = _lib.lib()
        _lib.check(L.tns_svd_sectors(self.handle, None, None, None, ctypes.byref(n)))
        k = np.zeros(max(1, n.value), np.int64)
        off = np.zeros(max(1, n.value), np.int64)
        _lib.check(L.tns_svd_sectors(self.handle, k.ctypes.data_as(_lib.c_i64p), off.ctypes.data_as(_lib.c_i64p),
                                     None, ctypes.byref(n)))
        return k[:n.value], off[:n.value]

    def svd_truncate_host(self, svals, remain_cut, relative_cut):
        svals = np.ascontiguousarray(svals, dtype=np.float64)
        _lib.check(_lib.lib().tns_svd_truncate(self.handle, svals.ctypes.data_as(_lib.c_dblp), int(remain_cut),
                                               float(relative_cut)))

    # -- execution ------------------------------------------------------------
    def run(self, dtype, inputs, outputs, workspace, stream):
        ins = (ctypes.c_void_p * max(1, len(inputs)))(*[int(p) for p in inputs])
        outs = (ctypes.c_void_p * max(1, len(outputs)))(*[int(p) for p in outputs])
        _lib.check(_lib.lib().tns_plan_run(self.handle, DTYPE_CODE[dtype], ins, outs, ctypes.c_void_p(int(workspace)),
                                           ctypes.c_void_p(int(stream))))

    def run_noalloc(self, dtype, inputs, outputs, stream):
        """run() with the workspace and task records in the library's static
        device array (tns_plan_run_noalloc): no device allocation."""
        ins = (ctypes.c_void_p * max(1, len(inputs)))(*[int(p) for p in inputs])
        outs = (ctypes.c_void_p * max(1, len(outputs)))(*[int(p) for p in outputs])
        _lib.check(_lib.lib().tns_plan_run_noalloc(self.handle, DTYPE_CODE[dtype], ins, outs,
                                                   ctypes.c_void_p(int(stream))))

    def svd_factor(self, dtype, inp, workspace, remain_cut, relative_cut, stream):
        _lib.check(_lib.lib().tns_svd_factor(self.handle, DTYPE_CODE[dtype], ctypes.c_void_p(int(inp)),
                                             ctypes.c_void_p(int(workspace)), int(remain_cut), float(relative_cut),
                                             ctypes.c_void_p(int(stream))))


def _ids(names):
    return NAMES.ids(list(names))


def edge_operator(shape, split_map=None, reversed_names=(), merge_map=None, new_names=(), apply_parity=False,
                  ex_split=(), ex_rev_before=(), ex_rev_after=(), ex_merge=()):
    """TAT edge_operator (TAT/include/TAT/implement/edge_operator.hpp:35-692).

    split_map: {old: [(new, edge_spec), ...]}, merge_map: {new: [old, ...]}."""
    split_map = split_map or {}
    merge_map = merge_map or {}
    kind = shape.kind
    s_old, s_cnt, s_names, s_edges = [], [], [], []
    for old, parts in split_map.items():
        s_old.append(old)
        s_cnt.append(len(parts))
        for nm, e in parts:
            s_names.append(nm)
            s_edges.append(Edge.build(kind, e))
    split_shape = Shape.create(kind, [f"\x01split{i}" for i in range(len(s_edges))], s_edges) if s_edges else None
    m_new, m_cnt, m_old = [], [], []
    for new, olds in merge_map.items():
        m_new.append(new)
        m_cnt.append(len(olds))
        m_old.extend(olds)
    h = ctypes.c_void_p()
    rev = list(reversed_names)
    L = _lib.lib()
    _lib.check(L.tns_plan_edge_operator(
        shape.handle, len(s_old), _lib.arr32(_ids(s_old)), _lib.arr32(s_cnt), _lib.arr32(_ids(s_names)),
        split_shape.handle if split_shape else None, len(rev), _lib.arr32(_ids(rev)), len(m_new),
        _lib.arr32(_ids(m_new)), _lib.arr32(m_cnt), _lib.arr32(_ids(m_old)), len(new_names),
        _lib.arr32(_ids(new_names)), int(bool(apply_parity)), len(ex_split), _lib.arr32(_ids(ex_split)),
        len(ex_rev_before), _lib.arr32(_ids(ex_rev_before)), len(ex_rev_after), _lib.arr32(_ids(ex_rev_after)),
        len(ex_merge), _lib.arr32(_ids(ex_merge)), ctypes.byref(h)))
    return Plan(h, "edge_operator", kind)


def contract(a, b, pairs):
    """TAT contract (TAT/include/TAT/implement/contract.hpp:307-620)."""
    flat = []
    for x, y in pairs:
        flat.extend([NAMES.id(x), NAMES.id(y)])
    h = ctypes.c_void_p()
    _lib.check(_lib.lib().tns_plan_contract(a.handle, b.handle, len(pairs), _lib.arr32(flat), ctypes.byref(h)))
    return Plan(h, "contract", a.kind)


def svd(t, free_u, cu, cv, su, sv):
    """TAT svd (TAT/include/TAT/implement/svd.hpp:260-538), factor phase."""
    free_u = list(free_u)
    h = ctypes.c_void_p()
    _lib.check(_lib.lib().tns_plan_svd(t.handle, len(free_u), _lib.arr32(_ids(free_u)), NAMES.id(cu), NAMES.id(cv),
                                       NAMES.id(su), NAMES.id(sv), ctypes.byref(h)))
    return Plan(h, "svd", t.kind)


def qr(t, direction, free_names, cq, cr):
    """TAT qr (TAT/include/TAT/implement/qr.hpp:310-508)."""
    free_names = list(free_names)
    h = ctypes.c_void_p()
    _lib.check(_lib.lib().tns_plan_qr(t.handle, direction.encode()[:1], len(free_names), _lib.arr32(_ids(free_names)),
                                      NAMES.id(cq), NAMES.id(cr), ctypes.byref(h)))
    return Plan(h, "qr", t.kind)


def conjugate(t, trivial_metric=False):
    """TAT conjugate (TAT/include/TAT/implement/conjugate.hpp:32-120)."""
    h = ctypes.c_void_p()
    _lib.check(_lib.lib().tns_plan_conjugate(t.handle, int(bool(trivial_metric)), ctypes.byref(h)))
    return Plan(h, "conjugate", t.kind)


def cut_params(cut):
    """PyTAT.hpp:1040-1047: cut >= 1 -> keep count, 0 < cut < 1 -> relative."""
    if cut is None or cut <= 0:
        return -1, 0.0
    if cut >= 1:
        return int(cut), 0.0
    return -1, float(cut)
