"""Python handle of the batched amplitude engine (libtnsgpu ``tns_engine_*``).

One engine holds one PEPS.  ``evaluate(configs)`` returns psi(s) for every
configuration: the native engine steps all of them through tetragono's
single-layer boundary-MPS op sequence in lockstep
(tetragono/tetragono/auxiliaries/single_layer_auxiliaries.py:463-498 per row,
:254-312 and :710-720 for the final row), planning each op per configuration
on host threads and executing one batched launch per kernel kind.
"""
import ctypes
import os

import numpy as np

from . import _lib
from .plan import DTYPE_CODE, cut_params
from .shape import NAMES
from .tensor import _stream

STAT_NAMES = ["plan_s", "exec_s", "total_s", "gemm_flops", "qr_flops", "svd_flops", "perm_elems", "launches",
              "device_bytes", "configs", "bind_s", "sync_s"]


class AmplitudeEngine:
    PRECISIONS = {"fp64": 0, "fp32": 1, "bf16x3": 2, "bf16x6": 3}

    def __init__(self, lattice, threads=None, chunk=1024, lanes=1, precision="fp64"):
        spec = lattice.spec
        self.lattice = lattice  # keeps the site buffers alive
        self.spec = spec
        sites = [lattice.sites[l1][l2] for l1 in range(spec.L1) for l2 in range(spec.L2)]
        shapes = (ctypes.c_void_p * len(sites))(*[t._shape.handle for t in sites])
        datas = (ctypes.c_void_p * len(sites))(*[t._data.data_ptr() for t in sites])
        names = ["P0", "T", "U", "D", "L", "R", "L1", "L2", "L3", "R1", "R2", "P", "Q"]
        names += [f"P_{l1}_{l2}_0" for l1 in range(spec.L1) for l2 in range(spec.L2)]
        if threads is None:
            threads = int(os.environ.get("TNS_THREADS", min(16, os.cpu_count() or 1)))
        h = ctypes.c_void_p()
        _lib.check(_lib.lib().tns_engine_create(sites[0].model.kind.id, spec.L1, spec.L2, shapes, datas,
                                                DTYPE_CODE[lattice.dtype], _lib.arr32(NAMES.ids(names)), len(names),
                                                int(threads), ctypes.byref(h)))
        self.handle = h
        self.set_chunk(chunk)
        self.set_lanes(lanes)
        self.set_precision(precision)

    def set_precision(self, precision):
        """GEMM precision of the contractions: "fp64", "fp32", "bf16x3" or
        "bf16x6" (split-bf16 products with fp32 accumulation).  QR, SVD and
        all stored tensors stay float64."""
        if precision not in self.PRECISIONS:
            raise ValueError(f"precision must be one of {list(self.PRECISIONS)}")
        _lib.check(_lib.lib().tns_engine_set_precision(self.handle, self.PRECISIONS[precision]))
        self.precision = precision

    def set_lanes(self, lanes):
        """Number of concurrent lockstep lanes (own host thread + stream)."""
        _lib.check(_lib.lib().tns_engine_set_lanes(self.handle, int(lanes)))

    def set_chunk(self, chunk):
        _lib.check(_lib.lib().tns_engine_set_chunk(self.handle, int(chunk)))

    def autotune(self, configs, cut=None):
        """Pick the GEMM tile variant per sector size class by timing every
        variant on the first contraction of ``configs`` (see
        ``tns_engine_autotune``); returns (table, ms per class and variant)."""
        cfg = np.ascontiguousarray(np.asarray(configs), dtype=np.int32)
        remain, relative = cut_params(self.spec.cut if cut is None else cut)
        table = (ctypes.c_int32 * 3)()
        ms = (ctypes.c_double * 9)()
        _lib.check(_lib.lib().tns_engine_autotune(self.handle, cfg.shape[0], cfg.ctypes.data_as(_lib.c_i32p), remain,
                                                  relative, table, ms, ctypes.c_void_p(_stream())))
        return list(table), np.array(list(ms)).reshape(3, 3)

    def evaluate(self, configs, cut=None):
        """psi for configurations of shape (n, L1, L2) (physical indices)."""
        cfg = np.ascontiguousarray(np.asarray(configs), dtype=np.int32)
        n = cfg.shape[0]
        remain, relative = cut_params(self.spec.cut if cut is None else cut)
        psi = np.zeros(n, dtype=np.float64)
        _lib.check(_lib.lib().tns_engine_evaluate(self.handle, n, cfg.ctypes.data_as(_lib.c_i32p), remain, relative,
                                                  psi.ctypes.data_as(_lib.c_dblp), ctypes.c_void_p(_stream())))
        return psi

    def refresh_sites(self):
        """The lattice's site tensors changed in place (``Lattice.apply_gradient``):
        project them again on the next evaluation."""
        _lib.check(_lib.lib().tns_engine_refresh_sites(self.handle))

    def trim(self):
        """Free the working device memory (it grows again on the next
        evaluation; the projected sites stay)."""
        _lib.check(_lib.lib().tns_engine_trim(self.handle))

    def stats(self):
        vals = (ctypes.c_double * len(STAT_NAMES))()
        _lib.check(_lib.lib().tns_engine_stats(self.handle, vals, len(STAT_NAMES)))
        return dict(zip(STAT_NAMES, list(vals)))

    KERNEL_KINDS = ["permute", "fill", "gemm", "qr", "svd", "diag", "qr_blocked", "svd_reg8", "svd_reg16", "svd_reg24"]

    def set_profile(self, on=True):
        _lib.check(_lib.lib().tns_engine_set_profile(self.handle, int(bool(on))))

    def kernel_stats(self):
        """{kernel: {"launches", "ms", "units"}}; units are bytes for permute /
        fill / diag and flops for gemm and the QR / SVD kernels."""
        n = 3 * len(self.KERNEL_KINDS)
        vals = (ctypes.c_double * n)()
        _lib.check(_lib.lib().tns_engine_kernel_stats(self.handle, vals, n))
        return {k: {"launches": vals[3 * i], "ms": vals[3 * i + 1], "units": vals[3 * i + 2]}
                for i, k in enumerate(self.KERNEL_KINDS)}

    def reset_stats(self):
        _lib.check(_lib.lib().tns_engine_reset_stats(self.handle))

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and _lib._lib is not None:
            _lib._lib.tns_engine_free(h)
            self.handle = None
