"""Drop-in module shaped like the reference binding ``TAT``
(PyTAT/PyTAT.cpp:58-166): ``TAT.<Symmetry>.<D|S>.Tensor``, ``.Edge``,
``.Symmetry`` and ``TAT.random``, for the symmetries No, BoseZ2, BoseU1,
FermiU1, FermiZ2, FermiU1BoseZ2, FermiU1BoseU1, FermiU1FermiU1 and the
scalar types float64 (D) and float32 (S).  Tensors live on the GPU.
"""
import numpy as np

from . import rng as _rng
from .shape import KINDS, Edge as _Edge
from .tensor import Tensor as _Tensor


class _ScalarModule:
    def __init__(self, model, dtype, btype):
        self.model = model
        self.dtype = dtype
        self.btype = btype

    def Tensor(self, *args, **kwargs):
        m, dt = self.model, self.dtype
        if args and isinstance(args[0], (int, float, np.integer, np.floating)) and not isinstance(args[0], bool):
            number = float(args[0])
            names = list(kwargs.get("names", args[1] if len(args) > 1 else []))
            syms = list(kwargs.get("edge_symmetry", args[2] if len(args) > 2 else []))
            arrows = list(kwargs.get("edge_arrow", args[3] if len(args) > 3 else []))
            syms = syms or [m.kind.zero()] * len(names)
            arrows = arrows or [False] * len(names)
            edges = [_Edge(m.kind, [(s, 1)], a) for s, a in zip(syms, arrows)]
            t = _Tensor(m, dt, names, edges)
            if t._data.numel() != 1:
                raise ValueError("symmetries of a one-element tensor must sum to zero")
            t._data.fill_(number)
            return t
        names = kwargs.get("names", args[0] if args else [])
        edges = kwargs.get("edges", args[1] if len(args) > 1 else [])
        return _Tensor(m, dt, list(names), list(edges))


class _SymmetryModule:
    def __init__(self, name):
        self.name = name
        self.kind = KINDS[name]
        self.D = self.float64 = _ScalarModule(self, "float64", "D")
        self.S = self.float32 = _ScalarModule(self, "float32", "S")

    def Tensor(self, *args, **kwargs):
        return self.D.Tensor(*args, **kwargs)

    def Edge(self, *args):
        if len(args) == 2:
            return _Edge.build(self.kind, (args[0], bool(args[1])))
        return _Edge.build(self.kind, args[0])

    def Symmetry(self, *args):
        return self.kind.display(self.kind.make(args[0] if len(args) == 1 else args))


No = _SymmetryModule("No")
BoseZ2 = _SymmetryModule("BoseZ2")
BoseU1 = _SymmetryModule("BoseU1")
FermiU1 = _SymmetryModule("FermiU1")
FermiZ2 = _SymmetryModule("FermiZ2")
FermiU1BoseZ2 = _SymmetryModule("FermiU1BoseZ2")
FermiU1BoseU1 = _SymmetryModule("FermiU1BoseU1")
FermiU1FermiU1 = _SymmetryModule("FermiU1FermiU1")
Normal = No
Z2 = BoseZ2
U1 = BoseU1
Fermi = FermiU1
Parity = FermiZ2


class random:
    seed = staticmethod(_rng.seed)
    uniform_int = staticmethod(_rng.uniform_int)
    uniform_real = staticmethod(_rng.uniform_real)
    normal = staticmethod(_rng.normal_real)
