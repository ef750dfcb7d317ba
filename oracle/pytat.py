"""PyTAT-shaped facade over the CPU oracle (TEST INFRASTRUCTURE ONLY).

Mirrors the module layout of the reference binding (PyTAT/PyTAT.cpp:68-159):
``No``, ``BoseZ2``, ``BoseU1``, ``FermiU1``, ``FermiZ2``, ``FermiU1BoseZ2``,
``FermiU1BoseU1``, ``FermiU1FermiU1``, each with ``.D`` / ``.S`` submodules
exposing ``Tensor`` and with ``Edge`` / ``Symmetry`` factories, so restated
reference tests can run unchanged against the oracle and against the GPU
package ``tnsp_amd.TAT``.
"""
import numpy as np

from . import tat


class _ScalarModule:
    def __init__(self, sym_module, dtype, btype):
        self.model = sym_module
        self.dtype = dtype
        self.btype = btype
        kind = sym_module.kind

        def make_tensor(*args, **kwargs):
            if args and isinstance(args[0], (int, float, np.floating, np.integer)) and not isinstance(args[0], bool):
                number = args[0]
                names = kwargs.get("names", args[1] if len(args) > 1 else [])
                syms = kwargs.get("edge_symmetry", args[2] if len(args) > 2 else [])
                arrows = kwargs.get("edge_arrow", args[3] if len(args) > 3 else [])
                return tat.Tensor.scalar(kind, number, names, [kind.make(s) for s in syms], arrows, dtype)
            names = kwargs.get("names", args[0] if args else [])
            edges = kwargs.get("edges", args[1] if len(args) > 1 else [])
            return tat.Tensor(kind, names, edges, dtype)

        self.Tensor = make_tensor


class _SymmetryModule:
    def __init__(self, kind):
        self.kind = kind
        self.D = self.float64 = _ScalarModule(self, np.float64, "D")
        self.S = self.float32 = _ScalarModule(self, np.float32, "S")
        self.Tensor = self.D.Tensor

    def Edge(self, *args):
        if len(args) == 2:
            return tat.Edge.build(self.kind, (list(args[0]) if not isinstance(args[0], int) else args[0], bool(args[1])))
        return tat.Edge.build(self.kind, args[0])

    def Symmetry(self, *args):
        return self.kind.make(args if len(args) != 1 else args[0])


No = _SymmetryModule(tat.No)
BoseZ2 = _SymmetryModule(tat.BoseZ2)
BoseU1 = _SymmetryModule(tat.BoseU1)
FermiU1 = _SymmetryModule(tat.FermiU1)
FermiZ2 = _SymmetryModule(tat.FermiZ2)
FermiU1BoseZ2 = _SymmetryModule(tat.FermiU1BoseZ2)
FermiU1BoseU1 = _SymmetryModule(tat.FermiU1BoseU1)
FermiU1FermiU1 = _SymmetryModule(tat.FermiU1FermiU1)
Normal = No
Z2 = BoseZ2
U1 = BoseU1
Fermi = FermiU1
Parity = FermiZ2
