"""Reading tetragono's saved SamplingLattice (tnsp_amd/latticeio.py).

The reference tree holds no pickled lattice, so the files here are written
by CPython's pickler from stand-in classes registered under the qualified
names the real ones carry (``TAT.FermiU1.D.Tensor``, ``TAT.FermiU1.Edge``,
``TAT.FermiU1.Symmetry``, ``tetragono.sampling_lattice.lattice.
SamplingLattice``), whose ``__getstate__`` gives what the real ones give:
the tensor's binary dump (PyTAT/PyTAT.hpp:768-771), the edge's and the
symmetry's binary items (PyTAT/PyTAT.hpp:175-188, 295-307) and the lattice's
state dict (sampling_lattice/lattice.py:739-742).  The container is CPython's
protocol for such objects (NEWOBJ of the class, then BUILD with the state),
and the tensor bytes are TAT's format as pinned by the golden dumps."""
import pickle
import struct
import sys
import types

import numpy as np
import pytest

from tnsp_amd import latticeio, peps, tatio
from tnsp_amd.shape import KINDS, Edge, Shape

K = KINDS["FermiU1"]


def _modules(names):
    saved = {n: sys.modules.get(n) for n in names}
    for n in names:
        sys.modules[n] = types.ModuleType(n)
    return saved


def _restore(saved):
    for n, m in saved.items():
        if m is None:
            sys.modules.pop(n, None)
        else:
            sys.modules[n] = m


def _stand_in(module, name, state):
    cls = getattr(sys.modules[module], name, None)
    if cls is None:
        cls = type(name, (), {"__getstate__": lambda self: self._state, "__module__": module, "__qualname__": name})
        setattr(sys.modules[module], name, cls)
    obj = object.__new__(cls)
    obj._state = state
    return obj, cls


def _edge_bytes(e):
    out = bytearray(b"\x01" if e.arrow else b"\x00")
    out += struct.pack("<Q", len(e.segments))
    for q, d in e.segments:
        out += tatio._segment_bytes(K, q, d)
    return bytes(out)


def _lattice_pickle(spec, storages, version=6, symmetry_module="FermiU1", protocol=4):
    """Pickle bytes of a SamplingLattice state in tetragono's layout."""
    mods = ["TAT", f"TAT.{symmetry_module}", f"TAT.{symmetry_module}.D", "tetragono", "tetragono.sampling_lattice",
            "tetragono.sampling_lattice.lattice"]
    saved = _modules(mods)
    try:
        tmod, emod = f"TAT.{symmetry_module}.D", f"TAT.{symmetry_module}"
        sites, vbonds, pedges = [], [], []
        tensor_cls = None
        for l1 in range(spec.L1):
            srow, vrow, prow = [], [], []
            for l2 in range(spec.L2):
                names, edges = peps.site_layout(spec, l1, l2)
                t, tensor_cls = _stand_in(tmod, "Tensor", tatio.serialize(K, names, edges,
                                                                           storages[l1 * spec.L2 + l2]))
                srow.append(t)
                bonds = {}
                for n, e in zip(names, edges):
                    if n in ("U", "D", "L", "R"):
                        bonds[n] = _stand_in(emod, "Edge", _edge_bytes(e))[0]
                vrow.append(bonds)
                prow.append({0: _stand_in(emod, "Edge", _edge_bytes(edges[0]))[0]})
            sites.append(srow)
            vbonds.append(vrow)
            pedges.append(prow)
        ham_names = ["O0", "I0", "O1", "I1"]
        ham_edges = [Edge(K, [(0, 1), (1, 1)], False), Edge(K, [(0, 1), (-1, 1)], True)] * 2
        nh = Shape.create(K, ham_names, ham_edges).size
        ham = _stand_in(tmod, "Tensor", tatio.serialize(K, ham_names, ham_edges, np.arange(1.0, nh + 1)))[0]
        state = {"Tensor": tensor_cls, "L1": spec.L1, "L2": spec.L2, "_physics_edges": pedges,
                 "_hamiltonians": {((0, 0, 0), (0, 1, 0)): ham},
                 "_total_symmetry": _stand_in(emod, "Symmetry", struct.pack("<i", spec.N))[0],
                 "_site_number": spec.L1 * spec.L2, "data_version": version, "attribute": {"alpha": 1.0},
                 "_virtual_bond": vbonds, "_lattice": sites}
        if version <= 4:
            state["_virtual_bond"] = {(l1, l2): vbonds[l1][l2] for l1 in range(spec.L1) for l2 in range(spec.L2)}
            del state["attribute"]
        if version <= 2:
            for n in ("L1", "L2", "Tensor"):
                state["_" + n] = state.pop(n)
        if version == 0:
            del state["data_version"]
            state["_cut_dimension"] = 7
        lat, _ = _stand_in("tetragono.sampling_lattice.lattice", "SamplingLattice", state)
        return pickle.dumps(lat, protocol=protocol)
    finally:
        _restore(saved)


SPEC = peps.LatticeSpec("FermiU1", 2, 3, (1, 2, 1), 3, 8, seed=5)


@pytest.mark.parametrize("version", [6, 4, 2, 0])
@pytest.mark.parametrize("protocol", [2, 4, 5])
def test_saved_lattice_round_trip(version, protocol):
    storages = peps.random_storages(SPEC)
    got = latticeio.loads(_lattice_pickle(SPEC, storages, version=version, protocol=protocol))
    assert isinstance(got, latticeio.SavedLattice)
    assert (got.kind, got.dtype, got.L1, got.L2, got.data_version) == ("FermiU1", "float64", 2, 3, 6)
    assert got.attribute == ({"alpha": 1.0} if version >= 5 else {})
    assert "_cut_dimension" not in got.__dict__
    assert got._total_symmetry == (SPEC.N,)
    for l1 in range(SPEC.L1):
        for l2 in range(SPEC.L2):
            names, edges = peps.site_layout(SPEC, l1, l2)
            site = got.site(l1, l2)
            assert site.names == names and site.edges == edges
            assert np.array_equal(site.storage, storages[l1 * SPEC.L2 + l2])
            assert got._virtual_bond[l1][l2] == {n: e for n, e in zip(names, edges) if n in ("U", "D", "L", "R")}
            assert got._physics_edges[l1][l2] == {0: edges[0]}
    ham = got._hamiltonians[((0, 0, 0), (0, 1, 0))]
    assert ham.names == ["O0", "I0", "O1", "I1"] and ham.storage[0] == 1.0


def test_saved_lattice_legacy_module_names():
    """Files from before TAT 0.3.15 name ``TAT.Fermi``: read as FermiU1
    (utility.py:178-197, without PYTAT_OLD_SYMMETRY_NAME)."""
    storages = peps.random_storages(SPEC)
    got = latticeio.loads(_lattice_pickle(SPEC, storages, symmetry_module="Fermi"))
    assert got.kind == "FermiU1"
    assert np.array_equal(got.site(1, 2).storage, storages[-1])


def test_foreign_globals_are_refused():
    """A pickle naming anything else is refused before any call: no
    function from the file runs."""
    calls = []

    class Evil:
        def __reduce__(self):
            return (print, ("ran",))

    with pytest.raises(pickle.UnpicklingError):
        latticeio.loads(pickle.dumps(Evil()))
    with pytest.raises(pickle.UnpicklingError):
        latticeio.loads(pickle.dumps({"x": calls.append}))
    assert calls == []
    with pytest.raises(pickle.UnpicklingError):
        latticeio.loads(pickle.dumps(np.zeros(3)))  # numpy's reconstructor is not part of a lattice


def test_read_lattice_refuses_other_objects(tmp_path):
    p = tmp_path / "x.dat"
    p.write_bytes(pickle.dumps([1, 2, 3]))
    with pytest.raises(ValueError):
        latticeio.read_lattice(str(p))


@pytest.mark.gpu
def test_saved_lattice_drives_the_engine():
    """A lattice read from a saved file gives the amplitudes of the lattice
    it was saved from (4x4 FermiU1, D = 4, cut 16), bit for bit."""
    from tnsp_amd.engine import AmplitudeEngine

    spec = peps.C2
    storages = peps.random_storages(spec)
    saved = latticeio.loads(_lattice_pickle(spec, storages))
    lat = peps.Lattice.from_saved(saved, cut=spec.cut)
    assert (lat.spec.N, lat.spec.L1, lat.spec.L2, lat.spec.kind) == (spec.N, 4, 4, "FermiU1")
    cfg = peps.sample_configurations(spec, 16, seed=3)
    got = AmplitudeEngine(lat, chunk=16).evaluate(cfg)
    want = AmplitudeEngine(peps.Lattice(spec, storages), chunk=16).evaluate(cfg)
    assert np.array_equal(got, want) and np.any(got != 0)
