"""tetragono's saved sampling lattices, read without running the file.

tetragono saves a ``SamplingLattice`` with ``pickle.dump``
(tetragono/tetragono/utility.py:217-226, paths relative to the TNSP
repository root) and reads it back with a plain unpickler that also renames
the symmetry modules of TAT before 0.3.15 (``RenamingUnpickler``,
utility.py:171-214).  A pickle may name any importable callable, so this
reader resolves only the globals such a file holds, each to a class of its
own that takes the pickled state as data; every other global is refused
before anything is called:

* ``TAT.<symmetry>.<scalar>.Tensor``: pybind11 pickles a PyTAT tensor
  through ``__getstate__`` as its binary dump (PyTAT/PyTAT.hpp:768-771),
  parsed here by ``tatio.parse``;
* ``TAT.<symmetry>.Edge`` and ``TAT.<symmetry>.Symmetry``: their binary
  items (PyTAT/PyTAT.hpp:175-188, 295-307; TAT/include/TAT/miscellaneous/
  io.hpp:456-473): the arrow byte for fermionic symmetries, then the
  segment list, and the raw symmetry;
* ``tetragono.sampling_lattice.lattice.SamplingLattice``: its state dict
  (``__getstate__``, lattice.py:739-742), brought to data version 6 by the
  upgrade steps of ``__setstate__`` (lattice.py:706-737, with
  abstract_state.py:288-308 and abstract_lattice.py:116-126);
* ``_codecs.encode`` with "latin1", which protocol 2 uses for bytes.

The site tensors stay host data until ``to_tensor`` / ``peps.Lattice``
moves them to the GPU.
"""
import io
import pickle
import re
import struct


from . import tatio
from .shape import KINDS, Edge

# TAT 0.3.15 renamed the symmetry modules (utility.py:178-184)
_LEGACY = {"Z2": "BoseZ2", "U1": "BoseU1", "Fermi": "FermiU1", "Parity": "FermiZ2",
           "FermiFermi": "FermiU1FermiU1"}
_SCALARS = {"D": "float64", "S": "float32"}
_LATTICE_MODULES = ("tetragono.sampling_lattice.lattice", "tetragono.sampling_lattice")


def _kind_name(name):
    name = _LEGACY.get(name, name)
    if name not in KINDS:
        raise pickle.UnpicklingError(f"unknown TAT symmetry module {name}")
    return name


class SavedTensor:
    """A PyTAT tensor from a saved file: ``names``, ``edges`` (``shape.Edge``)
    and ``storage`` (numpy, TAT's storage order) on the host."""
    kind_name = None
    dtype = None

    def __setstate__(self, state):
        if not isinstance(state, (bytes, bytearray)):
            raise pickle.UnpicklingError("a TAT tensor's state is its binary dump")
        kind = KINDS[self.kind_name]
        names, edges, storage, version = tatio.parse(kind, state, self.dtype)
        if version == 0:
            from .shape import Shape

            blocks = Shape.create(kind, names, edges).blocks()
            storage = tatio.v0_to_v1(blocks, storage)
        self.names, self.edges, self.storage = names, edges, storage

    def to_tensor(self):
        """The drop-in ``Tensor`` (on the GPU) with this structure and data."""
        from . import TAT

        mod = getattr(TAT, self.kind_name)
        t = (mod.D if self.dtype == "float64" else mod.S).Tensor(self.names, self.edges)
        t.storage = self.storage
        return t


class SavedEdge:
    kind_name = None

    def __setstate__(self, state):
        kind = KINDS[self.kind_name]
        data, pos = bytes(state), 0
        self.arrow = False
        if kind.is_fermi:
            self.arrow = data[0] != 0
            pos = 1
        (count,) = struct.unpack_from("<Q", data, pos)
        pos += 8
        segs = []
        for _ in range(count):
            segs.append(tatio._segment_value(kind, data[pos:pos + 16]))
            pos += 16
        if pos != len(data):
            raise pickle.UnpicklingError("trailing bytes after a TAT edge")
        self.edge = Edge(kind, segs, self.arrow)


class SavedSymmetry:
    kind_name = None

    def __setstate__(self, state):
        kind = KINDS[self.kind_name]
        size, offs = tatio._sym_layout(kind)
        data = bytes(state)
        if len(data) != size:
            raise pickle.UnpicklingError("a TAT symmetry of the wrong size")
        self.value = tuple(int(data[o] != 0) if z2 else struct.unpack_from("<i", data, o)[0] for o, z2 in offs)


class SavedLattice:
    """The state of a saved ``SamplingLattice`` at data version 6:
    ``Tensor`` (the tensor type), ``L1``, ``L2``, ``_physics_edges``,
    ``_hamiltonians``, ``_total_symmetry``, ``_site_number``,
    ``attribute``, ``_virtual_bond`` and ``_lattice`` (site tensors)."""

    def __setstate__(self, state):
        # before the data_version mechanism the state was (None, state)
        if isinstance(state, tuple):
            state = state[1]
        if not isinstance(state, dict):
            raise pickle.UnpicklingError("a SamplingLattice's state is a dict")
        state = dict(state)
        state.setdefault("data_version", 0)
        if state["data_version"] == 0:
            state["data_version"] = 1
            state.pop("_cut_dimension", None)
        if state["data_version"] == 1:
            state["data_version"] = 2
        if state["data_version"] == 2:
            for n in ("L1", "L2", "Tensor"):
                state[n] = state.pop("_" + n)
            state["data_version"] = 3
        if state["data_version"] == 3:
            state["data_version"] = 4
        if state["data_version"] == 4:
            vb = state["_virtual_bond"]
            if isinstance(vb, dict):
                state["_virtual_bond"] = [[vb[l1, l2] for l2 in range(state["L2"])] for l1 in range(state["L1"])]
            state["data_version"] = 5
        if state["data_version"] == 5:
            state["attribute"] = {}
            state["data_version"] = 6
        self.__dict__.update(state)

    @property
    def kind(self):
        return self.Tensor.kind_name

    @property
    def dtype(self):
        return self.Tensor.dtype

    def site(self, l1, l2):
        return self._lattice[l1][l2]


_TYPES = {}


def _stand_in(base, kind_name, dtype=None):
    key = (base, kind_name, dtype)
    if key not in _TYPES:
        _TYPES[key] = type(f"{base.__name__}[{kind_name}{'.' + dtype if dtype else ''}]", (base,),
                           {"kind_name": kind_name, "dtype": dtype})
    return _TYPES[key]


def _latin1(text, encoding):
    if not isinstance(text, str) or encoding != "latin1":
        raise pickle.UnpicklingError("_codecs.encode is allowed for protocol 2 bytes only")
    return text.encode("latin1")


class _Reader(pickle.Unpickler):
    def find_class(self, module, name):
        m = re.fullmatch(r"TAT\.(\w+)\.(\w+)", module)
        if m and name == "Tensor" and m.group(2) in _SCALARS:
            return _stand_in(SavedTensor, _kind_name(m.group(1)), _SCALARS[m.group(2)])
        m = re.fullmatch(r"TAT\.(\w+)", module)
        if m and name == "Edge":
            return _stand_in(SavedEdge, _kind_name(m.group(1)))
        if m and name == "Symmetry":
            return _stand_in(SavedSymmetry, _kind_name(m.group(1)))
        if module in _LATTICE_MODULES and name == "SamplingLattice":
            return SavedLattice
        if module == "_codecs" and name == "encode":
            # protocol 2 writes bytes as _codecs.encode(text, "latin1")
            return _latin1
        raise pickle.UnpicklingError(f"global {module}.{name} is not part of a saved lattice; refused")


def _plain(x):
    """Edges and symmetries as ``shape.Edge`` / symmetry tuples, recursively."""
    if isinstance(x, SavedEdge):
        return x.edge
    if isinstance(x, SavedSymmetry):
        return x.value
    if isinstance(x, list):
        return [_plain(v) for v in x]
    if isinstance(x, tuple):
        return tuple(_plain(v) for v in x)
    if isinstance(x, dict):
        return {_plain(k): _plain(v) for k, v in x.items()}
    return x


def loads(data):
    """A saved lattice (or any object made of the classes above) from
    pickle bytes."""
    obj = _Reader(io.BytesIO(bytes(data))).load()
    if isinstance(obj, SavedLattice):
        for k, v in list(obj.__dict__.items()):
            obj.__dict__[k] = _plain(v)
    return _plain(obj)


def read_lattice(path):
    """``utility.read_from_file`` for a saved SamplingLattice."""
    with open(path, "rb") as f:
        obj = loads(f.read())
    if not isinstance(obj, SavedLattice):
        raise ValueError(f"{path} holds a {type(obj).__name__}, not a SamplingLattice")
    return obj
