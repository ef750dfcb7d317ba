"""Device-resident block-sparse Tensor with PyTAT's Python API.

Mirrors ``TAT.<Sym>.<D|S>.Tensor`` of the reference binding
(PyTAT/PyTAT.hpp:570-1147).  Metadata (names, edges, block table) lives on
the host in a native ``tns_shape``; the storage is one contiguous torch
buffer on the GPU holding every allowed block in TAT's order
(TAT/include/TAT/structure/core.hpp:152-191).  All structural operations
(transpose / merge / split / reverse, contract, svd, qr, conjugate) run as
planned task lists on the hand-written HIP kernels of libtnsgpu.so; torch is
used only for buffer allocation and trivial elementwise arithmetic.
"""
import numpy as np

from . import plan as P
from .shape import Edge, Shape


def _torch():
    import torch

    return torch


_DEVICE = None


def device():
    global _DEVICE
    if _DEVICE is None:
        torch = _torch()
        if not torch.cuda.is_available():
            raise RuntimeError("tnsp_amd tensors live on the GPU and no HIP device is visible")
        _DEVICE = torch.device("cuda", torch.cuda.current_device())
    return _DEVICE


def _stream():
    return _torch().cuda.current_stream().cuda_stream


def _tdtype(dtype):
    torch = _torch()
    return torch.float64 if dtype == "float64" else torch.float32


def _workspace(nbytes):
    torch = _torch()
    if nbytes <= 0:
        return None, 0
    buf = torch.empty(int(nbytes), dtype=torch.uint8, device=device())
    return buf, buf.data_ptr()


class Storage:
    """Write-through handle on a tensor's device storage (PyTAT's ``storage``
    is a numpy view; here reads copy to the host and writes upload)."""

    __slots__ = ("_t",)

    def __init__(self, tensor):
        self._t = tensor

    def numpy(self):
        return self._t._data.detach().cpu().numpy()

    def __array__(self, dtype=None, copy=None):
        a = self.numpy()
        return a.astype(dtype) if dtype is not None else a

    def __len__(self):
        return self._t._data.numel()

    @property
    def size(self):
        return self._t._data.numel()

    @property
    def shape(self):
        return (self.size,)

    def __getitem__(self, idx):
        return self.numpy()[idx]

    def __setitem__(self, idx, value):
        torch = _torch()
        host = self.numpy()
        host[idx] = value
        self._t._data.copy_(torch.from_numpy(host))

    def __iter__(self):
        return iter(self.numpy())

    def __eq__(self, other):
        return self.numpy() == np.asarray(other)

    def __repr__(self):
        return repr(self.numpy())

    def tolist(self):
        return self.numpy().tolist()


class _Blocks:
    __slots__ = ("_t",)

    def __init__(self, t):
        self._t = t

    def _locate(self, key):
        t = self._t
        items = list(key)
        given = [n for n, _ in items]
        mapping = dict(items)
        pos = []
        for n, e in zip(t.names, t.edges):
            pos.append(e.position_by_symmetry(mapping[n]))
        off, shp = t._shape.block_of(pos)
        perm = [t.names.index(n) for n in given]
        return off, shp, perm

    def __getitem__(self, key):
        off, shp, perm = self._locate(key)
        n = int(np.prod(shp, dtype=np.int64))
        blk = self._t._data[off:off + n].detach().cpu().numpy().reshape(shp)
        return blk.transpose(perm) if perm else blk

    def __setitem__(self, key, value):
        torch = _torch()
        off, shp, perm = self._locate(key)
        n = int(np.prod(shp, dtype=np.int64))
        given_shape = [shp[p] for p in perm] if perm else list(shp)
        val = np.broadcast_to(np.asarray(value, dtype=self._t._np_dtype()), given_shape)
        if perm:
            val = val.transpose(np.argsort(perm))
        self._t._data[off:off + n] = torch.from_numpy(np.ascontiguousarray(val).reshape(-1)).to(self._t._data.device)


class Tensor:
    """Block-sparse tensor on the GPU (see module docstring)."""

    __slots__ = ("_shape", "_data", "model", "dtype", "__weakref__")

    # ------------------------------------------------------------ creation
    def __init__(self, model, dtype, names=None, edges=None, *, _shape=None, _data=None):
        self.model = model
        self.dtype = dtype
        if _shape is not None:
            self._shape = _shape
            self._data = _data
            return
        torch = _torch()
        self._shape = Shape.create(model.kind, list(names or []), list(edges or []))
        self._data = torch.zeros(self._shape.size, dtype=_tdtype(dtype), device=device())

    @classmethod
    def _wrap(cls, model, dtype, shape, data):
        return cls(model, dtype, _shape=shape, _data=data)

    def _new(self, shape, data=None):
        torch = _torch()
        if data is None:
            data = torch.empty(shape.size, dtype=self._data.dtype, device=self._data.device)
        return Tensor._wrap(self.model, self.dtype, shape, data)

    def _np_dtype(self):
        return np.float64 if self.dtype == "float64" else np.float32

    # ------------------------------------------------------------ metadata
    @property
    def names(self):
        return list(self._shape.names)

    @property
    def edges(self):
        return list(self._shape.edges)

    @property
    def rank(self):
        return len(self._shape.names)

    def edge_by_name(self, name):
        return self._shape.edges[self._shape.names.index(name)]

    @property
    def storage(self):
        return Storage(self)

    @storage.setter
    def storage(self, value):
        torch = _torch()
        arr = np.asarray(value, dtype=self._np_dtype()).reshape(-1)
        if arr.size != self._data.numel():
            raise ValueError("storage size mismatch")
        self._data.copy_(torch.from_numpy(arr))

    @property
    def blocks(self):
        return _Blocks(self)

    const_blocks = blocks

    def _coords(self, mapping):
        pos, offs = [], []
        for n, e in zip(self.names, self.edges):
            v = mapping[n]
            if isinstance(v, tuple):
                q, off = v
                p = e.position_by_symmetry(q)
            else:
                p, off = e.coord_by_index(int(v))
            pos.append(p)
            offs.append(off)
        boff, shp = self._shape.block_of(pos)
        lin = 0
        for o, d in zip(offs, shp):
            lin = lin * d + o
        return boff + lin

    def __getitem__(self, mapping):
        return float(self._data[self._coords(mapping)].item())

    def __setitem__(self, mapping, value):
        self._data[self._coords(mapping)] = float(value)

    # ------------------------------------------------------------ fills
    def zero_(self):
        self._data.zero_()
        return self

    def range_(self, first=0, step=1):
        torch = _torch()
        n = self._data.numel()
        self._data.copy_(first + step * torch.arange(n, dtype=self._data.dtype, device=self._data.device))
        return self

    def randn_(self, mean=0.0, stddev=1.0):
        from . import rng

        torch = _torch()
        self._data.copy_(torch.from_numpy(rng.normal(self._data.numel(), mean, stddev, self._np_dtype())))
        return self

    def rand_(self, min=0.0, max=1.0):
        from . import rng

        torch = _torch()
        self._data.copy_(torch.from_numpy(rng.uniform(self._data.numel(), min, max, self._np_dtype())))
        return self

    def set_(self, function):
        torch = _torch()
        vals = np.array([function() for _ in range(self._data.numel())], dtype=self._np_dtype())
        self._data.copy_(torch.from_numpy(vals))
        return self

    def copy(self):
        return self._new(self._shape, self._data.clone())

    __copy__ = copy

    def __deepcopy__(self, memo):
        return self.copy()

    def same_shape(self):
        return self._new(self._shape)

    def to(self, new_type):
        s = str(new_type)
        dtype = "float32" if ("32" in s or s in ("S",)) else "float64"
        mod = self.model.S if dtype == "float32" else self.model.D
        return Tensor._wrap(self.model, dtype, self._shape, self._data.to(_tdtype(dtype)).clone()) if mod else None

    # ------------------------------------------------------------ norms
    def norm_max(self):
        return float(self._data.abs().max().item()) if self._data.numel() else 0.0

    def norm_2(self):
        return float(self._data.double().square().sum().sqrt().item())

    def norm_sum(self):
        return float(self._data.abs().sum().item())

    def norm_num(self):
        return float(self._data.numel())

    def __float__(self):
        if self._data.numel() != 1:
            raise RuntimeError("not a one-element tensor")
        return float(self._data.item())

    # ------------------------------------------------------------ arithmetic
    def _zip(self, other, op, inplace=False):
        if isinstance(other, Tensor):
            if other.names != self.names and other.rank == self.rank:
                other = other.transpose(self.names)
            if other._data.numel() != self._data.numel():
                raise RuntimeError("elementwise operation on tensors of different shape")
            rhs = other._data.to(self._data.dtype)
        else:
            rhs = float(other)
        if inplace:
            op(self._data, rhs, out=self._data)
            return self
        res = self._new(self._shape, op(self._data, rhs))
        return res

    def __add__(self, o):
        return self._zip(o, _torch().add)

    def __sub__(self, o):
        return self._zip(o, _torch().sub)

    def __mul__(self, o):
        return self._zip(o, _torch().mul)

    def __truediv__(self, o):
        return self._zip(o, _torch().div)

    def __radd__(self, o):
        return self._zip(o, _torch().add)

    def __rmul__(self, o):
        return self._zip(o, _torch().mul)

    def __rsub__(self, o):
        return self._new(self._shape, float(o) - self._data)

    def __rtruediv__(self, o):
        return self._new(self._shape, float(o) / self._data)

    def __neg__(self):
        return self._new(self._shape, -self._data)

    def __iadd__(self, o):
        return self._zip(o, _torch().add, True)

    def __isub__(self, o):
        return self._zip(o, _torch().sub, True)

    def __imul__(self, o):
        return self._zip(o, _torch().mul, True)

    def __itruediv__(self, o):
        return self._zip(o, _torch().div, True)

    # ------------------------------------------------------------ structure ops
    def _run1(self, plan):
        """Run a one-input one-output permute plan."""
        out = self._new(plan.output(0))
        ws, wp = _workspace(plan.workspace(self.dtype))
        plan.run(self.dtype, [self._data.data_ptr()], [out._data.data_ptr()], wp, _stream())
        return out

    def edge_rename(self, dictionary):
        """New names on the SAME storage (tensor.hpp:760-784)."""
        return Tensor._wrap(self.model, self.dtype, self._shape.rename(dict(dictionary)), self._data)

    def transpose(self, target_names):
        target_names = list(target_names)
        if target_names == self.names:
            return Tensor._wrap(self.model, self.dtype, self._shape, self._data)
        return self._run1(P.edge_operator(self._shape, new_names=target_names))

    def reverse_edge(self, reversed_names, apply_parity=False, parity_exclude_names=()):
        return self._run1(P.edge_operator(self._shape, reversed_names=list(reversed_names), new_names=self.names,
                                          apply_parity=apply_parity, ex_rev_before=list(parity_exclude_names)))

    def merge_edge(self, merge, apply_parity=False, parity_exclude_names_merge=(), parity_exclude_names_reverse=()):
        """Merged edge lands where the last member of its group was
        (TAT/include/TAT/implement/split_and_merge.hpp:55-116)."""
        target = []
        for name in reversed(self.names):
            found = False
            for new, olds in merge.items():
                if name in olds:
                    if name == olds[-1]:
                        target.append(new)
                    found = True
                    break
            if not found:
                target.append(name)
        for new, olds in merge.items():
            if len(olds) == 0:
                target.append(new)
        target.reverse()
        return self._run1(P.edge_operator(self._shape, merge_map={k: list(v) for k, v in merge.items()},
                                          new_names=target, apply_parity=apply_parity,
                                          ex_merge=list(parity_exclude_names_merge),
                                          ex_rev_after=list(parity_exclude_names_reverse)))

    def split_edge(self, split, apply_parity=False, parity_exclude_names_split=()):
        target = []
        for name in self.names:
            if name in split:
                target.extend(n for n, _ in split[name])
            else:
                target.append(name)
        return self._run1(P.edge_operator(self._shape, split_map={k: list(v) for k, v in split.items()},
                                          new_names=target, apply_parity=apply_parity,
                                          ex_split=list(parity_exclude_names_split)))

    def edge_operator(self, split_map, reversed_names, merge_map, new_names, apply_parity=False,
                      parity_exclude_names_split=(), parity_exclude_names_reverse_before_transpose=(),
                      parity_exclude_names_reverse_after_transpose=(), parity_exclude_names_merge=()):
        return self._run1(P.edge_operator(
            self._shape, split_map={k: list(v) for k, v in split_map.items()}, reversed_names=list(reversed_names),
            merge_map={k: list(v) for k, v in merge_map.items()}, new_names=list(new_names),
            apply_parity=apply_parity, ex_split=list(parity_exclude_names_split),
            ex_rev_before=list(parity_exclude_names_reverse_before_transpose),
            ex_rev_after=list(parity_exclude_names_reverse_after_transpose),
            ex_merge=list(parity_exclude_names_merge)))

    def contract(self, another_tensor, contract_pairs, fuse_names=()):
        if fuse_names:
            raise NotImplementedError("fuse_names is outside the supported path")
        other = another_tensor
        if other.dtype != self.dtype:
            other = other.to(self.dtype)
        plan = P.contract(self._shape, other._shape, list(contract_pairs))
        out = self._new(plan.output(0))
        ws, wp = _workspace(plan.workspace(self.dtype))
        plan.run(self.dtype, [self._data.data_ptr(), other._data.data_ptr()], [out._data.data_ptr()], wp, _stream())
        return out

    def svd(self, free_names_u, common_name_u, common_name_v, singular_name_u, singular_name_v, cut=-1):
        plan = P.svd(self._shape, list(free_names_u), common_name_u, common_name_v, singular_name_u, singular_name_v)
        ws, wp = _workspace(plan.workspace(self.dtype))
        remain, relative = P.cut_params(cut)
        plan.svd_factor(self.dtype, self._data.data_ptr(), wp, remain, relative, _stream())
        outs = [self._new(plan.output(i)) for i in range(3)]
        plan.run(self.dtype, [self._data.data_ptr()], [o._data.data_ptr() for o in outs], wp, _stream())
        return tuple(outs)

    def qr(self, free_names_direction, free_names, common_name_q, common_name_r):
        plan = P.qr(self._shape, free_names_direction, list(free_names), common_name_q, common_name_r)
        outs = [self._new(plan.output(i)) for i in range(2)]
        ws, wp = _workspace(plan.workspace(self.dtype))
        plan.run(self.dtype, [self._data.data_ptr()], [o._data.data_ptr() for o in outs], wp, _stream())
        return tuple(outs)

    # ------------------------------------------------------------ binary I/O
    def dump(self):
        """PyTAT ``dump`` (PyTAT/PyTAT.hpp:757-760): TAT's binary format
        (TAT/include/TAT/miscellaneous/io.hpp:687-705) as bytes."""
        from . import tatio

        return tatio.serialize(self.model.kind, self.names, self.edges, np.asarray(self.storage))

    def load(self, data):
        """PyTAT ``load`` (PyTAT/PyTAT.hpp:761-765): this tensor becomes the
        one stored in ``data`` (binary version 1, or version 0 whose blocks
        are reordered as io.hpp:742-749 does); returns self."""
        from . import tatio

        names, edges, storage, version = tatio.parse(self.model.kind, data, self.dtype)
        t = Tensor(self.model, self.dtype, names, edges)
        if storage.size != t._shape.size:
            raise RuntimeError("TAT tensor bytes: storage size does not match the edges")
        if version == 0:
            blocks = [(tuple(edges[i].segments[p][0] for i, p in enumerate(pos)), int(np.prod(shape)))
                      for pos, _, shape in t._shape.blocks()]
            storage = tatio.v0_to_v1(blocks, storage)
        t.storage = storage
        self._shape, self._data = t._shape, t._data
        return self

    def __reduce__(self):
        # PyTAT pickles through dump / load (PyTAT/PyTAT.hpp:768-771)
        return (_from_bytes, (self.model.kind.name, self.dtype, self.dump()))

    def conjugate(self, trivial_metric=False):
        if self.model.kind.length == 0:
            return self.copy()
        return self._run1(P.conjugate(self._shape, trivial_metric))

    # ------------------------------------------------------------ display
    def __repr__(self):
        return f"{self.model.name}.{self.dtype}.Tensor(names={self.names}, edges={self.edges})"

    __str__ = __repr__


def _from_bytes(kind_name, dtype, data):
    from . import TAT

    sym = getattr(TAT, kind_name)
    return (sym.D if dtype == "float64" else sym.S).Tensor([], []).load(data)
