"""Sampled-configuration single-layer PEPS amplitude on the GPU.

The lattice follows tetragono's construction (paths relative to the TNSP
repository root):

* site tensor names ``P0`` then ``T, U, D, L, R`` as present
  (tetragono/tetragono/abstract_lattice.py:129-183), partner bonds conjugated
  (abstract_lattice.py:214-242), ``T = [(-N, 1)]`` on site (0, 0)
  (abstract_state.py:411-421);
* FermiU1 bond charges as the tV model lays them out
  (tetraku/tetraku/models/tV/__init__.py:49-77): three segments
  ``{Q-1, Q, Q+1}`` around a background Q that counts the particles
  downstream of the bond (Q = 0 on vertical bonds off column 0);
* FermiZ2 bonds ``{even, odd}`` (tetraku/tetraku/models/tV/parity/__init__.py).

``amplitude_plain`` evaluates psi(s) for ONE configuration through the
drop-in ``TAT`` API, op by op, in exactly tetragono's order
(``_shrink_configuration`` sampling_lattice/lattice.py:319-339,
``_two_line_to_one_line`` auxiliaries/single_layer_auxiliaries.py:463-498,
``hole(())`` :710-720): this is the plain per-configuration GPU path.  The
batched engine (``tnsp_amd.engine``) runs the same op sequence for thousands
of configurations in lockstep.
"""
import dataclasses

import numpy as np

from . import TAT
from .shape import KINDS, Edge


@dataclasses.dataclass(frozen=True)
class LatticeSpec:
    kind: str  # "FermiU1" or "FermiZ2"
    L1: int
    L2: int
    segments: tuple  # FermiU1: (d_{Q-1}, d_Q, d_{Q+1}); FermiZ2: (d_even, d_odd)
    N: int  # particle number (FermiU1 total charge); FermiZ2 uses its parity
    cut: int  # boundary-MPS bond dimension Dc
    seed: int = 0

    @property
    def D(self):
        return int(sum(self.segments))


# the workloads named by the benchmark configurations
C2 = LatticeSpec("FermiU1", 4, 4, (2, 1, 1), 8, 16, seed=2)
C3 = LatticeSpec("FermiZ2", 8, 8, (3, 3), 32, 36, seed=3)
C4 = LatticeSpec("FermiU1", 8, 8, (3, 3, 2), 32, 64, seed=4)


def site_layout(spec, l1, l2):
    """(names, edges) of the raw site tensor at (l1, l2)."""
    kind = KINDS[spec.kind]
    L1, L2, segs = spec.L1, spec.L2, spec.segments
    bonds = {}
    if spec.kind == "FermiU1":
        N = spec.N
        phys = Edge(kind, [(0, 1), (1, 1)], False)

        def d_bond(a, b):
            Q = int(N * (L1 - a - 1) / L1) if b == 0 else 0
            return Edge(kind, [(Q - 1, segs[0]), (Q, segs[1]), (Q + 1, segs[2])], False)

        def r_bond(a, b):
            Q = int((N / L1) * (L2 - b - 1) / L2)
            return Edge(kind, [(Q - 1, segs[0]), (Q, segs[1]), (Q + 1, segs[2])], False)

        if l1 == 0 and l2 == 0:
            bonds["T"] = Edge(kind, [(-N, 1)], False)
        if l1 != 0:
            bonds["U"] = d_bond(l1 - 1, l2).conjugate()
        if l1 != L1 - 1:
            bonds["D"] = d_bond(l1, l2)
        if l2 != 0:
            bonds["L"] = r_bond(l1, l2 - 1).conjugate()
        if l2 != L2 - 1:
            bonds["R"] = r_bond(l1, l2)
    elif spec.kind == "FermiZ2":
        phys = Edge(kind, [(0, 1), (1, 1)], False)
        z2 = Edge(kind, [(0, segs[0]), (1, segs[1])], False)
        if l1 == 0 and l2 == 0:
            bonds["T"] = Edge(kind, [(spec.N % 2, 1)], False)
        if l1 != 0:
            bonds["U"] = z2
        if l1 != L1 - 1:
            bonds["D"] = z2.conjugate()
        if l2 != 0:
            bonds["L"] = z2
        if l2 != L2 - 1:
            bonds["R"] = z2.conjugate()
    else:
        raise ValueError("lattice kind must be FermiU1 or FermiZ2")
    names, edges = ["P0"], [phys]
    for n in ("T", "U", "D", "L", "R"):
        if n in bonds:
            names.append(n)
            edges.append(bonds[n])
    return names, edges


def site_sizes(spec):
    from .shape import Shape

    out = []
    for l1 in range(spec.L1):
        for l2 in range(spec.L2):
            names, edges = site_layout(spec, l1, l2)
            out.append(Shape.create(KINDS[spec.kind], names, edges).size)
    return out


def random_storages(spec):
    """Site storages, blocks i.i.d. N(0, 1)/sqrt(D), filled in TAT storage
    order site by site (row-major) from numpy PCG64(spec.seed)."""
    gen = np.random.Generator(np.random.PCG64(spec.seed))
    return [gen.standard_normal(n) / np.sqrt(spec.D) for n in site_sizes(spec)]


def sample_configurations(spec, count, seed, particles=None):
    """Occupation maps (count, L1, L2) of int8.  FermiU1: exactly N particles
    placed uniformly.  FermiZ2: particle number drawn uniformly from the even
    values of ``particles`` (a (lo, hi) range), matching the even total
    parity of the lattice."""
    gen = np.random.Generator(np.random.PCG64(seed))
    L = spec.L1 * spec.L2
    out = np.zeros((count, L), dtype=np.int8)
    for c in range(count):
        if spec.kind == "FermiU1":
            n = spec.N
        else:
            lo, hi = particles or (spec.N, spec.N)
            choices = [x for x in range(lo, hi + 1) if x % 2 == spec.N % 2]
            n = int(gen.choice(choices))
        out[c, gen.choice(L, n, replace=False)] = 1
    return out.reshape(count, spec.L1, spec.L2)


class Lattice:
    """Device-resident PEPS site tensors (drop-in ``TAT`` tensors)."""

    def __init__(self, spec, storages=None, dtype="float64"):
        self.spec = spec
        self.dtype = dtype
        storages = storages if storages is not None else random_storages(spec)
        mod = getattr(TAT, spec.kind)
        smod = mod.D if dtype == "float64" else mod.S
        self.sites = []
        for l1 in range(spec.L1):
            row = []
            for l2 in range(spec.L2):
                names, edges = site_layout(spec, l1, l2)
                t = smod.Tensor(names, edges)
                t.storage = storages[l1 * spec.L2 + l2]
                row.append(t)
            self.sites.append(row)
        self._projected = {}

    @classmethod
    def from_saved(cls, saved, cut, dtype=None):
        """The lattice of a saved tetragono ``SamplingLattice``
        (``latticeio.read_lattice``) on the GPU: its site tensors as they
        are, one physical orbit per site (edge ``P0``), a FermiU1 or FermiZ2
        symmetry.  ``cut`` is the boundary-MPS bond dimension (tetragono
        keeps it outside the lattice); the particle number comes from the
        ``T`` edge of site (0, 0).  ``spec.segments`` is empty: the bonds are
        whatever the file holds."""
        if saved.kind not in ("FermiU1", "FermiZ2"):
            raise ValueError(f"the engine takes FermiU1 or FermiZ2 lattices, not {saved.kind}")
        s00 = saved.site(0, 0)
        if "T" not in s00.names:
            raise ValueError("site (0, 0) has no T edge")
        for l1 in range(saved.L1):
            for l2 in range(saved.L2):
                names = saved.site(l1, l2).names
                if "P0" not in names or any(n.startswith("P") and n != "P0" for n in names):
                    raise ValueError(f"site ({l1}, {l2}) must have exactly one physical edge P0, has {names}")
        q = s00.edges[s00.names.index("T")].segments[0][0][0]
        N = -q if saved.kind == "FermiU1" else q
        self = cls.__new__(cls)
        self.spec = LatticeSpec(saved.kind, saved.L1, saved.L2, (), int(N), int(cut))
        self.dtype = dtype or saved.dtype
        self.sites = []
        for l1 in range(saved.L1):
            row = []
            for l2 in range(saved.L2):
                t = saved.site(l1, l2).to_tensor()
                if t.dtype != self.dtype:
                    t = t.to(self.dtype)
                row.append(t)
            self.sites.append(row)
        self._projected = {}
        return self

    def projected(self, l1, l2, occupation):
        """``_shrink_configuration`` (lattice.py:291-339): contract P0 with a
        one-hot shrinker; the remaining dimension-one leg is P_{l1}_{l2}_0.
        Cached: a projected site depends only on the site and occupation."""
        key = (l1, l2, int(occupation))
        t = self._projected.get(key)
        if t is None:
            sh = self.shrinker(l1, l2, occupation)
            t = self.sites[l1][l2].contract(sh.edge_rename({"P": f"P_{l1}_{l2}_0"}), {("P0", "Q")})
            self._projected[key] = t
        return t

    def apply_gradient(self, gradient, step_size):
        """``SamplingLattice.apply_gradient`` (lattice.py:936-948): every site
        tensor minus ``step_size`` times its gradient, in place (the same
        device buffers, so an engine built on this lattice only needs
        ``refresh_sites``).  Projected sites are recomputed on demand."""
        for l1 in range(self.spec.L1):
            for l2 in range(self.spec.L2):
                site = self.sites[l1][l2]
                g = gradient[l1][l2]
                if g.names != site.names:
                    g = g.transpose(site.names)
                if g._data.numel() != site._data.numel():
                    raise ValueError("gradient and site differ in shape")
                site._data.sub_(g._data.to(site._data.dtype) * float(step_size))
        self._projected = {k: v for k, v in self._projected.items() if k[0] == "shrinker"}

    def shrinker(self, l1, l2, occupation):
        """``_get_shrinker`` (lattice.py:291-318): ``P`` the dimension-one
        edge of the occupation, ``Q`` the conjugate physical edge, one unit
        entry."""
        key = ("shrinker", l1, l2, int(occupation))
        sh = self._projected.get(key)
        if sh is None:
            site = self.sites[l1][l2]
            kind = KINDS[self.spec.kind]
            mod = getattr(TAT, self.spec.kind)
            smod = mod.D if self.dtype == "float64" else mod.S
            phys = site.edge_by_name("P0")
            q = kind.make(occupation)
            sh = smod.Tensor(["P", "Q"], [Edge(kind, [(q, 1)], False), phys.conjugate()]).zero_()
            sh[{"Q": (kind.neg(q), 0), "P": (q, 0)}] = 1.0
            self._projected[key] = sh
        return sh


def _safe_contract(t1, t2, pairs):
    """tetragono/tetragono/utility.py:340-350."""
    real = {(a, b) for a, b in pairs if a in t1.names and b in t2.names}
    return t1.contract(t2, real)


def _safe_rename(t, mapping):
    return t.edge_rename({k: v for k, v in mapping.items() if k in t.names})


def _one(spec, dtype):
    mod = getattr(TAT, spec.kind)
    return (mod.D if dtype == "float64" else mod.S).Tensor(1.0)


def two_line_to_one_line(line_1, line_2, cut):
    """single_layer_auxiliaries.py:463-498 with udlr = "DULR"."""
    up, down, left, right = "DULR"
    l1n, l2n, r1n, r2n = "L1", "L2", "R1", "R2"
    n = len(line_1)
    dl = [_safe_contract(_safe_rename(line_1[i], {left: l1n, right: r1n}),
                         _safe_rename(line_2[i], {left: l2n, right: r2n}), {(down, up)}) for i in range(n)]
    for i in range(n - 1):
        q, r = dl[i].qr("r", {x for x in (r1n, r2n) if x in dl[i].names}, right, left)
        dl[i] = q
        dl[i + 1] = _safe_contract(dl[i + 1], r, {(l1n, r1n), (l2n, r2n)})
    for i in reversed(range(1, n)):
        u, s, v = dl[i].svd({left}, right, left, left, right, cut)
        dl[i] = v
        dl[i - 1] = _safe_contract(_safe_contract(dl[i - 1], u, {(right, left)}), s, {(right, left)})
    return dl


def amplitude_plain(lattice, config):
    """psi(s) for one configuration; returns (value, one-element tensor)."""
    spec = lattice.spec
    L1, L2 = spec.L1, spec.L2
    proj = [[lattice.projected(l1, l2, config[l1][l2]) for l2 in range(L2)] for l1 in range(L1)]
    one = _one(spec, lattice.dtype)
    mps = [one] * L2
    for l1 in reversed(range(1, L1)):
        mps = two_line_to_one_line(mps, proj[l1], spec.cut)
    inline = one
    for l2 in reversed(range(L2)):
        tailed = _safe_rename(_safe_contract(inline, mps[l2], {("L3", "R")}), {"L": "L3"})
        res = _safe_contract(tailed, _safe_rename(proj[0][l2], {"L": "L2"}), {("L2", "R"), ("U", "D")})
        res = _safe_contract(res, _safe_rename(one, {"L": "L1"}), {("L1", "R"), ("U", "D")})
        inline = res
    value = float(np.asarray(inline.storage)[0]) if np.asarray(inline.storage).size else 0.0
    return value, inline
