"""CPU oracle for tetragono's sampled-configuration PEPS amplitude.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Restates, on top of ``oracle.tat``:

* site tensor construction: names ``P0`` then ``T, U, D, L, R`` as present
  (tetragono/tetragono/abstract_lattice.py:129-183), partner bonds are the
  ``conjugate()`` of the set side (abstract_lattice.py:214-242), the ``T``
  edge ``[(-N, 1)]`` at (0, 0) (abstract_state.py:411-421);
* the tV charge layout of virtual bonds
  (tetraku/tetraku/models/tV/__init__.py:49-77 for FermiU1,
  tetraku/tetraku/models/tV/parity/__init__.py:31-67 for FermiZ2);
* configuration projection ``_shrink_configuration``
  (tetragono/tetragono/sampling_lattice/lattice.py:291-339);
* the boundary-MPS row ``_two_line_to_one_line``
  (tetragono/tetragono/auxiliaries/single_layer_auxiliaries.py:463-498) and the
  final row ``hole(())`` -> ``_inline_right_to_left[0, 0]`` (:254-312, 710-720);
* the Metropolis acceptance probability of ``SweepSampling._single_term``
  (tetragono/tetragono/sampling_lattice/sampling.py:117-143).
"""
import numpy as np

from . import tat


def safe_contract(t1, t2, pairs):
    """tetragono/tetragono/utility.py:340-350."""
    real = {(a, b) for a, b in pairs if a in t1.names and b in t2.names}
    return t1.contract(t2, real)


def safe_rename(t, mapping):
    """tetragono/tetragono/utility.py:353-358."""
    return t.edge_rename({k: v for k, v in mapping.items() if k in t.names})


def one(kind):
    return tat.Tensor.scalar(kind, 1.0)


def two_line_to_one_line(line_1, line_2, cut, udlr="DULR"):
    """single_layer_auxiliaries.py:463-498 (normalize=False)."""
    up, down, left, right = udlr
    left1, left2, right1, right2 = left + "1", left + "2", right + "1", right + "2"
    if len(line_1) != len(line_2):
        raise ValueError("Different Length in Two Line to One Line")
    length = len(line_1)
    dl = [safe_contract(safe_rename(line_1[i], {left: left1, right: right1}),
                        safe_rename(line_2[i], {left: left2, right: right2}), {(down, up)}) for i in range(length)]
    for i in range(length - 1):
        q, r = dl[i].qr("r", {n for n in (right1, right2) if n in dl[i].names}, right, left)
        dl[i] = q
        dl[i + 1] = safe_contract(dl[i + 1], r, {(left1, right1), (left2, right2)})
    for i in reversed(range(1, length)):
        u, s, v = dl[i].svd({left}, right, left, left, right, cut)
        dl[i] = v
        dl[i - 1] = safe_contract(safe_contract(dl[i - 1], u, {(right, left)}), s, {(right, left)})
    return dl


def amplitude(lattice, cut):
    """psi(s) as a one-element tensor: ``hole(())`` with hint ("H", 0).

    ``lattice[l1][l2]`` are the projected (shrunk) site tensors."""
    L1, L2 = len(lattice), len(lattice[0])
    kind = lattice[0][0].kind
    o = one(kind)
    mps = [o] * L2
    for l1 in reversed(range(1, L1)):
        mps = two_line_to_one_line(mps, lattice[l1], cut)
    inline = o  # _inline_right_to_left[0, L2]
    for l2 in reversed(range(L2)):
        tailed = safe_rename(safe_contract(inline, mps[l2], {("L3", "R")}), {"L": "L3"})
        res = safe_contract(tailed, safe_rename(lattice[0][l2], {"L": "L2"}), {("L2", "R"), ("U", "D")})
        res = safe_contract(res, safe_rename(o, {"L": "L1"}), {("L1", "R"), ("U", "D")})
        inline = res
    return inline


def exact_amplitude(lattice):
    """Exact fermionic contraction of the projected lattice in plain
    row-major site order (no truncation, no boundary MPS).  Used to pin
    ``amplitude`` with a large cut: the two agree up to the order of the
    one-dimensional edges, so compare after transposing to one name order."""
    L1, L2 = len(lattice), len(lattice[0])
    acc = None
    for l1 in range(L1):
        for l2 in range(L2):
            t = lattice[l1][l2]
            t = safe_rename(t, {"U": f"V{l1}_{l2}", "D": f"V{l1 + 1}_{l2}", "L": f"H{l1}_{l2}", "R": f"H{l1}_{l2 + 1}"})
            if acc is None:
                acc = t
            else:
                pairs = {(n, n) for n in t.names if n in acc.names}
                # contracted edge names coincide; rename one side to keep them unique
                ren = {n: n + "'" for n, _ in pairs}
                acc = acc.contract(t.edge_rename(ren), {(n, n + "'") for n, _ in pairs})
    return acc


# ----------------------------------------------------------------------------
# lattice layouts
# ----------------------------------------------------------------------------
def site_names_edges(spec, l1, l2):
    """Names and edges of the raw (unprojected) site tensor at (l1, l2).

    FermiU1 (tV/__init__.py:49-77): ``D`` bonds carry ``[(Q-1, d0), (Q, d1),
    (Q+1, d2)]`` with Q = int(N (L1-l1-1)/L1) in column 0 and Q = 0 elsewhere
    (the reference uses one charge-0 segment there; the segment spec of the
    workload asks for the same three offsets on every bond), ``R`` bonds use
    Q = int((N/L1) (L2-l2-1)/L2).  The partner ``U`` / ``L`` edges are the
    conjugates.  FermiZ2 (tV/parity/__init__.py:31-67): ``U`` and ``L`` carry
    ``[(False, De), (True, Do)]``, partners ``D`` / ``R`` are conjugates, the
    total parity is even."""
    kind = tat.KINDS[spec["kind"]]
    L1, L2 = spec["L1"], spec["L2"]
    segs = spec["segments"]
    names, edges = ["P0"], []
    if spec["kind"] == "FermiU1":
        N = spec["N"]
        edges.append(tat.Edge(kind, [(0, 1), (1, 1)], False))

        def d_edge(a, b):  # bond below site (a, b): set on the D side
            Q = int(N * (L1 - a - 1) / L1) if b == 0 else 0
            return tat.Edge(kind, [(Q - 1, segs[0]), (Q, segs[1]), (Q + 1, segs[2])], False)

        def r_edge(a, b):
            tt = N / L1
            Q = int(tt * (L2 - b - 1) / L2)
            return tat.Edge(kind, [(Q - 1, segs[0]), (Q, segs[1]), (Q + 1, segs[2])], False)

        bonds = {}
        if l1 == 0 and l2 == 0:
            bonds["T"] = tat.Edge(kind, [(-N, 1)], False)
        if l1 != 0:
            bonds["U"] = d_edge(l1 - 1, l2).conjugate()
        if l1 != L1 - 1:
            bonds["D"] = d_edge(l1, l2)
        if l2 != 0:
            bonds["L"] = r_edge(l1, l2 - 1).conjugate()
        if l2 != L2 - 1:
            bonds["R"] = r_edge(l1, l2)
    elif spec["kind"] == "FermiZ2":
        edges.append(tat.Edge(kind, [(0, 1), (1, 1)], False))
        z2 = tat.Edge(kind, [(0, segs[0]), (1, segs[1])], False)
        bonds = {}
        if l1 == 0 and l2 == 0:
            bonds["T"] = tat.Edge(kind, [(0, 1)], False)
        if l1 != 0:
            bonds["U"] = z2
        if l1 != L1 - 1:
            bonds["D"] = z2.conjugate()
        if l2 != 0:
            bonds["L"] = z2
        if l2 != L2 - 1:
            bonds["R"] = z2.conjugate()
    else:
        raise ValueError("unsupported lattice kind")
    for n in ("T", "U", "D", "L", "R"):
        if n in bonds:
            names.append(n)
            edges.append(bonds[n])
    return kind, names, edges


def build_sites(spec, storages):
    """Raw site tensors with the given storages (row-major site order)."""
    L1, L2 = spec["L1"], spec["L2"]
    out = []
    for l1 in range(L1):
        row = []
        for l2 in range(L2):
            kind, names, edges = site_names_edges(spec, l1, l2)
            t = tat.Tensor(kind, names, edges)
            t.data[:] = storages[l1 * L2 + l2]
            row.append(t)
        out.append(row)
    return out


def shrink(site, l1, l2, occupation):
    """lattice.py:291-339: contract the physical leg with a one-hot shrinker
    and name the remaining dimension-one leg ``P_{l1}_{l2}_0``."""
    kind = site.kind
    phys = site.edge_by_name("P0")
    q = kind.make(occupation)
    sh = tat.Tensor(kind, ["P", "Q"], [tat.Edge(kind, [(q, 1)], False), phys.conjugate()])
    sh[{"Q": (kind.neg(q), 0), "P": (q, 0)}] = 1.0
    return site.contract(sh.edge_rename({"P": f"P_{l1}_{l2}_0"}), {("P0", "Q")})


def project(sites, config):
    L1, L2 = len(sites), len(sites[0])
    return [[shrink(sites[l1][l2], l1, l2, int(config[l1][l2])) for l2 in range(L2)] for l1 in range(L1)]


def psi(sites, config, cut):
    """psi(s) value read from the one-element result tensor."""
    t = amplitude(project(sites, config), cut)
    return float(t.data[0]) if t.size else 0.0, t


def metropolis_probability(psi_new, psi_old, n_hops_old, n_hops_new, alpha=1.0):
    """sampling.py:136-137: p = |psi'/psi|^(2 alpha) * n / n'."""
    return abs(psi_new / psi_old) ** (2 * alpha) * n_hops_old / n_hops_new
