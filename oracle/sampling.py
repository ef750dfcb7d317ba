"""CPU oracle for tetragono's sweep sampling and local energy.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Restates, on top of ``oracle.tat`` and ``oracle.peps`` (paths relative to the
TNSP repository root):

* the spinless-fermion operators of ``tetragono/tetragono/common_tensor/
  Fermi.py:18-37`` (CP, CM, CC = C0C1 + C1C0, N) and the tV bond Hamiltonian
  ``H = -t CC + V NN`` of ``tetraku/tetraku/models/tV/__init__.py:39-45``
  (``rename_io`` / ``kronecker_product`` of ``common_tensor/tensor_toolkit.py:
  34-47``);
* ``calculate_element`` of ``tetragono/tetragono/tensor_element.py:84-111``:
  the sparse table in-occupations -> {out-occupations: element tensor};
* the sweep order of ``SweepSampling._get_proper_position_order``
  (``sampling_lattice/sampling.py:156-190``) for nearest-neighbour bonds;
* one term of the observer's local energy, ``observer.py:350-379``:
  ``<psi|s'|H|s|psi> / <psi|s|psi>`` as a contraction of one-element
  fermionic tensors.
"""
import numpy as np

from . import peps, tat



def rename_io(t, m):
    """common_tensor/tensor_toolkit.py:34-41."""
    if not isinstance(m, dict):
        m = {i: j for i, j in enumerate(m)}
    res = {}
    for i, j in m.items():
        res[f"I{i}"] = f"I{j}"
        res[f"O{i}"] = f"O{j}"
    return t.edge_rename({k: v for k, v in res.items() if k in t.names})


def fermi_operators(kind_name="FermiU1"):
    """common_tensor/Fermi.py:18-37 (FermiU1) and common_tensor/Parity.py:
    22-38 (FermiZ2), real part: CP = c^dagger, CM = c with the charge carried
    by the one-dimensional ``T`` edge."""
    K = tat.KINDS[kind_name]
    if kind_name == "FermiU1":
        EF = tat.Edge(K, [((0,), 1), ((1,), 1)], False)
        ET = tat.Edge(K, [((0,), 1), ((-1,), 1)], True)
        CP = tat.Tensor(K, ["O0", "I0", "T"], [EF, ET, tat.Edge(K, [((-1,), 1)], False)]).range_(1, 0)
        CM = tat.Tensor(K, ["O0", "I0", "T"], [EF, ET, tat.Edge(K, [((1,), 1)], True)]).range_(1, 0)
    elif kind_name == "FermiZ2":
        EF = tat.Edge(K, [((0,), 1), ((1,), 1)], False)
        ET = tat.Edge(K, [((0,), 1), ((1,), 1)], True)
        CP = tat.Tensor(K, ["O0", "I0", "T"], [EF, ET, tat.Edge(K, [((1,), 1)], False)])
        CP[{"O0": ((1,), 0), "I0": ((0,), 0), "T": ((1,), 0)}] = 1.0
        CM = tat.Tensor(K, ["O0", "I0", "T"], [EF, ET, tat.Edge(K, [((1,), 1)], True)])
        CM[{"O0": ((0,), 0), "I0": ((1,), 0), "T": ((1,), 0)}] = 1.0
    else:
        raise ValueError("fermion operators for FermiU1 and FermiZ2 only")
    C0C1 = rename_io(CP, [0]).contract(rename_io(CM, [1]), {("T", "T")})
    C1C0 = rename_io(CP, [1]).contract(rename_io(CM, [0]), {("T", "T")})
    CC = C0C1 + C1C0
    N = rename_io(CP, [0]).contract(rename_io(CM, [0]), {("T", "T"), ("I0", "O0")})
    return CP, CM, CC, N


def tv_hamiltonian(t=1.0, V=1.0, kind_name="FermiU1"):
    """tetraku/tetraku/models/tV/__init__.py:39-45 (and tV/parity for
    FermiZ2): H = -t CC + V NN with NN = N_0 (x) N_1."""
    _, _, CC, N = fermi_operators(kind_name)
    NN = rename_io(N, [0]).contract(rename_io(N, [1]), set())
    return CC * (-t) + NN * V


def element_pool(H):
    """tensor_element.py:84-111: {in occupations: {out occupations: element
    tensor}}; element tensors have one-element edges (symmetry of the point,
    the operator's arrows).  Occupations are the charges of the points
    (U(1): particle number, Z(2): parity, both the occupation here)."""
    body = H.rank // 2
    result = {}
    for pos_idx in H.block_order:
        syms = [H.edges[i].segments[p][0] for i, p in enumerate(pos_idx)]
        dims = [H.edges[i].segments[p][1] for i, p in enumerate(pos_idx)]
        block = H.block(pos_idx)
        for idx in np.ndindex(*dims):
            value = block[idx]
            if value == 0:
                continue
            edges = [tat.Edge(H.kind, [(syms[i], 1)], H.edges[i].arrow) for i in range(H.rank)]
            tmpl = tat.Tensor(H.kind, list(H.names), edges)
            tmpl.data[:] = value
            point = {H.names[i]: (syms[i], idx[i]) for i in range(H.rank)}
            edge_in = tuple(H.kind.neg(point[f"I{i}"][0])[0] for i in range(body))
            edge_out = tuple(point[f"O{i}"][0][0] for i in range(body))
            result.setdefault(edge_in, {})[edge_out] = tmpl
    return result


def bond_order(L1, L2):
    """sampling.py:156-190 for a Hamiltonian on every nearest-neighbour bond:
    horizontal bonds row-major, then vertical bonds column-major."""
    order = []
    for l1 in range(L1):
        for l2 in range(L2 - 1):
            order.append(((l1, l2), (l1, l2 + 1)))
    for l2 in range(L2):
        for l1 in range(L1 - 1):
            order.append(((l1, l2), (l1 + 1, l2)))
    return order


def local_energy_term(sites, config, cut, positions, H, pool=None, wss_of=None):
    """observer.py:350-379 for one bond: sum over s' of
    ``inv_ws_conj . <s'|H|s> . conj(wss)`` contracted over T and every
    physical name.  ``positions`` = ((l1, l2), (l1', l2')).  ``wss_of(cfg2,
    occupations)`` gives <s'|psi> (default: the amplitude of s' from
    scratch; tetragono's observer uses ``Configuration.replace``)."""
    pool = element_pool(H) if pool is None else pool
    L1, L2 = len(sites), len(sites[0])
    occ = tuple(int(config[p[0]][p[1]]) for p in positions)
    if occ not in pool:
        return 0.0
    _, ws = peps.psi(sites, config, cut)
    w = float(ws.data[0])
    inv_ws_conj = ws * (1.0 / (w * w))
    all_name = {("T", "T")} | {(f"P_{a}_{b}_0", f"P_{a}_{b}_0") for a in range(L1) for b in range(L2)}
    physics_names = [f"P_{p[0]}_{p[1]}_0" for p in positions]
    total = 0.0
    for occ_s, elem in pool[occ].items():
        cfg2 = np.array(config, copy=True)
        for i, p in enumerate(positions):
            cfg2[p[0]][p[1]] = occ_s[i]
        wss = peps.psi(sites, cfg2, cut)[1] if wss_of is None else wss_of(cfg2, occ_s)
        if wss.size == 0 or float(np.max(np.abs(wss.data))) == 0:
            continue
        v = inv_ws_conj.contract(elem, {(physics_names[i], f"I{i}") for i in range(len(positions))})
        v = v.edge_rename({f"O{i}": physics_names[i] for i in range(len(positions))})
        v = v.contract(wss.conjugate(), all_name)
        total += float(v.data[0])
    return total


def local_energy(sites, config, cut, H):
    pool = element_pool(H)
    L1, L2 = len(sites), len(sites[0])
    return sum(local_energy_term(sites, config, cut, pos, H, pool) for pos in bond_order(L1, L2))
