"""CPU oracle for tetragono's cached-environment ``replace`` and the sweep
sampling that drives it.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Restates, on top of ``oracle.tat`` and ``oracle.peps`` (paths relative to the
TNSP repository root, file ``tetragono/tetragono/auxiliaries/
single_layer_auxiliaries.py`` unless named):

* the four boundary MPS ``_up_to_down`` ("UDLR"), ``_down_to_up`` ("DULR"),
  ``_left_to_right`` ("LRUD"), ``_right_to_left`` ("RLUD") (:402-432), each a
  chain of ``_two_line_to_one_line`` (:463-498) from the ones row / column;
* the inline environments within a row, ``_inline_left_to_right[_tailed]``
  and ``_inline_right_to_left[_tailed]`` (:254-312), and within a column,
  ``_inline_up_to_down[_tailed]`` and ``_inline_down_to_up[_tailed]``
  (:314-372);
* ``replace`` for two neighbouring sites (:593-708), horizontal pair
  (:593-603, 641-651) and vertical pair (:652-687), the branches a sweep
  reaches: the four-row variants read environments that nothing in a sweep
  ever evaluates;
* ``SweepSampling._single_term`` / ``__call__`` (sampling_lattice/
  sampling.py:117-154) for the tV bond Hamiltonian, including its diagonal
  target on a doubly occupied bond: ``Configuration.replace`` with no site
  changed (sampling_lattice/lattice.py:231-267) returns ``hole(())``, the
  amplitude of the current configuration (:500-510, 710-720).

Environments are recomputed from the current lattice at each use: the lazy
graph of the reference caches exactly these values and invalidates them when
a site they depend on changes, so a from-scratch value is what every cache
hit returns.
"""
import numpy as np

from .peps import one, project, safe_contract, safe_rename, two_line_to_one_line


class Environments:
    """Environments of one projected lattice (``lat[l1][l2]``: shrunk site
    tensors) at boundary-MPS bond dimension ``cut``."""

    def __init__(self, lat, cut):
        self.lat = [list(row) for row in lat]
        self.L1, self.L2 = len(lat), len(lat[0])
        self.cut = cut
        self.one = one(lat[0][0].kind)

    # -- boundary MPS (:402-432) ------------------------------------------------
    def up_to_down(self, l1):
        """_up_to_down[l1]: rows 0..l1 absorbed from the top; -1 is ones."""
        mps = [self.one] * self.L2
        for r in range(l1 + 1):
            mps = two_line_to_one_line(mps, self.lat[r], self.cut, "UDLR")
        return mps

    def down_to_up(self, l1):
        mps = [self.one] * self.L2
        for r in reversed(range(l1, self.L1)):
            mps = two_line_to_one_line(mps, self.lat[r], self.cut, "DULR")
        return mps

    def column(self, l2):
        return [self.lat[l1][l2] for l1 in range(self.L1)]

    def left_to_right(self, l2):
        mps = [self.one] * self.L1
        for c in range(l2 + 1):
            mps = two_line_to_one_line(mps, self.column(c), self.cut, "LRUD")
        return mps

    def right_to_left(self, l2):
        mps = [self.one] * self.L1
        for c in reversed(range(l2, self.L2)):
            mps = two_line_to_one_line(mps, self.column(c), self.cut, "RLUD")
        return mps

    # -- row inline environments (:254-312) --------------------------------------
    def inline_left_tailed(self, l1, l2, up, down):
        """_inline_left_to_right_tailed[l1, l2]; ``up`` = _up_to_down[l1-1],
        ``down`` = _down_to_up[l1+1]."""
        tailed = None
        for c in range(-1, l2 + 1):
            if c == -1:
                left = self.one
            else:
                left = safe_contract(tailed, safe_rename(self.lat[l1][c], {"R": "R2"}), {("R2", "L"), ("D", "U")})
                left = safe_contract(left, safe_rename(down[c], {"R": "R3"}), {("R3", "L"), ("D", "U")})
            if c == self.L2 - 1:
                tailed = left
            else:
                tailed = safe_rename(safe_contract(left, up[c + 1], {("R1", "L")}), {"R": "R1"})
        return tailed

    def inline_right(self, l1, l2, up, down, tailed_result=True):
        """_inline_right_to_left_tailed[l1, l2] (or the untailed
        _inline_right_to_left[l1, l2] with ``tailed_result=False``)."""
        tailed = None
        for c in reversed(range(l2, self.L2 + 1)):
            if c == self.L2:
                right = self.one
            else:
                right = safe_contract(tailed, safe_rename(self.lat[l1][c], {"L": "L2"}), {("L2", "R"), ("U", "D")})
                right = safe_contract(right, safe_rename(up[c], {"L": "L1"}), {("L1", "R"), ("U", "D")})
            if c == l2 and not tailed_result:
                return right
            if c == 0:
                tailed = right
            else:
                tailed = safe_rename(safe_contract(right, down[c - 1], {("L3", "R")}), {"L": "L3"})
        return tailed

    # -- column inline environments (:314-372) -----------------------------------
    def inline_up_tailed(self, l1, l2, left, right):
        """_inline_up_to_down_tailed[l1, l2]; ``left`` = _left_to_right[l2-1],
        ``right`` = _right_to_left[l2+1]."""
        tailed = None
        for r in range(-1, l1 + 1):
            if r == -1:
                inl = self.one
            else:
                inl = safe_contract(tailed, safe_rename(self.lat[r][l2], {"D": "D2"}), {("D2", "U"), ("R", "L")})
                inl = safe_contract(inl, safe_rename(right[r], {"D": "D3"}), {("D3", "U"), ("R", "L")})
            # r == L1 - 1 is never the argument of a tailed read in a sweep
            tailed = safe_rename(safe_contract(inl, left[r + 1], {("D1", "U")}), {"D": "D1"})
        return tailed

    def inline_down_tailed(self, l1, l2, left, right):
        """_inline_down_to_up_tailed[l1, l2]."""
        tailed = None
        for r in reversed(range(l1, self.L1 + 1)):
            if r == self.L1:
                inl = self.one
            else:
                inl = safe_contract(tailed, safe_rename(self.lat[r][l2], {"U": "U2"}), {("U2", "D"), ("L", "R")})
                inl = safe_contract(inl, safe_rename(left[r], {"U": "U1"}), {("U1", "D"), ("L", "R")})
            if r == 0:
                tailed = inl
            else:
                tailed = safe_rename(safe_contract(inl, right[r - 1], {("U3", "D")}), {"U": "U3"})
        return tailed

    # -- replace (:593-708) -------------------------------------------------------
    def replace_horizontal(self, l1, l2, t0, t1):
        """<s'|psi> with sites (l1, l2), (l1, l2 + 1) replaced by t0, t1."""
        up = self.up_to_down(l1 - 1)
        down = self.down_to_up(l1 + 1)
        left_part = self.inline_left_tailed(l1, l2 - 1, up, down)
        left_dot = down[l2]
        right_part = self.inline_right(l1, l2 + 2, up, down)
        right_dot = up[l2 + 1]
        res = safe_contract(left_part, safe_rename(t0, {"R": "R2"}), {("D", "U"), ("R2", "L")})
        res = safe_contract(res, safe_rename(left_dot, {"R": "R3"}), {("D", "U"), ("R3", "L")})
        res = safe_contract(res, safe_rename(right_dot, {"R": "R1"}), {("R1", "L")})
        res = safe_contract(res, safe_rename(t1, {"R": "R2"}), {("D", "U"), ("R2", "L")})
        return safe_contract(res, right_part, {("R1", "L1"), ("R2", "L2"), ("D", "U"), ("R3", "L3")})

    def replace_vertical(self, l1, l2, t0, t1):
        """<s'|psi> with sites (l1, l2), (l1 + 1, l2) replaced by t0, t1."""
        left = self.left_to_right(l2 - 1)
        right = self.right_to_left(l2 + 1)
        up_part = self.inline_up_tailed(l1 - 1, l2, left, right)
        up_dot = right[l1]
        down_part = self.inline_down_tailed(l1 + 2, l2, left, right)
        down_dot = left[l1 + 1]
        res = safe_contract(up_part, safe_rename(t0, {"D": "D2"}), {("R", "L"), ("D2", "U")})
        res = safe_contract(res, safe_rename(up_dot, {"D": "D3"}), {("R", "L"), ("D3", "U")})
        res = safe_contract(res, safe_rename(down_dot, {"D": "D1"}), {("D1", "U")})
        res = safe_contract(res, safe_rename(t1, {"D": "D2"}), {("R", "L"), ("D2", "U")})
        return safe_contract(res, down_part, {("D1", "U1"), ("D2", "U2"), ("R", "L"), ("D3", "U3")})

    def hole(self):
        """hole(()) with hint ("H", 0): _inline_right_to_left[0, 0]."""
        return self.inline_right(0, 0, [self.one] * self.L2, self.down_to_up(1), tailed_result=False)


def value(t):
    return float(t.data[0]) if t.size else 0.0


def replace(sites, config, cut, positions, occupations):
    """``Configuration.replace`` for a neighbouring pair ``positions`` =
    ((l1, l2), (l1', l2')) set to ``occupations``: the result tensor."""
    from .peps import shrink

    (a1, a2), (b1, b2) = positions
    env = Environments(project(sites, config), cut)
    t0 = shrink(sites[a1][a2], a1, a2, occupations[0])
    t1 = shrink(sites[b1][b2], b1, b2, occupations[1])
    if a1 == b1 and b2 == a2 + 1:
        return env.replace_horizontal(a1, a2, t0, t1)
    if a2 == b2 and b1 == a1 + 1:
        return env.replace_vertical(a1, a2, t0, t1)
    raise ValueError("replace of this site pair is outside the sweep's bonds")


def sweep(sites, config, cut, uniforms, alpha=1.0, t=1.0, V=1.0, table=None, target_uniforms=None):
    """One ``SweepSampling.__call__`` (sampling.py:117-154): bonds in
    ``bond_order``; the bond's occupations select their targets in the
    element table (``table``: {in occupations: {out occupations: element}},
    the keys of ``sampling.element_pool``; default the tV model's: a hop
    between a filled and an empty site when t != 0, the diagonal of a doubly
    occupied bond when V != 0); occupations without elements draw nothing.
    One target is drawn uniformly (``target_uniforms[k]`` picks it when a key
    has several: index floor(u n)); a target equal to the configuration is
    ``hole(())``, another one ``replace``; the acceptance is
    |wss/ws|^(2 alpha) n / n' against ``uniforms[k]``.  Returns the final
    configuration, the per-bond trace of (|ws| before, |wss| or None,
    accepted) and the final |ws|."""
    from .sampling import bond_order

    if table is None:
        table = {}
        if t != 0:
            table[(0, 1)] = {(1, 0): None}
            table[(1, 0)] = {(0, 1): None}
        if V != 0:
            table[(1, 1)] = {(1, 1): None}
    cfg = np.array(config, dtype=np.int64).copy()
    env = Environments(project(sites, cfg), cut)
    ws = value(env.hole())
    trace = []
    for k, (a, b) in enumerate(bond_order(len(cfg), len(cfg[0]))):
        key = (int(cfg[a]), int(cfg[b]))
        if key not in table or not table[key]:
            trace.append((abs(ws), None, False))
            continue
        outs = list(table[key])
        n = len(outs)
        target = outs[0] if n == 1 else outs[min(int(target_uniforms[k] * n), n - 1)]
        ns = len(table[target])
        if target == key:
            wss = value(Environments(project(sites, cfg), cut).hole())
        else:
            wss = value(replace(sites, cfg, cut, (a, b), target))
        p = abs(wss / ws) ** (2 * alpha) * n / ns
        acc = bool(uniforms[k] < p)
        trace.append((abs(ws), abs(wss), acc))
        if acc:
            ws = wss
            cfg[a], cfg[b] = target
    return cfg, trace, abs(ws)


def local_energy(sites, config, cut, H):
    """tetragono's observer energy (sampling_lattice/observer.py:313-389) for
    one configuration: ws = hole(()); each term's wss from ``replace`` (for
    a target equal to the configuration, ``hole(())``).  Returns None when
    ws is zero (the observer skips the configuration)."""
    from .peps import psi
    from .sampling import bond_order, element_pool, local_energy_term

    cfg = np.array(config, dtype=np.int64)
    if psi(sites, cfg, cut)[0] == 0:
        return None
    pool = element_pool(H)
    total = 0.0
    for pos in bond_order(len(cfg), len(cfg[0])):
        def wss_of(cfg2, occ_s, pos=pos):
            if all(int(cfg2[p]) == int(cfg[p]) for p in pos):
                return Environments(project(sites, cfg), cut).hole()
            return replace(sites, cfg, cut, pos, occ_s)

        total += local_energy_term(sites, cfg, cut, pos, H, pool, wss_of=wss_of)
    return total


def shrinker(site, occupation):
    """lattice.py:291-318 ``_get_shrinker``: ``P`` the dimension-one edge of
    the occupation, ``Q`` the conjugate of the physical edge, one unit entry."""
    from . import tat

    kind = site.kind
    phys = site.edge_by_name("P0")
    q = kind.make(occupation)
    sh = tat.Tensor(kind, ["P", "Q"], [tat.Edge(kind, [(q, 1)], False), phys.conjugate()])
    sh[{"Q": (kind.neg(q), 0), "P": (q, 0)}] = 1.0
    return sh


def holes(sites, config, cut):
    """``Configuration.holes`` (sampling_lattice/lattice.py:362-416) with
    ``hole(((l1, l2),))`` from the row environments (hint "H",
    single_layer_auxiliaries.py:721-729): <psi|s|d_x psi> / <psi|s|psi> for
    every site x, with the site's own edge names and its full physical
    edge.  Returns a list of rows of tensors."""
    cfg = np.array(config, dtype=np.int64)
    L1, L2 = len(cfg), len(cfg[0])
    env = Environments(project(sites, cfg), cut)
    ws = env.hole()
    w = value(ws)
    inv_ws = (ws * (1.0 / (w * w))).conjugate()
    all_name = {("T", "T")} | {(f"P_{a}_{b}_0", f"P_{a}_{b}_0") for a in range(L1) for b in range(L2)}
    out = []
    for l1 in range(L1):
        up = env.up_to_down(l1 - 1)
        down = env.down_to_up(l1 + 1)
        row = []
        for l2 in range(L2):
            left = env.inline_left_tailed(l1, l2 - 1, up, down)
            right = env.inline_right(l1, l2 + 1, up, down)
            big = safe_contract(left, right, {("R1", "L1"), ("R3", "L3")})
            hole = safe_rename(big, {"R2": "R0", "L2": "L0", "U": "U0", "D": "D0"})
            names = set(all_name)
            names.remove((f"P_{l1}_{l2}_0", f"P_{l1}_{l2}_0"))
            if "T" not in hole.names:
                names.remove(("T", "T"))
            hole = hole.contract(inv_ws, names)
            hole = safe_rename(hole, {"L0": "R", "R0": "L", "U0": "D", "D0": "U", f"P_{l1}_{l2}_0": "P0"})
            hole = hole.contract(shrinker(sites[l1][l2], int(cfg[l1][l2])), {("P0", "P")}).edge_rename({"Q": "P0"})
            row.append(hole)
        out.append(row)
    return out


def energy_gradient(sites, configs, cut, H, reweight=None):
    """observer.py:396-407, 543-555: 2 conj(<E Delta> - E <Delta>) per site
    (conj with the trivial metric, utility.py:230-231)
    with Delta = ``holes``, E the observer energy through ``replace``,
    averages weighted by ``reweight``; a configuration with psi = 0 is
    skipped.  Returns (gradient rows, energy)."""
    es, hs, ws = [], [], []
    for i, cfg in enumerate(configs):
        e = local_energy(sites, cfg, cut, H)
        if e is None:
            continue
        es.append(e)
        hs.append(holes(sites, cfg, cut))
        ws.append(1.0 if reweight is None else float(reweight[i]))
    ws = np.array(ws) / np.sum(ws)
    energy = float(np.sum(ws * np.array(es)))
    out = []
    for l1 in range(len(hs[0])):
        row = []
        for l2 in range(len(hs[0][0])):
            g = hs[0][l1][l2] * 0.0
            for w, e, h in zip(ws, es, hs):
                g = g + h[l1][l2] * (2.0 * w * (e - energy))
            # lattice_conjugate (utility.py:230-231): conjugate with the trivial metric
            row.append(g.conjugate(True))
        out.append(row)
    return out, energy


def natural_gradient(sites, configs, cut, H, step, error=0.0, reweight=None):
    """observer.py:576-675 (real tensors): conjugate gradient on
    (D^T D) x = D^T E with D = sqrt(w_s) (Delta_s - <Delta>), E = sqrt(w_s)
    (E_s - <E>); returns (2 x as site tensors through the trivial-metric
    conjugate, energy, steps, the D matrix, the E vector and 2 x itself)."""
    es, hs, ws = [], [], []
    for i, cfg in enumerate(configs):
        e = local_energy(sites, cfg, cut, H)
        if e is None:
            continue
        es.append(e)
        hs.append(holes(sites, cfg, cut))
        ws.append(1.0 if reweight is None else float(reweight[i]))
    ws = np.array(ws) / np.sum(ws)
    es = np.array(es)
    energy = float(np.sum(ws * es))
    L1, L2 = len(hs[0]), len(hs[0][0])
    arr = np.array([np.concatenate([h[l1][l2].data for l1 in range(L1) for l2 in range(L2)]) for h in hs])
    mean = (ws[:, None] * arr).sum(0)
    D = (arr - mean[None, :]) * np.sqrt(ws)[:, None]
    E = np.sqrt(ws) * (es - energy)
    b = D.T @ E
    b2 = float(b @ b)
    x = np.zeros_like(b)
    r = b.copy()
    p = r.copy()
    r2 = float(r @ r)
    t = 0
    while t < step and not (error != 0.0 and error ** 2 > r2 / b2):
        Dp = D @ p
        alpha = r2 / float(Dp @ Dp)
        x = x + alpha * p
        r = r - alpha * (D.T @ Dp)
        nr2 = float(r @ r)
        p = r + (nr2 / r2) * p
        r2 = nr2
        t += 1
    x = 2 * x
    out, off = [], 0
    for l1 in range(L1):
        row = []
        for l2 in range(L2):
            g = hs[0][l1][l2].copy()
            g.data[:] = x[off:off + g.size]
            off += g.size
            row.append(g.conjugate(True))
        out.append(row)
    return out, energy, t, D, E, x
