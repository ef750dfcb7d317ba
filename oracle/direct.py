"""CPU oracle for tetragono's direct sampling.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Restates, on top of ``oracle.tat`` and ``oracle.peps`` (paths relative to the
TNSP repository root, directory ``tetragono/tetragono/``):

* the double-layer boundary MPS ``_two_line_to_one_line``
  (auxiliaries/double_layer_auxiliaries.py:415-496): the ket layer absorbed
  and canonicalised first, then the bra layer with the physical and ``T``
  edges traced, both at cut ``cut2`` and with the singular values divided by
  their sum (``normalize=True``, sampling_lattice/sampling.py:267-272); the
  chain ``_down_to_up`` (:387-392) from the ones row;
* the three-line auxiliaries of one row (auxiliaries/
  three_line_auxiliaries.py): ``_two_line_to_one_line`` (:67-101) for "LR"
  and "RL", the chains ``_left_to_right`` / ``_right_to_left`` (:103-115)
  and ``hole`` (:117-153);
* ``DirectSampling.__call__`` (sampling_lattice/sampling.py:276-360) with
  ``_choice`` (:362-371) for sites of one orbit (the tV model): the row
  above comes from the single-layer boundary MPS of the sampled rows
  (``Configuration._up_to_down_site``), the rows below from the double
  layer, and each site's orbit is drawn from the diagonal of the site hole.
  After the last orbit of a site the reference also updates ``site_hole``,
  which nothing reads afterwards; that update is left out.

``possibility`` is the probability with which the configuration was drawn;
with no truncation and ``alpha = 1`` it equals |psi(s)|^2 / sum |psi|^2.
"""
import numpy as np

from .peps import one, safe_contract, safe_rename, shrink, two_line_to_one_line


def _contract_all(t1, t2, pairs):
    """``safe_contract(..., contract_all_physics_edges=True)``
    (utility.py:340-350): also every common name starting with ``P``."""
    extra = {(n, n) for n in t1.names if str(n)[0] == "P" and n in t2.names}
    return safe_contract(t1, t2, set(pairs) | extra)


def double_layer_two_line(udlr, line_1, line_2, cut, normalize=True):
    """double_layer_auxiliaries.py:415-496; ``line_2[i]`` = (ket, bra)."""
    up, down, left, right = udlr
    down_n, down_c = down + "N", down + "C"
    left0, left1, right0, right1 = left + "0", left + "1", right + "0", right + "1"
    length = len(line_1)

    def sweep(stage):
        for i in range(length - 1):
            q, r = stage[i].qr("r", {n for n in (right1, right0) if n in stage[i].names}, right, left)
            stage[i] = q
            stage[i + 1] = safe_contract(stage[i + 1], r, {(left1, right1), (left0, right0)})
        for i in reversed(range(1, length)):
            u, s, v = stage[i].svd({left}, right, left, left, right, cut)
            if normalize:
                s = s / s.norm_sum()
            stage[i] = v
            stage[i - 1] = safe_contract(safe_contract(stage[i - 1], u, {(right, left)}), s, {(right, left)})
        return stage

    stage_1 = sweep([safe_contract(safe_rename(line_1[i], {left: left1, right: right1}),
                                   safe_rename(line_2[i][0], {left: left0, right: right0, down: down_n}),
                                   {(down_n, up)}) for i in range(length)])
    return sweep([_contract_all(safe_rename(stage_1[i], {left: left1, right: right1}),
                                safe_rename(line_2[i][1], {left: left0, right: right0, down: down_c}),
                                {("T", "T"), (down_c, up)}) for i in range(length)])


def double_layer_down_to_up(sites, cut2):
    """``_down_to_up[l1]`` for l1 = L1 .. 0 (L1 is the ones row) of the
    double layer of the raw sites."""
    L1, L2 = len(sites), len(sites[0])
    o = one(sites[0][0].kind)
    out = {L1: [o] * L2}
    for l1 in reversed(range(L1)):
        row = [(sites[l1][l2], sites[l1][l2].conjugate()) for l2 in range(L2)]
        out[l1] = double_layer_two_line("DULR", out[l1 + 1], row, cut2)
    return out


def three_line_two_line(lr, line_1, line_2, cut):
    """three_line_auxiliaries.py:67-101; ``line_2`` = (n0, n1, t2, c1, c0),
    ``line_1`` the five tensors of the previous column."""
    left, right = lr
    right_n, right_c = right + "N", right + "C"
    rn0, nr = line_2[0].qr("q", {n for n in (right,) if n in line_2[0].names}, "D", "U")
    n12 = safe_contract(safe_rename(nr, {"D": "D2"}), safe_rename(line_1[0], {"D": "D1"}), {(left, right)})
    n12 = safe_contract(n12, safe_rename(line_1[1], {"D": "D1"}), {("D1", "U")})
    n12 = safe_contract(n12, safe_rename(line_2[1], {"D": "D2"}), {(right, left), ("D2", "U")})

    # the reference tests the names of line_2[0] here too (:78)
    rc0, cr = line_2[4].qr("q", {n for n in (right,) if n in line_2[0].names}, "D", "U")
    c12 = safe_contract(safe_rename(cr, {"D": "D2"}), safe_rename(line_1[4], {"D": "D1"}), {(left, right)})
    c12 = safe_contract(c12, safe_rename(line_1[3], {"D": "D1"}), {("D1", "U")})
    c12 = safe_contract(c12, safe_rename(line_2[3], {"D": "D2"}), {(right, left), ("D2", "U")})

    big = safe_contract(safe_rename(n12, {right: right_n, "U": "UN"}), safe_rename(line_1[2], {"UC": "UC1"}),
                        {("D1", "UN")})
    big = safe_contract(big, safe_rename(line_2[2], {"UC": "UC2"}), {("D2", "UN"), (right, left)})
    big = _contract_all(big, safe_rename(c12, {right: right_c, "U": "UC"}), {("UC1", "D1"), ("UC2", "D2"), ("T", "T")})

    u, s, v = big.svd({"UN", right_n}, "D", "UN", "UN", "D", cut)
    rn1 = safe_rename(u, {"UN": "U", right_n: right})
    big = safe_contract(v, s, {("UN", "D")})
    u, s, v = big.svd({"UC", right_c}, "D", "UC", "UC", "D", cut)
    rc1 = safe_rename(u, {"UC": "U", right_c: right})
    big = safe_contract(v, s, {("UC", "D")})
    return [rn0, rn1, big, rc1, rc0]


def three_line_hole(left, line, right):
    """three_line_auxiliaries.py:117-153: ``line`` = (n0, n1, t2, c1, c0) of
    the site, its physical edges already renamed ``O*`` (ket) / ``I*``
    (bra); ``left`` / ``right`` the column environments beside it."""
    n0, n1, t2, c1, c0 = line
    res = safe_contract(safe_rename(left[0], {"D": "D1"}), safe_rename(n0, {"D": "D2"}), {("R", "L")})
    res = safe_contract(res, safe_rename(right[0], {"D": "D3"}), {("R", "L")})
    res = safe_contract(res, safe_rename(left[1], {"D": "D1"}), {("D1", "U")})
    res = safe_contract(res, safe_rename(n1, {"D": "D2"}), {("D2", "U"), ("R", "L")})
    res = safe_contract(res, safe_rename(right[1], {"D": "D3"}), {("D3", "U"), ("R", "L")})
    res = safe_contract(res, safe_rename(left[2], {"UC": "U1"}), {("D1", "UN")})
    res = safe_contract(res, safe_rename(t2, {"UC": "U2"}), {("D2", "UN"), ("R", "L")})
    res = safe_contract(res, safe_rename(right[2], {"UC": "U3"}), {("D3", "UN"), ("R", "L")})
    res = safe_contract(res, safe_rename(left[3], {"U": "U1"}), {("U1", "D")})
    res = safe_contract(res, safe_rename(c1, {"U": "U2"}), {("U2", "D"), ("R", "L"), ("T", "T")})
    res = safe_contract(res, safe_rename(right[3], {"U": "U3"}), {("U3", "D"), ("R", "L")})
    res = safe_contract(res, safe_rename(left[4], {"U": "U1"}), {("U1", "D")})
    res = _contract_all(res, safe_rename(c0, {"U": "U2"}), {("U2", "D"), ("R", "L"), ("T", "T")})
    res = safe_contract(res, safe_rename(right[4], {"U": "U3"}), {("U3", "D"), ("R", "L")})
    return res


def shrinker(site, occupation):
    """lattice.py:291-318 ``_get_shrinker`` for orbit 0."""
    from .environment import shrinker as _shrinker

    return _shrinker(site, occupation)


def site_rho(hole):
    """sampling.py:318-331: the diagonal of the site hole, block by block
    in the order of the ket edge's segments, negative rounding clipped."""
    h = hole.edge_rename({"I0": "I", "O0": "O"}).transpose(["I", "O"])
    edge = h.edge_by_name("O")
    kind = h.kind
    rho = []
    for q, _ in edge.segments:
        rho.extend(np.diagonal(h.block_by_symmetries({"I": kind.neg(q), "O": q})))
    return np.maximum(np.real(np.array(rho)), 0), edge


def choice(p, rho):
    """sampling.py:362-371."""
    for i, r in enumerate(rho):
        p -= r
        if p < 0:
            return i
    return i


def direct_sample(sites, cut, cut2, uniforms=None, alpha=1.0, forced=None, down=None):
    """One ``DirectSampling.__call__``.  ``uniforms`` (L1*L2 numbers in
    [0, 1), site by site in row-major order) drive the draws; ``forced``
    (an L1 x L2 occupation array) replaces each draw by the given
    occupation, which gives the probability of drawing that configuration.
    Returns (possibility, configuration, rhos) -- ``rhos`` the normalised
    distribution at each site -- or (0, None, rhos) where every weight of a
    site vanished (the reference then draws again from scratch)."""
    L1, L2 = len(sites), len(sites[0])
    kind = sites[0][0].kind
    o = one(kind)
    if down is None:
        down = double_layer_down_to_up(sites, cut2)
    cfg = np.zeros((L1, L2), dtype=np.int64)
    possibility = 1.0
    rhos = []
    up = [o] * L2  # single-layer _up_to_down[l1 - 1]
    k = 0
    for l1 in range(L1):
        n0 = list(up)
        c0 = [t.conjugate() for t in n0]
        n1 = [sites[l1][l2] for l2 in range(L2)]
        t2 = down[l1 + 1]
        ones = [o] * 5
        column = lambda c: (n0[c], n1[c], t2[c], n1[c].conjugate(), c0[c])  # noqa: E731
        right_env = {L2: ones}
        for c in reversed(range(L2)):
            right_env[c] = three_line_two_line("RL", right_env[c + 1], column(c), cut)
        left_env = ones  # _left_to_right[l2 - 1]
        for l2 in range(L2):
            ket = safe_rename(n1[l2], {n: "O" + n[1:] for n in n1[l2].names if n.startswith("P")})
            bra = n1[l2].conjugate()
            bra = safe_rename(bra, {n: "I" + n[1:] for n in bra.names if n.startswith("P")})
            hole = three_line_hole(left_env, (n0[l2], ket, t2[l2], bra, c0[l2]), right_env[l2 + 1])
            rho, edge = site_rho(hole)
            if np.sum(rho) == 0:
                rhos.append(rho)
                return 0.0, None, rhos
            rho = rho ** alpha
            rho = rho / np.sum(rho)
            rhos.append(rho)
            if forced is not None:
                q = kind.make(int(forced[l1][l2]))
                pick = edge.index_by_point((q, 0)) if edge.position_by_symmetry(q) is not None else None
                if pick is None:
                    return 0.0, None, rhos
            else:
                pick = choice(uniforms[k], rho)
            k += 1
            possibility *= rho[pick]
            occ = int(np.asarray(edge.point_by_index(pick)[0]).reshape(-1)[0])
            cfg[l1, l2] = occ
            n1[l2] = n1[l2].contract(shrinker(sites[l1][l2], occ).edge_rename({"P": "P0"}), {("P0", "Q")})
            left_env = three_line_two_line("LR", left_env, column(l2), cut)
        row = [shrink(sites[l1][l2], l1, l2, int(cfg[l1, l2])) for l2 in range(L2)]
        up = two_line_to_one_line(up, row, cut, "UDLR")
    return possibility, cfg, rhos
