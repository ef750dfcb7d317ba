"""Sweep sampling with cached environments, many chains in lockstep.

tetragono evaluates a proposal psi(s') with ``Configuration.replace``
(tetragono/tetragono/auxiliaries/single_layer_auxiliaries.py:500-708, paths
relative to the TNSP repository root): only the two changed sites are
contracted, between environments cached from earlier proposals.  Along a
sweep (sampling_lattice/sampling.py:145-190: horizontal bonds row by row,
then vertical bonds column by column) the caches that ``replace`` reads are:

* horizontal pass, row l1: the boundary MPS of the rows above
  (``_up_to_down[l1-1]``, rebuilt after each row) and below
  (``_down_to_up[l1+1]``, built once at the start of the sweep), the inline
  environment right of the bond (``_inline_right_to_left_tailed``, built
  once per row from the row as it was when the row started: the columns
  right of the bond are unchanged until the sweep reaches them) and the one
  left of it (``_inline_left_to_right_tailed``, extended by one column after
  each bond with the column as decided);
* vertical pass, column l2: the same with the column MPS ``_left_to_right``
  / ``_right_to_left`` and the column inline environments.

``EnvironmentSampler`` keeps exactly those tensors for every chain, as
batches (``tnsp_amd.batch``): each contraction, QR and SVD of a step runs
for all chains in one launch per kernel kind.  The results equal the lazy
graph's, which recomputes an invalidated environment from the same inputs.

A doubly occupied bond of the tV model has only its diagonal element, so
``_single_term`` proposes the configuration itself; ``replace`` with no site
changed is ``hole(())`` (sampling_lattice/lattice.py:231-267), the
amplitude computed afresh, and the draw compares it with the carried psi.
``diagonal="hole"`` does that through the amplitude engine;
``diagonal="keep"`` skips it (the two amplitudes differ only by the
boundary-MPS truncation, so the step accepts with probability one up to
that difference).

Step by step, a refresh in the horizontal pass at row l1 restarts from the
boundary MPS of the rows below l1 built at the sweep start (they have not
moved since), as the lazy graph recomputes only the invalidated nodes; in
the vertical pass every row may have moved, and each lane evaluates the
amplitude with an engine of its own, so concurrent lanes do not queue on
one engine handle.

A hole(()) changes only the carried psi, never the configuration, so no
environment depends on it: only later acceptance tests read it.  With
``speculate=True`` the ``"hole"`` sweep runs speculatively: the lockstep pass
decides with the carried psi and records every hop (its replace value and
decision) and every doubly occupied bond (a hole(()) request whenever the
chain moved since its last one); all requests of the sweep are then
evaluated by the engine in one batch, and the decisions are replayed in bond
order with the true psi.  A chain whose replayed decisions all agree
followed tetragono's trajectory exactly and takes the replayed psi; a chain
with any disagreement is swept again from its start on the step-by-step path
with the same uniforms.  Either way every chain ends where the step-by-step
sweep ends, with the same psi.  This pays only where hole(()) and the
carried psi nearly agree (a mild truncation): on the random 8x8 D=8 PEPS at
Dc=64 they differ by a median factor of e^1.7 and nearly every chain is swept
again, so the default is step by step.
"""
import threading
from concurrent.futures import ThreadPoolExecutor

import numpy as np

from . import batch as Bt
from .elements import Moves, tensor_element, tv_table
from .peps import _one
from .sampler import bond_order


class _Speculation:
    """Record of one lane's speculative ``"hole"`` sweep: per bond the hop
    chains with their replace values and decisions, the doubly occupied
    chains with the hole(()) request each one reads; request r of chain c is
    the configuration c had when it reached a doubly occupied bond after
    moving (requests 0 .. n-1: the sweep-start psi of the n chains, which is
    hole(()) itself)."""

    def __init__(self, psi0):
        n = len(psi0)
        self.psi0 = psi0.copy()
        self.req = np.arange(n)  # current request of each chain, -1 after a move
        self.owner = [np.arange(n)]
        self.cfgs = []
        self.nreq = n
        self.events = []  # (bond, kind, chains, values, decisions)

    def hop(self, k, idx, wss, acc, n, ns):
        self.events.append((k, 0, idx, wss, acc, n, ns))
        self.req[idx[acc]] = -1

    def diag(self, k, idx, configs, n, ns):
        new = idx[self.req[idx] < 0]
        if len(new):
            self.req[new] = self.nreq + np.arange(len(new))
            self.nreq += len(new)
            self.owner.append(new)
            self.cfgs.append(configs[new].copy())
        self.events.append((k, 1, idx, self.req[idx].copy(), None, n, ns))

    def replay(self, u, alpha, hole):
        """Decisions in bond order with the true psi (``hole``: |hole(())| of
        every request).  Returns the final psi, the chains whose replayed
        decisions differ from the recorded ones anywhere, and, counted over
        the others, proposals, acceptances and engine refreshes."""
        psi = self.psi0.copy()
        bad = np.zeros(len(psi), dtype=bool)
        acc_n = np.zeros(len(psi), dtype=np.int64)
        prop_n = np.zeros(len(psi), dtype=np.int64)
        with np.errstate(divide="ignore", invalid="ignore"):
            for k, kind, idx, val, acc, n, ns in self.events:
                wss = val if kind == 0 else hole[val]
                # the expression of the step-by-step path, elementwise
                t = u[k, idx] < (wss / psi[idx]) ** (2 * alpha) * n / ns
                if kind == 0:
                    bad[idx[t != acc]] = True
                    prop_n[idx] += 1
                    acc_n[idx] += t
                psi[idx[t]] = wss[t]
        owner = np.concatenate(self.owner)
        good = ~bad
        refreshed = int(good[owner[len(psi):]].sum())
        return psi, bad, int(prop_n[good].sum()), int(acc_n[good].sum()), refreshed


class EnvironmentSampler:
    """Lockstep sweep sampling of B chains of a FermiU1 or FermiZ2 PEPS with
    the tV Hamiltonian (hops t, density interaction V).

    ``sweep()`` is ``SweepSampling.__call__``; ``local_energy()`` the
    observer's energy and ``holes()`` ``Configuration.holes``, all over the
    same cached environments.  ``cut`` overrides the PEPS's boundary-MPS
    bond dimension; ``diagonal`` chooses how diagonal targets (a doubly
    occupied bond of the tV model) are treated and ``speculate`` how
    ``"hole"`` is executed (see the module notes; ``respun`` counts chains
    swept again after a disagreement); ``lanes`` splits the chains into
    groups that run on their own host threads and HIP streams.

    ``hamiltonian``: the bond operator whose element table
    (``elements.tensor_element`` of a drop-in ``Tensor`` with edges I0, I1,
    O0, O1, or the table itself as a dict) drives the sweep; by default the
    tV model's (t, V).  A table with several targets per key draws the target
    from a second uniform per chain and bond, from a generator seeded with
    (seed, 1).  The observer's energy and the gradient use the tV elements
    and refuse another operator."""

    def __init__(self, engine, configs, t=1.0, V=1.0, alpha=1.0, seed=0, cut=None, diagonal="hole", lanes=1,
                 speculate=False, hamiltonian=None):
        if diagonal not in ("hole", "keep"):
            raise ValueError('diagonal must be "hole" or "keep"')
        self.lanes = max(1, int(lanes))
        self.engine = engine
        self.lattice = engine.lattice
        spec = engine.spec
        self.L1, self.L2 = spec.L1, spec.L2
        self.cut = spec.cut if cut is None else cut
        self.t, self.V, self.alpha = float(t), float(V), float(alpha)
        self.diagonal = diagonal
        self.speculate = bool(speculate)
        self._spec = None
        self.respun = 0
        self.configs = np.ascontiguousarray(np.asarray(configs), dtype=np.int8).copy()
        if self.configs.ndim != 3 or self.configs.shape[1:] != (self.L1, self.L2):
            raise ValueError("configs must have shape (chains, L1, L2)")
        self.gen = np.random.Generator(np.random.PCG64(seed))
        self.gen_target = np.random.Generator(np.random.PCG64(np.random.SeedSequence((seed, 1))))
        if hamiltonian is None:
            table = tv_table(self.t, self.V)
        elif isinstance(hamiltonian, dict):
            table = hamiltonian
        else:
            table = tensor_element(hamiltonian)
        self.custom_operator = hamiltonian is not None
        self.moves = Moves(table)
        self._v = None
        self.bonds = bond_order(self.L1, self.L2)
        self._bond_index = {bd: k for k, bd in enumerate(self.bonds)}
        self._proj = [[(self.lattice.projected(l1, l2, 0), self.lattice.projected(l1, l2, 1))
                       for l2 in range(self.L2)] for l1 in range(self.L1)]
        self._one = _one(spec, self.lattice.dtype)
        # made here, on the caller's stream, before any lane reads them
        self._shr = [[(self.lattice.shrinker(l1, l2, 0), self.lattice.shrinker(l1, l2, 1))
                      for l2 in range(self.L2)] for l1 in range(self.L1)]
        self.psi = np.abs(engine.evaluate(self.configs, cut=self.cut))
        # the refreshes evaluate a few configurations at a time: give the
        # memory of this full batch back to the sampler's environments
        engine.trim()
        if np.any(self.psi == 0):
            raise ValueError("initial configurations with psi = 0")
        self.proposals = 0
        self.accepted = 0
        self.replaced = 0
        self.refreshed = 0
        # called as recorder(a, b, before, psi_old, psi_new) for every bond
        # with proposals: flat configurations before the hop and |psi|
        self.recorder = None
        self._lock = threading.Lock()
        self._u = None
        self._pool = None
        self._streams = []
        # hole(()) engines of lanes 1, 2, ... (lane 0 uses ``engine``): the
        # step-by-step refreshes of concurrent lanes do not queue on one handle
        self._lane = 0
        self._lane_engines = {}
        self._hole_rows = None
        from .sampler import _fermion_order_index, hole_name_order

        self._hole_pos = {n: i for i, n in enumerate(hole_name_order(self.L1, self.L2))}
        # parity of the T edge at site (0, 0): -N for FermiU1, even for FermiZ2
        self._t_parity = spec.N % 2 if spec.kind == "FermiU1" else 0
        self._site_pos, self._t_pos = _fermion_order_index(self.L1, self.L2)

    def refresh_sites(self):
        """The lattice changed (``Lattice.apply_gradient``): project the sites
        again here and in the engine; chains keep their configurations and
        psi is recomputed."""
        self.engine.refresh_sites()
        for eng in self._lane_engines.values():
            eng.refresh_sites()
        self._proj = [[(self.lattice.projected(l1, l2, 0), self.lattice.projected(l1, l2, 1))
                       for l2 in range(self.L2)] for l1 in range(self.L1)]
        self.psi = np.abs(self.engine.evaluate(self.configs, cut=self.cut))

    # -- batches of site tensors --------------------------------------------------
    def _site(self, l1, l2, occ):
        """Projected site (l1, l2) for occupations ``occ`` (one per item)."""
        return Bt.TensorBatch.select(self._proj[l1][l2], occ)

    def _row(self, l1):
        return [self._site(l1, l2, self.configs[:, l1, l2]) for l2 in range(self.L2)]

    def _column(self, l2):
        return [self._site(l1, l2, self.configs[:, l1, l2]) for l1 in range(self.L1)]

    def _ones(self, n):
        return Bt.TensorBatch.repeat(self._one, n)

    # -- boundary MPS (single_layer_auxiliaries.py:463-498) -------------------------
    def _two_line(self, udlr, line_1, line_2):
        up, down, left, right = udlr
        l1n, l2n, r1n, r2n = left + "1", left + "2", right + "1", right + "2"
        n = len(line_1)
        dl = [Bt.contract(line_1[i].rename({left: l1n, right: r1n}), line_2[i].rename({left: l2n, right: r2n}),
                          [(down, up)]) for i in range(n)]
        for i in range(n - 1):
            q, r = Bt.qr(dl[i], "r", [x for x in (r1n, r2n) if x in dl[i].names], right, left)
            dl[i] = q
            dl[i + 1] = Bt.contract(dl[i + 1], r, [(l1n, r1n), (l2n, r2n)])
        for i in reversed(range(1, n)):
            u, s, v = Bt.svd(dl[i], [left], right, left, left, right, self.cut)
            dl[i] = v
            dl[i - 1] = Bt.contract(Bt.contract(dl[i - 1], u, [(right, left)]), s, [(right, left)])
        return dl

    # -- one bond ---------------------------------------------------------------------
    def _term(self, a, b, replace):
        """``_single_term`` on bond (a, b) for every chain: the bond's
        occupations select their targets in the element table, one target is
        drawn per chain and accepted with |psi(s')/psi(s)|^(2 alpha) n / n'
        against one uniform per chain.  An off-diagonal target reads psi(s')
        from ``replace(idx, t0, t1)``; a target equal to the configuration
        reads hole(()) (module notes)."""
        na = self.configs[:, a[0], a[1]].copy()
        nb = self.configs[:, b[0], b[1]].copy()
        k = self._bond_index[(a, b)]
        u = self._u[k]
        active, ta, tb, n, ns = self.moves.draw(na, nb, None if self._v is None else self._v[k])
        moved = (ta != na) | (tb != nb)
        hop = np.nonzero(active & moved)[0]
        dbl = np.nonzero(active & ~moved)[0]
        if len(hop):
            t0 = self._site(a[0], a[1], ta[hop])
            t1 = self._site(b[0], b[1], tb[hop])
            wss = np.abs(replace(hop, t0, t1).values())
            if self.recorder is not None:
                with self._lock:
                    self.recorder(a, b, self.configs[hop].reshape(len(hop), -1).copy(), self.psi[hop].copy(), wss)
            self.replaced += len(hop)
            self.proposals += len(hop)
            with np.errstate(divide="ignore", invalid="ignore"):
                # a carried psi of zero (every sector cut away) accepts any
                # non-zero proposal, as the tensor ratio of the reference does;
                # the factor n / n' in tetragono's order of operations
                acc = u[hop] < (wss / self.psi[hop]) ** (2 * self.alpha) * n[hop] / ns[hop]
            take = hop[acc]
            if self._spec is not None:
                self._spec.hop(k, hop, wss, acc, n[hop], ns[hop])
            self.configs[take, a[0], a[1]], self.configs[take, b[0], b[1]] = ta[take], tb[take]
            self.psi[take] = wss[acc]
            self.accepted += int(acc.sum())
            self._hole_cache[take] = np.nan
        if len(dbl) and self.diagonal == "hole" and self._spec is not None:
            self._spec.diag(k, dbl, self.configs, n[dbl], ns[dbl])
        elif len(dbl) and self.diagonal == "hole":
            # hole(()) is recomputed only for chains that moved since it was
            # last known, as the lazy graph's cache does
            need = dbl[np.isnan(self._hole_cache[dbl])]
            if len(need):
                if self._hole_rows is not None:
                    self._hole_cache[need] = self._hole_from_rows(need, *self._hole_rows)
                else:
                    self._hole_cache[need] = np.abs(self._engine_of_lane().evaluate(self.configs[need], cut=self.cut))
                self.refreshed += len(need)
            wss = self._hole_cache[dbl]
            with np.errstate(divide="ignore", invalid="ignore"):
                acc = u[dbl] < (wss / self.psi[dbl]) ** (2 * self.alpha) * n[dbl] / ns[dbl]
            self.psi[dbl[acc]] = wss[acc]

    def _hole_from_rows(self, idx, l1, down):
        """|hole(())| of chains ``idx`` during the horizontal pass at row l1:
        _inline_right_to_left[0, 0] over _down_to_up[1], whose rows below l1
        are unchanged since ``down`` was built at the sweep start, so the
        chain restarts from down[l1 + 1] (the lazy graph recomputes only the
        nodes the moved sites invalidated: rows l1 .. 1 and row 0)."""
        L2 = self.L2
        one = self._ones(len(idx))
        d = [t.take(idx) for t in down[l1 + 1]]
        for k in range(l1, 0, -1):
            d = self._two_line("DULR", d, [self._site(k, c, self.configs[idx, k, c]) for c in range(L2)])
        row0 = [self._site(0, c, self.configs[idx, 0, c]) for c in range(L2)]
        return np.abs(self._right_tailed(row0, [one] * L2, d, one)[0].values())

    def _engine_of_lane(self):
        """The amplitude engine this lane evaluates hole(()) with: the
        sampler's own for lane 0, one of the same PEPS and precision per
        further lane (made on first use), so that lanes refresh
        concurrently on their own streams."""
        if self._lane == 0:
            return self.engine
        with self._lock:
            eng = self._lane_engines.get(self._lane)
            if eng is None:
                from .engine import AmplitudeEngine

                # a few configurations per call: planned on this thread and
                # one helper, not on a pool of its own per lane
                eng = AmplitudeEngine(self.lattice, threads=2, chunk=256, lanes=1, precision=self.engine.precision)
                self._lane_engines[self._lane] = eng
        return eng

    # -- passes -----------------------------------------------------------------------
    # -- row environments (single_layer_auxiliaries.py:254-312, 437-442) --------------
    def _down_mps(self, one):
        """_down_to_up[l1] for l1 = L1 .. 1 from the current rows."""
        down = {self.L1: [one] * self.L2}
        for l1 in reversed(range(1, self.L1)):
            down[l1] = self._two_line("DULR", down[l1 + 1], self._row(l1))
        return down

    def _right_tailed(self, row, up, dn, one):
        """_inline_right_to_left_tailed[l1, c] for c = L2 .. 0 (:283-312);
        ``up`` = _up_to_down[l1-1], ``dn`` = _down_to_up[l1+1]."""
        rt, tailed = {}, None
        for c in reversed(range(self.L2 + 1)):
            if c == self.L2:
                right = one
            else:
                right = Bt.contract(tailed, row[c].rename({"L": "L2"}), [("L2", "R"), ("U", "D")])
                right = Bt.contract(right, up[c].rename({"L": "L1"}), [("L1", "R"), ("U", "D")])
            tailed = right if c == 0 else Bt.contract(right, dn[c - 1], [("L3", "R")]).rename({"L": "L3"})
            rt[c] = tailed
        return rt

    @staticmethod
    def _left_start(one, up):
        """_inline_left_to_right_tailed[l1, -1] (:254-282)."""
        return Bt.contract(one, up[0], [("R1", "L")]).rename({"R": "R1"})

    @staticmethod
    def _left_extend(lt, site, dn_c, up_next):
        """_inline_left_to_right_tailed[l1, c] from the one at c - 1 and the
        site at column c (``dn_c`` = _down_to_up_site[l1+1, c], ``up_next`` =
        _up_to_down_site[l1-1, c+1])."""
        left = Bt.contract(lt, site.rename({"R": "R2"}), [("R2", "L"), ("D", "U")])
        left = Bt.contract(left, dn_c.rename({"R": "R3"}), [("R3", "L"), ("D", "U")])
        return Bt.contract(left, up_next, [("R1", "L")]).rename({"R": "R1"})

    def _horizontal(self, log=None, observe=False):
        L1, L2 = self.L1, self.L2
        B = len(self.configs)
        one = self._ones(B)
        down = self._down_mps(one)
        up = [one] * L2
        for l1 in range(L1):
            # rows below l1 are as they were when down was built
            self._hole_rows = (l1, down)
            dn = down[l1 + 1]
            row = self._row(l1)
            rt = self._right_tailed(row, up, dn, one)
            if l1 == 0:
                # hole(()) = _inline_right_to_left[0, 0]: psi at the start of the sweep
                if observe:
                    self._hole = rt[0].values()
                else:
                    self.psi = np.abs(rt[0].values())
                    self._hole_cache = self.psi.copy()
                    if self._spec is not None:
                        self._spec = _Speculation(self.psi)
            lt = self._left_start(one, up)
            for l2 in range(L2 - 1):
                def replace(idx, t0, t1, lt=lt, l2=l2):
                    # single_layer_auxiliaries.py:641-651
                    res = Bt.contract(lt.take(idx), t0.rename({"R": "R2"}), [("D", "U"), ("R2", "L")])
                    res = Bt.contract(res, dn[l2].take(idx).rename({"R": "R3"}), [("D", "U"), ("R3", "L")])
                    res = Bt.contract(res, up[l2 + 1].take(idx).rename({"R": "R1"}), [("R1", "L")])
                    res = Bt.contract(res, t1.rename({"R": "R2"}), [("D", "U"), ("R2", "L")])
                    return Bt.contract(res, rt[l2 + 2].take(idx), [("R1", "L1"), ("R2", "L2"), ("D", "U"),
                                                                    ("R3", "L3")])

                (self._observe if observe else self._term)((l1, l2), (l1, l2 + 1), replace)
                # the column as decided
                lt = self._left_extend(lt, self._site(l1, l2, self.configs[:, l1, l2]), dn[l2], up[l2 + 1])
            if l1 < L1 - 1:
                up = self._two_line("UDLR", up, self._row(l1))
            if log:
                log(f"row {l1} done")
        self._hole_rows = None

    def _vertical(self, log=None, observe=False):
        L1, L2 = self.L1, self.L2
        B = len(self.configs)
        one = self._ones(B)
        rcol = {L2: [one] * L1}
        for l2 in reversed(range(1, L2)):
            rcol[l2] = self._two_line("RLUD", rcol[l2 + 1], self._column(l2))
        lcol = [one] * L1
        for l2 in range(L2):
            rc = rcol[l2 + 1]
            col = self._column(l2)
            # _inline_down_to_up_tailed[r, l2] for r = L1 .. 0 (:343-372)
            dt, tailed = {}, None
            for r in reversed(range(L1 + 1)):
                if r == L1:
                    inl = one
                else:
                    inl = Bt.contract(tailed, col[r].rename({"U": "U2"}), [("U2", "D"), ("L", "R")])
                    inl = Bt.contract(inl, lcol[r].rename({"U": "U1"}), [("U1", "D"), ("L", "R")])
                tailed = inl if r == 0 else Bt.contract(inl, rc[r - 1], [("U3", "D")]).rename({"U": "U3"})
                dt[r] = tailed
            # _inline_up_to_down_tailed[-1, l2] (:314-342)
            ut = Bt.contract(one, lcol[0], [("D1", "U")]).rename({"D": "D1"})
            for l1 in range(L1 - 1):
                def replace(idx, t0, t1, ut=ut, l1=l1):
                    # single_layer_auxiliaries.py:652-687
                    res = Bt.contract(ut.take(idx), t0.rename({"D": "D2"}), [("R", "L"), ("D2", "U")])
                    res = Bt.contract(res, rc[l1].take(idx).rename({"D": "D3"}), [("R", "L"), ("D3", "U")])
                    res = Bt.contract(res, lcol[l1 + 1].take(idx).rename({"D": "D1"}), [("D1", "U")])
                    res = Bt.contract(res, t1.rename({"D": "D2"}), [("R", "L"), ("D2", "U")])
                    return Bt.contract(res, dt[l1 + 2].take(idx), [("D1", "U1"), ("D2", "U2"), ("R", "L"),
                                                                    ("D3", "U3")])

                (self._observe if observe else self._term)((l1, l2), (l1 + 1, l2), replace)
                site = self._site(l1, l2, self.configs[:, l1, l2])
                inl = Bt.contract(ut, site.rename({"D": "D2"}), [("D2", "U"), ("R", "L")])
                inl = Bt.contract(inl, rc[l1].rename({"D": "D3"}), [("D3", "U"), ("R", "L")])
                ut = Bt.contract(inl, lcol[l1 + 1], [("D1", "U")]).rename({"D": "D1"})
            if l2 < L2 - 1:
                lcol = self._two_line("LRUD", lcol, self._column(l2))
            if log:
                log(f"column {l2} done")

    # -- local energy (observer.py:313-389) ------------------------------------------
    def _reorder_sign(self, names, flat):
        """Sign that brings a one-element tensor with edges ``names`` (all
        of dimension one: P_{l1}_{l2}_0 and T) to the edge order of
        ``hole(())``: (-1)^(pairs of odd edges whose order flips), the
        transpose sign of TAT's edge operator for such tensors.  ``flat``:
        occupations of the configuration the tensor belongs to."""
        pos = np.empty(len(names), dtype=np.int64)
        odd = np.empty((len(flat), len(names)), dtype=np.int64)
        for j, n in enumerate(names):
            pos[j] = self._hole_pos[n]
            if n == "T":
                odd[:, j] = self._t_parity
            elif n.startswith("P_"):
                _, a, b, _ = n.split("_")
                odd[:, j] = flat[:, int(a) * self.L2 + int(b)]
            else:
                raise ValueError(f"replace left a virtual edge {n} uncontracted")
        flips = np.triu(pos[:, None] > pos[None, :], 1).astype(np.int64)
        count = np.einsum("ci,ij,cj->c", odd, flips, odd)
        return 1 - 2 * (count % 2)

    def _observe(self, a, b, replace):
        """One bond of the energy observer: -t (-1)^k psi(s^ab) / psi(s) for
        every chain with one particle on the bond (psi(s^ab) from
        ``replace``, brought to the edge order of ``hole(())``)."""
        na = self.configs[:, a[0], a[1]]
        nb = self.configs[:, b[0], b[1]]
        hop = np.nonzero(na != nb)[0]
        if len(hop) == 0:
            return
        res = replace(hop, self._site(a[0], a[1], nb[hop]), self._site(b[0], b[1], na[hop]))
        before = self.configs[hop].reshape(len(hop), -1).astype(np.int64)
        after = before.copy()
        ia, ib = a[0] * self.L2 + a[1], b[0] * self.L2 + b[1]
        after[:, [ia, ib]] = after[:, [ib, ia]]
        psi_new = res.values() * self._reorder_sign(res.names, after)
        coef = -self.t * self._hop_sign(before, ia, ib)
        with np.errstate(divide="ignore", invalid="ignore"):
            self._energy[hop] += coef * psi_new / self._hole[hop]

    def _hop_sign(self, flat, ia, ib):
        pa, pb = self._site_pos[ia], self._site_pos[ib]
        lo, hi = min(pa, pb), max(pa, pb)
        between = (self._site_pos > lo) & (self._site_pos < hi)
        k = flat[:, between].sum(1)
        if lo < self._t_pos < hi:
            k = k + self._t_parity
        return 1 - 2 * (k % 2)

    def local_energy(self, configs=None, log=None):
        """E_loc of every chain as tetragono's observer computes it
        (observer.py:313-389): psi(s) = hole(()), each hop term's psi(s')
        from ``replace`` over the cached environments, the diagonal term
        V n_i n_j.  Chains whose psi(s) is exactly zero get NaN (the observer
        skips such configurations).  ``configs`` defaults to the chains."""
        if self.custom_operator:
            raise NotImplementedError("the observer's energy is the tV model's; this sampler was given another "
                                      "bond operator")
        cfg = self.configs if configs is None else np.ascontiguousarray(np.asarray(configs), dtype=np.int8)
        out = np.zeros(len(cfg))

        def run(child, idx):
            child.configs = cfg[idx].copy()
            flat = child.configs.reshape(len(idx), -1).astype(np.int64)
            child._energy = np.zeros(len(idx))
            for (a, b) in child.bonds:
                ia, ib = a[0] * child.L2 + a[1], b[0] * child.L2 + b[1]
                child._energy += child.V * flat[:, ia] * flat[:, ib]
            child._horizontal(log, observe=True)
            child._vertical(log, observe=True)
            e = child._energy
            e[child._hole == 0] = np.nan
            out[idx] = e

        self._in_lanes(len(cfg), run)
        return out

    def sweep(self, log=None):
        """One ``SweepSampling.__call__`` for every chain; returns the
        configurations.  ``self.psi`` holds |psi| as carried by the sweep
        (the last accepted ``replace`` value of each chain).  ``log``, if
        given, is called with a progress line after every row and column.
        One uniform per chain and bond, drawn for all chains bond by bond,
        whatever the number of lanes."""
        self._u = self.gen.random((len(self.bonds), len(self.configs)))
        self._v = self.gen_target.random((len(self.bonds), len(self.configs))) if self.moves.multi else None
        diagonal_targets = any(k in outs for k, outs in self.moves.targets.items())
        if self.diagonal == "hole" and self.speculate and self.recorder is None and diagonal_targets:
            return self._sweep_speculative(log)

        def run(child, idx):
            child.configs = self.configs[idx].copy()
            child.psi = self.psi[idx].copy()
            child._hole_cache = np.full(len(idx), np.nan)
            child._u = self._u[:, idx]
            child._v = None if self._v is None else self._v[:, idx]
            child._spec = None
            child._horizontal(log)
            child._vertical(log)
            self.configs[idx] = child.configs
            self.psi[idx] = child.psi

        self._in_lanes(len(self.configs), run)
        return self.configs

    def _sweep_speculative(self, log=None):
        """The ``"hole"`` sweep run speculatively, then replayed (module notes)."""
        start = self.configs.copy()
        records = []

        def run(child, idx):
            child.configs = self.configs[idx].copy()
            child.psi = self.psi[idx].copy()
            child._hole_cache = np.full(len(idx), np.nan)
            child._u = self._u[:, idx]
            child._v = None if self._v is None else self._v[:, idx]
            child._spec = True  # replaced by the record once row 0 gives psi
            child._horizontal(log)
            child._vertical(log)
            records.append((idx, child._spec, child.configs))

        before = (self.proposals, self.replaced, self.accepted)
        self._in_lanes(len(self.configs), run)
        # the lanes counted speculative decisions: the replay counts again
        self.proposals, self.replaced, self.accepted = before
        cfgs = [c for _, rec, _ in records for c in rec.cfgs]
        if cfgs:
            with self._lock:
                fresh = np.abs(self.engine.evaluate(np.concatenate(cfgs), cut=self.cut))
        if log:
            log(f"hole(()) of {sum(len(c) for c in cfgs)} configurations evaluated")
        off = 0
        redo = []
        for idx, rec, cfg in records:
            nfresh = rec.nreq - len(idx)
            hole = np.concatenate([rec.psi0, fresh[off:off + nfresh]]) if nfresh else rec.psi0
            off += nfresh
            psi, bad, props, accs, refreshed = rec.replay(self._u[:, idx], self.alpha, hole)
            self.configs[idx] = cfg
            self.psi[idx] = psi
            self.proposals += props
            self.replaced += props
            self.accepted += accs
            self.refreshed += refreshed
            redo.append(idx[bad])
        redo = np.concatenate(redo)
        if len(redo):
            # a disagreement: these chains take the step-by-step sweep again
            # from their start, on the same uniforms
            self.respun += len(redo)
            child = self._child()
            child.configs = start[redo].copy()
            child.psi = self.psi[redo].copy()
            child._hole_cache = np.full(len(redo), np.nan)
            child._u = self._u[:, redo]
            child._v = None if self._v is None else self._v[:, redo]
            child._spec = None
            child._horizontal(log)
            child._vertical(log)
            self.configs[redo] = child.configs
            self.psi[redo] = child.psi
            self.proposals += child.proposals
            self.replaced += child.replaced
            self.accepted += child.accepted
            self.refreshed += child.refreshed
        return self.configs

    # -- holes (sampling_lattice/lattice.py:362-416) ------------------------------------
    def _holes_group(self):
        """Holes of this object's chains: every site's hole(((l1, l2),))
        from the row environments (single_layer_auxiliaries.py:721-729),
        then the normalisation, renaming and physical expansion of
        ``Configuration.holes``."""
        L1, L2 = self.L1, self.L2
        B = len(self.configs)
        one = self._ones(B)
        down = self._down_mps(one)
        up = [one] * L2
        raw = [[None] * L2 for _ in range(L1)]
        ws = None
        for l1 in range(L1):
            dn = down[l1 + 1]
            row = self._row(l1)
            rt = self._right_tailed(row, up, dn, one)
            if l1 == 0:
                ws = rt[0]
            lt = self._left_start(one, up)
            for l2 in range(L2):
                big = Bt.contract(lt, rt[l2 + 1], [("R1", "L1"), ("R3", "L3")])
                raw[l1][l2] = big.rename({"R2": "R0", "L2": "L0", "U": "U0", "D": "D0"})
                if l2 < L2 - 1:
                    lt = self._left_extend(lt, row[l2], dn[l2], up[l2 + 1])
            if l1 < L1 - 1:
                up = self._two_line("UDLR", up, row)
        w = ws.values()
        with np.errstate(divide="ignore"):
            factor = np.where(w != 0, 1.0 / (w * w), 0.0)
        inv_ws = Bt.conjugate(Bt.scale(ws, factor))
        physics = [f"P_{a}_{b}_0" for a in range(L1) for b in range(L2)]
        out = [[None] * L2 for _ in range(L1)]
        for l1 in range(L1):
            for l2 in range(L2):
                h = raw[l1][l2]
                own = f"P_{l1}_{l2}_0"
                pairs = [(n, n) for n in physics if n != own]
                if "T" in h.names:
                    pairs.append(("T", "T"))
                h = Bt.contract(h, inv_ws, pairs)
                h = h.rename({"L0": "R", "R0": "L", "U0": "D", "D0": "U", own: "P0"})
                sh = Bt.TensorBatch.select(self._shr[l1][l2], self.configs[:, l1, l2])
                out[l1][l2] = Bt.contract(h, sh, [("P0", "P")]).rename({"Q": "P0"})
        return out, w

    def holes(self, configs=None):
        """``Configuration.holes`` (lattice.py:362-416) of every chain:
        <psi|s|d_x psi> / <psi|s|psi> for each site x, a tensor with the
        site's own edge names and full physical edge.  Returns (holes,
        psi): ``holes[l1][l2]`` a TensorBatch over the chains, ``psi`` the
        hole(()) values (a chain with psi = 0 gets zero holes)."""
        cfg = self.configs if configs is None else np.ascontiguousarray(np.asarray(configs), dtype=np.int8)
        parts = {}

        def run(child, idx):
            child.configs = cfg[idx].copy()
            parts[int(idx[0])] = child._holes_group()

        self._in_lanes(len(cfg), run)
        order = [parts[k] for k in sorted(parts)]
        holes = [[Bt.TensorBatch.concat([p[0][l1][l2] for p in order]) for l2 in range(self.L2)]
                 for l1 in range(self.L1)]
        return holes, np.concatenate([p[1] for p in order])

    def energy_gradient(self, configs=None, reweight=None):
        """The observer's energy gradient (observer.py:396-407, 543-555) over
        the chains: 2 conj(<E Delta> - E <Delta>) per site (conj with the
        trivial metric, utility.py:230-231), Delta the holes,
        averages weighted by ``reweight`` (default 1: the chains sample
        |psi|^2); chains with psi(s) = 0 are skipped as the observer skips
        them.  Returns (gradient, energy): ``gradient[l1][l2]`` a drop-in
        Tensor with the site's edges, ``energy`` the weighted mean local
        energy."""
        torch = Bt._torch()
        holes, psi = self.holes(configs)
        e = self.local_energy(configs)
        w = np.ones(len(e)) if reweight is None else np.asarray(reweight, dtype=np.float64).copy()
        w[(psi == 0) | ~np.isfinite(e)] = 0.0
        if w.sum() <= 0:
            raise ValueError("no chain with a non-zero amplitude")
        w = w / w.sum()
        energy = float(np.sum(w * np.where(w > 0, e, 0.0)))
        coef = np.where(w > 0, 2.0 * w * (e - energy), 0.0)
        out = []
        for l1 in range(self.L1):
            row = []
            for l2 in range(self.L2):
                h = holes[l1][l2]
                if len(set(h.sizes.tolist())) != 1:
                    raise RuntimeError("holes of one site differ in structure between chains")
                n = int(h.sizes[0])
                stack = torch.stack([h.bufs[b][o:o + n] for b, o in zip(h.buf_idx, h.offs)])
                c = torch.as_tensor(coef, dtype=stack.dtype, device=stack.device)
                g = h.tensor(0)
                g._data = (c[:, None] * stack).sum(0)
                # lattice_conjugate (utility.py:230-231): the trivial metric
                row.append(g.conjugate(trivial_metric=True))
            out.append(row)
        return out, energy

    def natural_gradient(self, step, error=0.0, configs=None, reweight=None):
        """The observer's natural gradient by conjugate gradient
        (observer.py:576-675): solve (D^H D) x = D^H E with
        D = sqrt(w_s) (Delta_s - <Delta>) and E = conj(sqrt(w_s) (E_s - <E>))
        over the chains, ``step`` iterations at most, stopping early when
        |r|^2 / |b|^2 < error^2; the result 2 x goes back to site tensors and
        through the trivial-metric conjugate.  Returns (gradient, energy,
        steps taken)."""
        torch = Bt._torch()
        holes, psi = self.holes(configs)
        e = self.local_energy(configs)
        w = np.ones(len(e)) if reweight is None else np.asarray(reweight, dtype=np.float64).copy()
        w[(psi == 0) | ~np.isfinite(e)] = 0.0
        if w.sum() <= 0:
            raise ValueError("no chain with a non-zero amplitude")
        w = w / w.sum()
        energy = float(np.sum(w * np.where(w > 0, e, 0.0)))
        # Delta (chains x parameters): the holes of all sites side by side
        blocks, shapes = [], []
        for l1 in range(self.L1):
            for l2 in range(self.L2):
                h = holes[l1][l2]
                n = int(h.sizes[0])
                blocks.append(torch.stack([h.bufs[b][o:o + n] for b, o in zip(h.buf_idx, h.offs)]))
                shapes.append((l1, l2, n))
        delta = torch.cat(blocks, dim=1)
        wt = torch.as_tensor(w, dtype=delta.dtype, device=delta.device)
        mean = (wt[:, None] * delta).sum(0)
        D = (delta - mean[None, :]) * wt.sqrt()[:, None]
        E = torch.as_tensor(np.where(w > 0, np.sqrt(w) * (e - energy), 0.0), dtype=delta.dtype, device=delta.device)
        b = D.T @ E
        b2 = float(b @ b)
        x = torch.zeros_like(b)
        r = b.clone()
        p = r.clone()
        r2 = float(r @ r)
        t = 0
        while t < step and not (error != 0.0 and b2 > 0 and error ** 2 > r2 / b2):
            Dp = D @ p
            alpha = r2 / float(Dp @ Dp)
            x = x + alpha * p
            r = r - alpha * (D.T @ Dp)
            nr2 = float(r @ r)
            p = r + (nr2 / r2) * p
            r2 = nr2
            t += 1
        x = 2 * x
        out = [[None] * self.L2 for _ in range(self.L1)]
        off = 0
        for l1, l2, n in shapes:
            g = holes[l1][l2].tensor(0)
            g._data = x[off:off + n].clone()
            off += n
            out[l1][l2] = g.conjugate(trivial_metric=True)
        return out, energy, t

    # -- lanes --------------------------------------------------------------------------
    def _child(self):
        """A view of this sampler for one group of chains: shares the engine,
        projected sites, lock and recorder; counts its own proposals."""
        c = object.__new__(EnvironmentSampler)
        c.__dict__.update(self.__dict__)
        c.proposals = c.accepted = c.replaced = c.refreshed = 0
        return c

    def _in_lanes(self, n, run):
        """run(child, chain indices) for ``lanes`` contiguous groups of the
        n chains, each on its own host thread and HIP stream (the batched
        ops of one group overlap the host work and waits of the others);
        one group runs on the caller's thread and stream."""
        groups = [g for g in np.array_split(np.arange(n), min(self.lanes, max(1, n))) if len(g)]
        children = [self._child() for _ in groups]
        for i, c in enumerate(children):
            c._lane = i
        if len(groups) == 1:
            run(children[0], groups[0])
        else:
            import torch

            if self._pool is None or len(self._streams) < len(groups):
                # persistent workers: the native library keeps per-thread
                # task staging and pinned buffers
                self._pool = ThreadPoolExecutor(max_workers=len(groups))
                self._streams = [torch.cuda.Stream() for _ in groups]

            def work(child, idx, stream):
                with torch.cuda.stream(stream):
                    run(child, idx)
                stream.synchronize()

            main = torch.cuda.current_stream()
            streams = self._streams[:len(groups)]
            for st in streams:
                st.wait_stream(main)
            futures = [self._pool.submit(work, c, g, st) for c, g, st in zip(children, groups, streams)]
            errors = [f.exception() for f in futures]
            for st in streams:
                main.wait_stream(st)
            for e in errors:
                if e is not None:
                    raise e
        for c in children:
            self.proposals += c.proposals
            self.accepted += c.accepted
            self.replaced += c.replaced
            self.refreshed += c.refreshed
