"""Direct sampling of a PEPS, many samples in lockstep.

tetragono's ``DirectSampling`` (tetragono/tetragono/sampling_lattice/
sampling.py:252-371, paths relative to the TNSP repository root) draws a
configuration site by site, each occupation from its marginal given the
sites already drawn:

* the rows below the current row enter through the double-layer boundary MPS
  ``_down_to_up`` of <psi|psi> (auxiliaries/double_layer_auxiliaries.py:
  387-392, 415-496), built once per PEPS at bond dimension ``cut2`` with
  normalised singular values;
* the rows above through the single-layer boundary MPS of the sampled rows
  (``Configuration._up_to_down``, auxiliaries/single_layer_auxiliaries.py:
  463-498), ket and bra;
* within the row, the three-line auxiliaries (auxiliaries/
  three_line_auxiliaries.py) contract the column environments left and right
  of the site (``_left_to_right`` / ``_right_to_left``, :67-115) around the
  site's open ket and bra physical edges (``hole``, :117-153); the diagonal
  of that hole is the weight of each occupation (sampling.py:318-331).

The returned possibility is the probability with which the sample was drawn;
the observer weighs a sample by |psi(s)|^2 / possibility
(sampling_lattice/observer.py:340).

``DirectSampler`` draws B samples together: the double layer is computed once
and shared, every per-sample contraction, QR and SVD of a step runs for all
samples in one launch per kernel kind (``tnsp_amd.batch``).  A sample whose
weights all vanish at some site is drawn again, as the reference restarts
such a draw.
"""
import ctypes

import numpy as np

from . import _lib
from . import batch as Bt
from .peps import _one


def _block_key(handle):
    """Block positions and offsets of a native shape handle, as bytes: equal
    keys mean the same block list in the same storage places."""
    L = _lib.lib()
    h = ctypes.c_void_p(handle)
    rank, nb = ctypes.c_int32(), ctypes.c_int64()
    _lib.check(L.tns_shape_counts(h, None, ctypes.byref(rank), None, ctypes.byref(nb), None))
    pos = np.zeros(max(1, nb.value * rank.value), dtype=np.int64)
    off = np.zeros(max(1, nb.value), dtype=np.int64)
    _lib.check(L.tns_shape_blocks(h, pos.ctypes.data_as(_lib.c_i64p), off.ctypes.data_as(_lib.c_i64p)))
    return pos.tobytes() + off.tobytes()


def reweight(psi, possibility):
    """The observer's weight of a direct sample (observer.py:333-342):
    |psi(s)|^2 / p(s), zero where psi(s) is zero (the observer counts such a
    sample and adds no weight)."""
    psi = np.abs(np.asarray(psi, dtype=np.float64))
    p = np.asarray(possibility, dtype=np.float64)
    out = np.zeros(len(psi))
    nz = psi != 0
    out[nz] = psi[nz] ** 2 / p[nz]
    return out


def instability(weights):
    """observer.py:417-431: standard deviation over mean of the weights of
    all samples (zero weights included in the count)."""
    w = np.asarray(weights, dtype=np.float64)
    mean = w.sum() / len(w)
    var = (w * w).sum() / len(w) - mean * mean
    return (np.sqrt(var) if var > 0 else 0.0) / mean


def _contract_all(a, b, pairs):
    """safe_contract with ``contract_all_physics_edges`` (utility.py:
    340-350): every common name starting with ``P`` is contracted too."""
    extra = [(n, n) for n in a.names if n[0] == "P" and n in b.names]
    return Bt.contract(a, b, list(pairs) + extra)


class DirectSampler:
    """Direct sampling of a FermiU1 / FermiZ2 tV lattice (one orbit per
    site).  ``cut`` is the single-layer bond dimension of the sampled rows
    (the configuration's), ``cut2`` the double layer's
    (``direct_sampling_cut_dimension``)."""

    def __init__(self, lattice, cut=None, cut2=4, alpha=1.0, seed=0):
        self.lattice = lattice
        spec = lattice.spec
        self.L1, self.L2 = spec.L1, spec.L2
        self.cut = spec.cut if cut is None else int(cut)
        self.cut2 = int(cut2)
        self.alpha = float(alpha)
        self.gen = np.random.Generator(np.random.PCG64(seed))
        self._one = _one(spec, lattice.dtype)
        self._shr = [[(lattice.shrinker(l1, l2, 0), lattice.shrinker(l1, l2, 1)) for l2 in range(self.L2)]
                     for l1 in range(self.L1)]
        self.redrawn = 0
        self.refresh_all()

    # -- double layer (double_layer_auxiliaries.py:387-392, 415-496) -----------------
    def refresh_all(self):
        """Rebuilds the double-layer environments from the lattice's sites
        (``DirectSampling.refresh_all``, sampling.py:267-274)."""
        self._proj = [[(self.lattice.projected(l1, l2, 0), self.lattice.projected(l1, l2, 1))
                       for l2 in range(self.L2)] for l1 in range(self.L1)]
        ones = [Bt.TensorBatch.repeat(self._one, 1)] * self.L2
        down = {self.L1: ones}
        # (ket, bra) of every raw site, one item each: shared by the double
        # layer and by every draw's row before it is shrunk
        self._kb = [[None] * self.L2 for _ in range(self.L1)]
        for l1 in reversed(range(self.L1)):
            for l2 in range(self.L2):
                ket = Bt.TensorBatch.repeat(self.lattice.sites[l1][l2], 1)
                self._kb[l1][l2] = (ket, Bt.conjugate(ket))
            down[l1] = self._double_two_line("DULR", down[l1 + 1], self._kb[l1])
        self._down = down

    def _canonical(self, stage, left, right, left0, left1, right0, right1, cut, normalize):
        n = len(stage)
        for i in range(n - 1):
            q, r = Bt.qr(stage[i], "r", [x for x in (right1, right0) if x in stage[i].names], right, left)
            stage[i] = q
            stage[i + 1] = Bt.contract(stage[i + 1], r, [(left1, right1), (left0, right0)])
        for i in reversed(range(1, n)):
            u, s, v = Bt.svd(stage[i], [left], right, left, left, right, cut)
            if normalize:
                s = Bt.scale(s, 1.0 / np.array([s.tensor(k).norm_sum() for k in range(len(s))]))
            stage[i] = v
            stage[i - 1] = Bt.contract(Bt.contract(stage[i - 1], u, [(right, left)]), s, [(right, left)])
        return stage

    def _double_two_line(self, udlr, line_1, line_2):
        up, down, left, right = udlr
        dn, dc = down + "N", down + "C"
        l0, l1n, r0, r1n = left + "0", left + "1", right + "0", right + "1"
        st = [Bt.contract(line_1[i].rename({left: l1n, right: r1n}),
                          line_2[i][0].rename({left: l0, right: r0, down: dn}), [(dn, up)])
              for i in range(len(line_1))]
        st = self._canonical(st, left, right, l0, l1n, r0, r1n, self.cut2, True)
        st = [_contract_all(st[i].rename({left: l1n, right: r1n}),
                            line_2[i][1].rename({left: l0, right: r0, down: dc}), [("T", "T"), (dc, up)])
              for i in range(len(line_1))]
        return self._canonical(st, left, right, l0, l1n, r0, r1n, self.cut2, True)

    # -- single layer of the sampled rows (single_layer_auxiliaries.py:463-498) -------
    def _single_two_line(self, line_1, line_2):
        up, down, left, right = "UDLR"
        l1n, l2n, r1n, r2n = left + "1", left + "2", right + "1", right + "2"
        dl = [Bt.contract(line_1[i].rename({left: l1n, right: r1n}), line_2[i].rename({left: l2n, right: r2n}),
                          [(down, up)]) for i in range(len(line_1))]
        return self._canonical(dl, left, right, l2n, l1n, r2n, r1n, self.cut, False)

    # -- three-line auxiliaries (three_line_auxiliaries.py) ----------------------------
    def _three_two_line(self, lr, line_1, line_2):
        """:67-101; ``line_2`` = (n0, n1, t2, c1, c0) of a column."""
        left, right = lr
        rn, rc = right + "N", right + "C"
        cut = self.cut
        rn0, nr = Bt.qr(line_2[0], "q", [x for x in (right,) if x in line_2[0].names], "D", "U")
        n12 = Bt.contract(nr.rename({"D": "D2"}), line_1[0].rename({"D": "D1"}), [(left, right)])
        n12 = Bt.contract(n12, line_1[1].rename({"D": "D1"}), [("D1", "U")])
        n12 = Bt.contract(n12, line_2[1].rename({"D": "D2"}), [(right, left), ("D2", "U")])
        # the free names are those of line_2[0] here too (:78)
        rc0, cr = Bt.qr(line_2[4], "q", [x for x in (right,) if x in line_2[0].names], "D", "U")
        c12 = Bt.contract(cr.rename({"D": "D2"}), line_1[4].rename({"D": "D1"}), [(left, right)])
        c12 = Bt.contract(c12, line_1[3].rename({"D": "D1"}), [("D1", "U")])
        c12 = Bt.contract(c12, line_2[3].rename({"D": "D2"}), [(right, left), ("D2", "U")])
        big = Bt.contract(n12.rename({right: rn, "U": "UN"}), line_1[2].rename({"UC": "UC1"}), [("D1", "UN")])
        big = Bt.contract(big, line_2[2].rename({"UC": "UC2"}), [("D2", "UN"), (right, left)])
        big = _contract_all(big, c12.rename({right: rc, "U": "UC"}), [("UC1", "D1"), ("UC2", "D2"), ("T", "T")])
        u, s, v = Bt.svd(big, ["UN", rn], "D", "UN", "UN", "D", cut)
        rn1 = u.rename({"UN": "U", rn: right})
        big = Bt.contract(v, s, [("UN", "D")])
        u, s, v = Bt.svd(big, ["UC", rc], "D", "UC", "UC", "D", cut)
        rc1 = u.rename({"UC": "U", rc: right})
        big = Bt.contract(v, s, [("UC", "D")])
        return [rn0, rn1, big, rc1, rc0]

    @staticmethod
    def _hole(left, line, right):
        """:117-153."""
        n0, n1, t2, c1, c0 = line
        r = Bt.contract(left[0].rename({"D": "D1"}), n0.rename({"D": "D2"}), [("R", "L")])
        r = Bt.contract(r, right[0].rename({"D": "D3"}), [("R", "L")])
        r = Bt.contract(r, left[1].rename({"D": "D1"}), [("D1", "U")])
        r = Bt.contract(r, n1.rename({"D": "D2"}), [("D2", "U"), ("R", "L")])
        r = Bt.contract(r, right[1].rename({"D": "D3"}), [("D3", "U"), ("R", "L")])
        r = Bt.contract(r, left[2].rename({"UC": "U1"}), [("D1", "UN")])
        r = Bt.contract(r, t2.rename({"UC": "U2"}), [("D2", "UN"), ("R", "L")])
        r = Bt.contract(r, right[2].rename({"UC": "U3"}), [("D3", "UN"), ("R", "L")])
        r = Bt.contract(r, left[3].rename({"U": "U1"}), [("U1", "D")])
        r = Bt.contract(r, c1.rename({"U": "U2"}), [("U2", "D"), ("R", "L"), ("T", "T")])
        r = Bt.contract(r, right[3].rename({"U": "U3"}), [("U3", "D"), ("R", "L")])
        r = Bt.contract(r, left[4].rename({"U": "U1"}), [("U1", "D")])
        r = _contract_all(r, c0.rename({"U": "U2"}), [("U2", "D"), ("R", "L"), ("T", "T")])
        return Bt.contract(r, right[4].rename({"U": "U3"}), [("U3", "D"), ("R", "L")])

    @staticmethod
    def _layout(t, kind, swap):
        """For one site hole: the storage index and sign of each diagonal
        entry (index -1 and sign 0 where the contraction left the block
        empty: it weighs nothing) and the occupation each entry stands for."""
        edges = dict(zip(t.names, t.edges))
        eo, ei = edges["O0"], edges["I0"]
        blocks = {p: off for p, off, _ in t._shape.blocks()}
        idx, sgn, occ = [], [], []
        for q, d in eo.segments:
            pi, po = ei.position_by_symmetry(kind.neg(q)), eo.position_by_symmetry(q)
            off = blocks.get((pi, po) if t.names[0] == "I0" else (po, pi))
            s = -1.0 if swap and kind.parity(q) else 1.0
            for j in range(d):
                # entry (j, j) of a row-major d x d block
                idx.append(-1 if off is None else off + j * (d + 1))
                sgn.append(0.0 if off is None else s)
                occ.append(int(np.asarray(q).reshape(-1)[0]))
        return np.array(idx, dtype=np.int64), np.array(sgn), np.array(occ)

    def _rho(self, hole):
        """sampling.py:318-331 for every sample: the diagonal block of each
        segment of the ket edge ``O0``, in that edge's segment order, read
        in the order (I0, O0) (TAT's transpose of two odd edges changes the
        sign of their odd blocks), negative rounding clipped.  Returns
        (weights, occupation of each column).  Items are grouped by their
        block list (positions and offsets): one layout serves the items of a
        group; the weights of all samples come to the host in one copy."""
        if sorted(hole.names) != ["I0", "O0"]:
            raise RuntimeError(f"site hole has edges {hole.names}, expected I0 and O0")
        swap = hole.names == ["O0", "I0"]
        kind = hole.model.kind
        groups = {}
        for i in range(len(hole)):
            groups.setdefault(_block_key(int(hole.handles[i])), []).append(i)
        layouts, occ = {}, None
        for key, items in groups.items():
            idx, sgn, o = self._layout(hole.tensor(items[0]), kind, swap)
            if occ is None:
                occ = o
            elif not np.array_equal(occ, o):
                raise RuntimeError("site holes of one step differ in their physical segments")
            layouts[key] = (idx, sgn)
        rows = np.zeros((len(hole), len(occ)))
        hosts = {int(b): hole.bufs[b].double().cpu().numpy() for b in np.unique(hole.buf_idx)}
        for key, items in groups.items():
            idx, sgn = layouts[key]
            items = np.asarray(items)
            for b in np.unique(hole.buf_idx[items]):
                sel = items[hole.buf_idx[items] == b]
                host = hosts[int(b)]
                if np.any(idx >= hole.sizes[sel][:, None]):
                    raise RuntimeError("site hole layout exceeds the item's storage")
                at = hole.offs[sel][:, None] + np.maximum(idx, 0)[None, :]
                rows[sel] = np.where(idx[None, :] >= 0, host[at], 0.0) * sgn
        return np.maximum(rows, 0.0), occ

    # -- sampling (sampling.py:276-360) -----------------------------------------------------
    def _draw(self, B, uniforms, log=None):
        L1, L2 = self.L1, self.L2
        cfg = np.zeros((B, L1, L2), dtype=np.int8)
        poss = np.ones(B)
        alive = np.ones(B, dtype=bool)
        up = [Bt.TensorBatch.repeat(self._one, B)] * L2
        ones = [Bt.TensorBatch.repeat(self._one, B)] * 5
        k = 0
        for l1 in range(L1):
            n0 = up
            c0 = [Bt.conjugate(t) for t in n0]
            every = np.zeros(B, dtype=np.int64)
            n1 = [self._kb[l1][l2][0].take(every) for l2 in range(L2)]
            c1 = [self._kb[l1][l2][1].take(every) for l2 in range(L2)]
            t2 = [d.take(every) for d in self._down[l1 + 1]]
            right = {L2: ones}
            for c in reversed(range(L2)):
                right[c] = self._three_two_line("RL", right[c + 1], (n0[c], n1[c], t2[c], c1[c], c0[c]))
            left = ones
            for l2 in range(L2):
                ket = n1[l2].rename({n: "O" + n[1:] for n in n1[l2].names if n.startswith("P")})
                bra = c1[l2].rename({n: "I" + n[1:] for n in c1[l2].names if n.startswith("P")})
                rho, occ = self._rho(self._hole(left, (n0[l2], ket, t2[l2], bra, c0[l2]), right[l2 + 1]))
                total = rho.sum(axis=1)
                alive &= total > 0
                rho = rho ** self.alpha
                with np.errstate(invalid="ignore", divide="ignore"):
                    rho = rho / rho.sum(axis=1, keepdims=True)
                # _choice (:362-371): the first index where the running
                # subtraction of the weights goes negative, else the last
                run = uniforms[:, k][:, None] - np.cumsum(np.nan_to_num(rho), axis=1)
                neg = run < 0
                pick = np.where(neg.any(axis=1), neg.argmax(axis=1), rho.shape[1] - 1)
                k += 1
                poss *= np.where(alive, np.nan_to_num(rho[np.arange(B), pick]), 0.0)
                cfg[:, l1, l2] = occ[pick]
                if l2 == L2 - 1:
                    break  # nothing reads the row's environment past its last site
                shr = Bt.TensorBatch.select(self._shr[l1][l2], cfg[:, l1, l2]).rename({"P": "P0"})
                n1[l2] = Bt.contract(n1[l2], shr, [("P0", "Q")])
                c1[l2] = Bt.conjugate(n1[l2])
                left = self._three_two_line("LR", left, (n0[l2], n1[l2], t2[l2], c1[l2], c0[l2]))
            if l1 < L1 - 1:  # the boundary MPS of the drawn rows, for the next row
                row = [Bt.TensorBatch.select(self._proj[l1][l2], cfg[:, l1, l2]) for l2 in range(L2)]
                up = self._single_two_line(up, row)
            if log:
                log(f"row {l1} of {L1} drawn for {B} samples")
        return poss, cfg, alive

    def sample(self, count, uniforms=None, max_redraw=20, log=None, restrict=None):
        """``count`` independent draws: (possibility, configurations).  The
        uniforms (count x L1*L2, site by site in row-major order) default to
        this sampler's generator; draws that hit a site with no weight, and
        draws that ``restrict`` (configurations -> bool per draw, tetragono's
        ``restrict_subspace``, sampling.py:352-355) rejects, are drawn again
        from fresh uniforms, at most ``max_redraw`` times."""
        sites = self.L1 * self.L2
        if uniforms is None:
            uniforms = self.gen.random((count, sites))
        uniforms = np.asarray(uniforms, dtype=np.float64)
        if uniforms.shape != (count, sites):
            raise ValueError("uniforms must have shape (count, L1*L2)")
        if count == 0:
            return np.zeros(0), np.zeros((0, self.L1, self.L2), dtype=np.int8)

        def good(c, a):
            if restrict is None:
                return a
            return a & np.asarray(restrict(c), dtype=bool)

        poss, cfg, alive = self._draw(count, uniforms, log)
        todo = np.nonzero(~good(cfg, alive))[0]
        for _ in range(max_redraw):
            if not len(todo):
                break
            self.redrawn += len(todo)
            p2, c2, a2 = self._draw(len(todo), self.gen.random((len(todo), sites)), log)
            poss[todo], cfg[todo] = p2, c2
            todo = todo[~good(c2, a2)]
        if len(todo):
            raise RuntimeError(f"{len(todo)} direct samples found no accepted draw in {max_redraw} redraws")
        return poss, cfg
