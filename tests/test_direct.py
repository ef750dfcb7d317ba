"""Oracle of tetragono's direct sampling (oracle/direct.py).

Pins, on small lattices whose every configuration can be listed:
* with no truncation (single- and double-layer cuts above every bond
  dimension) and alpha = 1, the probability of drawing each configuration is
  |psi(s)|^2 / sum_s' |psi(s')|^2, with psi from the pinned amplitude
  (oracle.peps.psi); FermiU1 at fixed particle number and FermiZ2 at even
  parity -- a dropped conjugate, a wrong fermion sign in the double layer or a
  wrong block of the site hole breaks this;
* the double layer alone (uncut, unnormalised) contracts to sum |psi|^2;
* at truncating cuts the draw is still a normalised sequential distribution:
  the probabilities of all configurations add up to one, and a draw driven by
  uniforms reports the probability that the forced draw of its configuration
  gives;
* ``choice`` is tetragono's running subtraction (first index where it goes
  negative, the last index when rounding leaves it positive);
* the observer's sample weight and instability (tnsp_amd.direct) against
  their closed forms; the GPU sampler's host readout of the site holes
  against the oracle's, and its redraw loop with scripted draws.
"""
import itertools

import numpy as np
import pytest

from test_sampler import small_lattice


def z2_lattice(L1, L2, seed):
    from oracle import peps
    from tnsp_amd import peps as pp

    spec = pp.LatticeSpec("FermiZ2", L1, L2, (2, 1), 2, 1000, seed=seed)
    storages = pp.random_storages(spec)
    sd = {"kind": "FermiZ2", "L1": L1, "L2": L2, "N": 2, "segments": list(spec.segments)}
    return spec, storages, peps.build_sites(sd, storages)


def all_configs(L1, L2, kind, N):
    for flat in itertools.product([0, 1], repeat=L1 * L2):
        if (kind == "FermiU1" and sum(flat) == N) or (kind == "FermiZ2" and sum(flat) % 2 == 0):
            yield np.array(flat, dtype=np.int64).reshape(L1, L2)


@pytest.mark.parametrize("kind", ["FermiU1", "FermiZ2"])
def test_uncut_probability_is_born_rule(kind):
    from oracle import direct as Dr, peps

    L1, L2, N = 2, 3, 3
    if kind == "FermiU1":
        _, _, sites = small_lattice(L1, L2, N, seed=3)
    else:
        _, _, sites = z2_lattice(L1, L2, seed=5)
    big = 10_000
    down = Dr.double_layer_down_to_up(sites, big)
    p, w = [], []
    for cfg in all_configs(L1, L2, kind, N):
        pr, _, rhos = Dr.direct_sample(sites, big, big, forced=cfg, down=down)
        p.append(pr)
        w.append(peps.psi(sites, cfg, big)[0] ** 2)
        assert all(abs(r.sum() - 1) < 1e-12 for r in rhos)
    p, w = np.array(p), np.array(w)
    assert np.all(w > 0)
    assert np.allclose(p, w / w.sum(), rtol=1e-9, atol=1e-15)


def test_truncated_draw_is_normalised_and_consistent():
    from oracle import direct as Dr

    L1, L2 = 2, 3
    _, _, sites = z2_lattice(L1, L2, seed=7)
    cut, cut2 = 2, 2
    down = Dr.double_layer_down_to_up(sites, cut2)
    total = sum(Dr.direct_sample(sites, cut, cut2, forced=cfg, down=down)[0]
                for cfg in all_configs(L1, L2, "FermiZ2", 0))
    assert abs(total - 1) < 1e-12
    rng = np.random.default_rng(11)
    for _ in range(3):
        u = rng.random(L1 * L2)
        p, cfg, _ = Dr.direct_sample(sites, cut, cut2, uniforms=u, down=down)
        pf, cf, _ = Dr.direct_sample(sites, cut, cut2, forced=cfg, down=down)
        assert np.array_equal(cfg, cf)
        assert p == pytest.approx(pf, rel=1e-13)


def test_choice():
    from oracle.direct import choice

    rho = np.array([0.25, 0.5, 0.25])
    assert choice(0.1, rho) == 0
    assert choice(0.3, rho) == 1
    assert choice(0.76, rho) == 2
    assert choice(1.0, rho) == 2  # rounding: never beyond the last index


def test_reweight_and_instability():
    """observer.py:333-342, 417-431: weight |psi|^2 / p, none for psi = 0
    (counted all the same); instability = population standard deviation
    over the mean of all weights."""
    from tnsp_amd.direct import instability, reweight

    w = reweight([0.0, 2.0, -1.0], [0.5, 0.25, 0.5])
    assert np.array_equal(w, [0.0, 16.0, 2.0])
    assert instability(w) == pytest.approx(np.std(w) / np.mean(w), rel=1e-12)
    assert instability([3.0, 3.0, 3.0]) == 0.0


@pytest.mark.parametrize("kind_name,order", [("FermiU1", ["I0", "O0"]), ("FermiU1", ["O0", "I0"]),
                                             ("FermiZ2", ["O0", "I0"])])
def test_site_weight_readout(kind_name, order):
    """The GPU sampler's host readout of the site holes (DirectSampler._rho)
    gives the oracle's weights (edge rename, transpose to (I, O), diagonal of
    each block, clipped) for holes stored in either edge order; blocks of two
    states check the diagonal indexing, the (O0, I0) order the sign of odd
    blocks.  Storage is handed over as host tensors: only the block layout of
    the native shape planner and the readout run."""
    from types import SimpleNamespace

    import torch

    from oracle import direct as Dr, tat
    from tnsp_amd.direct import DirectSampler
    from tnsp_amd.shape import KINDS, Edge, Shape

    segs = [(0, 2), (1, 3)]
    kind, okind = KINDS[kind_name], tat.KINDS[kind_name]
    phys = Edge(kind, segs, kind_name == "FermiU1")
    edges = {"O0": phys, "I0": phys.conjugate()}
    shape = Shape.create(kind, order, [edges[n] for n in order])
    oedges = {"O0": tat.Edge(okind, segs, kind_name == "FermiU1")}
    oedges["I0"] = oedges["O0"].conjugate()
    rng = np.random.default_rng(3)
    n, size = 3, shape.size
    data = rng.normal(size=(n, size))
    # item 1 has its own handle of the same structure: it shares the layout
    shapes = [shape, Shape.create(kind, order, [edges[nm] for nm in order]), shape]

    class Hole:
        names = list(order)
        model = SimpleNamespace(kind=kind)
        sizes = np.full(n, size)
        bufs = [torch.from_numpy(data.reshape(-1).copy())]
        buf_idx = np.zeros(n, dtype=np.int64)
        offs = np.arange(n) * size

        @property
        def handles(self):
            return np.array([s.handle.value for s in shapes], dtype=np.uint64)

        def __len__(self):
            return n

        def tensor(self, i):
            return SimpleNamespace(names=shapes[i].names, edges=shapes[i].edges, _shape=shapes[i])

    hole = Hole()
    rho, occ = DirectSampler._rho(DirectSampler.__new__(DirectSampler), hole)
    assert list(occ) == [0, 0, 1, 1, 1]
    for i in range(n):
        t = tat.Tensor(okind, list(order), [oedges[nm] for nm in order])
        t.data[:] = data[i]
        want, _ = Dr.site_rho(t)
        np.testing.assert_array_equal(rho[i], want)
    # an item of the same size whose blocks lie elsewhere (segments in the
    # other order) is not read through item 0's layout
    rev = Edge(kind, segs[::-1], kind_name == "FermiU1")
    redges = {"O0": rev, "I0": rev.conjugate()}
    shapes[1] = Shape.create(kind, order, [redges[nm] for nm in order])
    assert shapes[1].size == size
    with pytest.raises(RuntimeError):
        DirectSampler._rho(DirectSampler.__new__(DirectSampler), hole)


def _stub_sampler(script):
    """A DirectSampler whose draws come from ``script`` (a list of
    (possibility, configuration, alive) per call); no device is touched."""
    from tnsp_amd.direct import DirectSampler

    smp = DirectSampler.__new__(DirectSampler)
    smp.L1, smp.L2 = 1, 2
    smp.gen = np.random.Generator(np.random.PCG64(0))
    smp.redrawn = 0
    calls = []

    def draw(B, uniforms, log=None):
        assert uniforms.shape == (B, 2)
        p, c, a = script[len(calls)]
        calls.append(B)
        return np.array(p, dtype=float)[:B], np.array(c, dtype=np.int8)[:B], np.array(a)[:B]

    smp._draw = draw
    return smp, calls


def test_sample_redraws_dead_and_restricted_draws():
    """Draws with no weight at some site and draws the restriction rejects
    are drawn again (only those), as DirectSampling.__call__ restarts them;
    the count of redraws is kept."""
    smp, calls = _stub_sampler([
        ([0.5, 0.0, 0.25], [[[1, 0]], [[0, 1]], [[1, 1]]], [True, False, True]),
        ([0.5, 0.125], [[[0, 1]], [[1, 0]]], [True, True]),
    ])
    poss, cfg = smp.sample(3, restrict=lambda c: c.reshape(len(c), -1).sum(axis=1) == 1)
    assert calls == [3, 2]
    assert smp.redrawn == 2
    assert np.array_equal(poss, [0.5, 0.5, 0.125])
    assert np.array_equal(cfg.reshape(3, -1), [[1, 0], [0, 1], [1, 0]])


def test_sample_gives_up_after_max_redraw():
    smp, calls = _stub_sampler([([0.0], [[[0, 0]]], [False])] * 3)
    with pytest.raises(RuntimeError, match="no accepted draw"):
        smp.sample(1, max_redraw=2)
    assert calls == [1, 1, 1]
    assert len(smp.sample(0)[0]) == 0


def test_double_layer_norm_uncut():
    """The double-layer boundary MPS of every row (no truncation, no
    normalisation) contracted along the row is <psi|psi> = sum_s psi(s)^2:
    the bra layer's conjugation and the traced physical and T edges carry
    the right fermion signs."""
    from oracle import direct as Dr, peps

    L1, L2, N = 2, 3, 3
    _, _, sites = small_lattice(L1, L2, N, seed=13)
    big = 10_000
    mps = [peps.one(sites[0][0].kind)] * L2
    for l1 in reversed(range(L1)):
        row = [(s, s.conjugate()) for s in sites[l1]]
        mps = Dr.double_layer_two_line("DULR", mps, row, big, normalize=False)
    acc = mps[0]
    for t in mps[1:]:
        acc = acc.contract(t, {("R", "L")})
    assert acc.names == []
    z = sum(peps.psi(sites, cfg, big)[0] ** 2 for cfg in all_configs(L1, L2, "FermiU1", N))
    assert float(acc.data[0]) == pytest.approx(z, rel=1e-10)
