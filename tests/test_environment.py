"""Oracle of the cached-environment ``replace`` (oracle/environment.py).

Pins, on small FermiU1 lattices:
* ``hole()`` through the row environments is the pinned amplitude
  (oracle.peps.amplitude, itself checked against TAT's golden vectors), at a
  truncating cut;
* without truncation, ``replace`` of every horizontal and vertical bond
  equals the amplitude of the changed configuration, sign included once both
  one-element tensors are transposed to one edge order;
* at a truncating cut, ``replace`` with the current sites is the same number
  for every bond of a row (a horizontal strip contracted exactly between two
  boundary MPS) and of a column; for row 0 it is ``hole(())``;
* the sweep with an uncut boundary MPS makes the same decisions as the
  Metropolis rule on exact amplitudes.
"""
import numpy as np
import pytest

from test_sampler import small_lattice


def configs(L1, L2, N, count, seed):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        flat = np.zeros(L1 * L2, dtype=np.int64)
        flat[rng.choice(L1 * L2, N, replace=False)] = 1
        out.append(flat.reshape(L1, L2))
    return out


def signed(t, names):
    return float(t.transpose(names).data[0])


@pytest.mark.parametrize("L1,L2,N,cut", [(3, 3, 4, 3), (4, 4, 8, 4)])
def test_hole_is_amplitude(L1, L2, N, cut):
    from oracle import environment as E, peps

    _, _, sites = small_lattice(L1, L2, N, seed=1)
    for cfg in configs(L1, L2, N, 2, seed=2):
        lat = peps.project(sites, cfg)
        a = peps.amplitude(lat, cut)
        h = E.Environments(lat, cut).hole()
        assert h.names == a.names
        assert np.allclose(h.data, a.data, rtol=1e-12, atol=0)


@pytest.mark.parametrize("L1,L2,N", [(3, 3, 4), (3, 4, 5)])
def test_replace_uncut_equals_amplitude(L1, L2, N):
    from oracle import environment as E, peps, sampling

    _, _, sites = small_lattice(L1, L2, N, seed=4)
    cut = 10_000
    (cfg,) = configs(L1, L2, N, 1, seed=5)
    checked = 0
    for a, b in sampling.bond_order(L1, L2):
        if cfg[a] == cfg[b]:
            continue
        new = cfg.copy()
        new[a], new[b] = cfg[b], cfg[a]
        want = peps.amplitude(peps.project(sites, new), cut)
        got = E.replace(sites, cfg, cut, (a, b), (int(new[a]), int(new[b])))
        assert sorted(got.names) == sorted(want.names)
        w = float(want.data[0])
        assert abs(signed(got, want.names) - w) <= 1e-10 * abs(w), (a, b)
        checked += 1
    assert checked >= 4


def test_replace_unchanged_is_row_and_column_invariant():
    from oracle import environment as E, peps

    L1, L2, N, cut = 4, 4, 8, 3
    _, _, sites = small_lattice(L1, L2, N, seed=6)
    (cfg,) = configs(L1, L2, N, 1, seed=7)
    lat = peps.project(sites, cfg)
    env = E.Environments(lat, cut)
    hole = E.value(env.hole())
    for l1 in range(L1):
        row = [abs(E.value(env.replace_horizontal(l1, l2, lat[l1][l2], lat[l1][l2 + 1]))) for l2 in range(L2 - 1)]
        assert np.allclose(row, row[0], rtol=1e-11, atol=0), l1
        if l1 == 0:
            assert abs(row[0] - abs(hole)) <= 1e-12 * abs(hole)
    cols = []
    for l2 in range(L2):
        col = [abs(E.value(env.replace_vertical(l1, l2, lat[l1][l2], lat[l1 + 1][l2]))) for l1 in range(L1 - 1)]
        assert np.allclose(col, col[0], rtol=1e-11, atol=0), l2
        cols.append(col[0])
    # at this cut the column contractions approximate the amplitude differently
    # from the row ones: a real truncation, so the invariances above are not
    # the uncut identity in disguise
    assert not np.allclose(cols, abs(hole), rtol=1e-3)


def test_sweep_uncut_follows_exact_metropolis():
    from oracle import environment as E, peps, sampling

    L1, L2, N = 3, 3, 4
    _, _, sites = small_lattice(L1, L2, N, seed=8)
    cut = 10_000
    (cfg,) = configs(L1, L2, N, 1, seed=9)
    bonds = sampling.bond_order(L1, L2)
    u = np.random.default_rng(10).random(len(bonds))
    final, trace, last = E.sweep(sites, cfg, cut, u)
    cur = cfg.copy()
    psi = abs(peps.psi(sites, cur, cut)[0])
    for k, (a, b) in enumerate(bonds):
        before, wss, acc = trace[k]
        assert abs(before - psi) <= 1e-10 * psi
        if cur[a] == cur[b]:
            if cur[a] == 1:
                assert abs(wss - psi) <= 1e-10 * psi
            continue
        new = cur.copy()
        new[a], new[b] = cur[b], cur[a]
        p_new = abs(peps.psi(sites, new, cut)[0])
        assert abs(wss - p_new) <= 1e-10 * p_new
        assert acc == (u[k] < (p_new / psi) ** 2)
        if acc:
            cur, psi = new, p_new
    assert np.array_equal(final, cur)
    assert abs(last - psi) <= 1e-10 * psi


def test_observer_energy_uncut_equals_scratch_energy():
    """The observer through replace equals the observer with every psi(s')
    from scratch when nothing is truncated (3x3, two configurations)."""
    from oracle import environment as E, sampling

    L1, L2, N = 3, 3, 4
    _, _, sites = small_lattice(L1, L2, N, seed=12)
    H = sampling.tv_hamiltonian(1.0, 1.0)
    for cfg in configs(L1, L2, N, 2, seed=13):
        got = E.local_energy(sites, cfg, 10_000, H)
        want = sampling.local_energy(sites, cfg, 10_000, H)
        assert abs(got - want) <= 1e-10 * max(1.0, abs(want))


def _pair(hole, site):
    """Full contraction of a hole with its site tensor (one number)."""
    t = hole.contract(site, {(n, n) for n in hole.names if n in site.names})
    return float(t.data[0])


@pytest.mark.parametrize("L1,L2,N,cut,rows", [(3, 3, 4, 10_000, None), (4, 4, 8, 3, [0])])
def test_holes_contract_to_one(L1, L2, N, cut, rows):
    """psi is linear in every site tensor, so <hole_x, x> = psi / psi = 1:
    for every site without truncation; at a truncating cut for the sites of
    row 0, whose environments are those of hole(()) itself."""
    from oracle import environment as E

    from oracle import peps

    _, _, sites = small_lattice(L1, L2, N, seed=14)
    cfg = next(c for c in configs(L1, L2, N, 8, seed=15) if peps.psi(sites, c, cut)[0] != 0)
    hs = E.holes(sites, cfg, cut)
    for l1 in (range(L1) if rows is None else rows):
        for l2 in range(L2):
            assert abs(_pair(hs[l1][l2], sites[l1][l2]) - 1.0) <= 1e-10, (l1, l2)
    if rows is not None:
        # the other rows use different truncations: not 1, which shows the cut bites
        assert any(abs(_pair(hs[l1][0], sites[l1][0]) - 1.0) > 1e-6 for l1 in range(1, L1))


def test_energy_gradient_matches_finite_differences():
    """The observer's gradient 2 conj(<E Delta> - E <Delta>) with exact
    enumeration weights |psi|^2 is the derivative of the variational energy
    E(T) = sum_s psi^2 E_loc / sum_s psi^2 with respect to each site tensor's
    storage (after bringing the gradient to the site's edge order); checked
    by central differences on a 2x3 lattice at half filling, no truncation."""
    import itertools

    from oracle import environment as E, peps, sampling

    from test_sampler import OracleAmplitudes
    from tnsp_amd.sampler import TVSampler

    L1, L2, N, cut = 2, 3, 3, 10_000
    spec, _, sites = small_lattice(L1, L2, N, seed=16)
    H = sampling.tv_hamiltonian(1.0, 1.0)
    cfgs = []
    for occ in itertools.combinations(range(L1 * L2), N):
        c = np.zeros(L1 * L2, dtype=np.int64)
        c[list(occ)] = 1
        cfgs.append(c.reshape(L1, L2))
    cfgs = [c for c in cfgs if peps.psi(sites, c, cut)[0] != 0]

    def energy(st):
        # the closed-form local energy (pinned against the observer contraction
        # in test_sampler) over every configuration, weights psi^2
        amps = OracleAmplitudes(spec, st)
        smp = TVSampler(amps, np.stack(cfgs))
        ps = smp.psi
        return float(np.sum(ps ** 2 * smp.local_energy()) / np.sum(ps ** 2)), ps

    e0, ps = energy(sites)
    grad, eg = E.energy_gradient(sites, cfgs, cut, H, reweight=ps ** 2)
    assert abs(eg - e0) <= 1e-10 * abs(e0)
    rng = np.random.default_rng(17)
    eps = 1e-5
    for l1, l2 in ((0, 0), (1, 1), (0, 2)):
        g = grad[l1][l2].transpose(sites[l1][l2].names)
        assert [e.segments for e in g.edges] == [e.segments for e in sites[l1][l2].edges]
        for i in rng.choice(sites[l1][l2].size, 2, replace=False):
            plus = [[t.copy() for t in row] for row in sites]
            minus = [[t.copy() for t in row] for row in sites]
            plus[l1][l2].data[i] += eps
            minus[l1][l2].data[i] -= eps
            fd = (energy(plus)[0] - energy(minus)[0]) / (2 * eps)
            assert abs(g.data[i] - fd) <= 1e-6 * max(1.0, abs(fd)), (l1, l2, i, g.data[i], fd)


def test_natural_gradient_is_min_norm_least_squares():
    """Conjugate gradient from zero on the normal equations converges to the
    minimum-norm least-squares solution: 2 x = 2 pinv(D) E (numpy lstsq),
    3x3 lattice, 6 configurations, truncating cut."""
    from oracle import environment as E, peps, sampling

    L1, L2, N, cut = 3, 3, 4, 3
    _, _, sites = small_lattice(L1, L2, N, seed=18)
    cfgs = [c for c in configs(L1, L2, N, 20, seed=19) if peps.psi(sites, c, cut)[0] != 0][:6]
    H = sampling.tv_hamiltonian(1.0, 1.0)
    # D has rank <= 5 (6 centred rows): the residual vanishes within 5 steps;
    # the error rule stops there (iterating on a zero residual divides by ~0,
    # in the reference as here)
    _, _, steps, D, Ev, x = E.natural_gradient(sites, cfgs, cut, H, step=50, error=1e-10)
    assert steps <= 6
    want = 2 * np.linalg.lstsq(D, Ev, rcond=None)[0]
    assert np.max(np.abs(x - want)) <= 1e-6 * np.max(np.abs(want))
