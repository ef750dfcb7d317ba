"""The element table of a bond operator (tnsp_amd/elements.py) against the
oracle's restatement of tetragono's ``calculate_element``
(tetragono/tetragono/tensor_element.py:84-111), and the oracle sweep with a
table of several targets per key against the Metropolis rule on exact
amplitudes."""
import numpy as np
import pytest

from test_environment import configs
from test_sampler import small_lattice


def _mu_operator(t, V, mu, kind="FermiU1"):
    """tV plus mu n_1 (site 1's number, identity on site 0): the key (0, 1)
    gets two targets, the hop and the diagonal, while (1, 0) keeps one."""
    from oracle import sampling, tat

    K = tat.KINDS[kind]
    H = sampling.tv_hamiltonian(t, V, kind)
    _, _, _, N = sampling.fermi_operators(kind)
    EF = tat.Edge(K, [((0,), 1), ((1,), 1)], False)
    if kind == "FermiU1":
        ET = tat.Edge(K, [((0,), 1), ((-1,), 1)], True)
        odd_in = ((-1,), 0)
    else:
        ET = tat.Edge(K, [((0,), 1), ((1,), 1)], True)
        odd_in = ((1,), 0)
    ident = tat.Tensor(K, ["O0", "I0"], [EF, ET])
    ident[{"O0": ((0,), 0), "I0": ((0,), 0)}] = 1.0
    ident[{"O0": ((1,), 0), "I0": odd_in}] = 1.0
    M = sampling.rename_io(ident, [0]).contract(sampling.rename_io(N, [1]), set())
    return H + M * mu


def _host_table(H):
    from tnsp_amd.elements import element_table
    from tnsp_amd.shape import KINDS

    blocks = [(p, H.block(p)) for p in H.block_order]
    return element_table(KINDS[H.kind.name], H.names, H.edges, blocks)


@pytest.mark.parametrize("kind", ["FermiU1", "FermiZ2"])
@pytest.mark.parametrize("mu", [0.0, 0.5])
def test_element_table_matches_oracle_pool(kind, mu):
    """Same keys, same targets in the same order, same element values."""
    from oracle import sampling

    H = _mu_operator(1.0, 0.7, mu, kind) if mu else sampling.tv_hamiltonian(1.0, 0.7, kind)
    pool = sampling.element_pool(H)
    got = _host_table(H)
    assert list(got) == list(pool)
    for k in pool:
        assert list(got[k]) == list(pool[k])
        for o in pool[k]:
            assert got[k][o] == float(pool[k][o].data[0])
    if mu:
        assert len(got[(0, 1)]) == 2 and len(got[(1, 0)]) == 1


def test_element_table_block_order_is_lexicographic():
    """Blocks handed over in any order give tetragono's target order (blocks
    by segment positions, last edge fastest)."""
    from oracle import sampling
    from tnsp_amd.elements import element_table
    from tnsp_amd.shape import KINDS

    H = _mu_operator(1.0, 0.7, 0.5)
    blocks = [(p, H.block(p)) for p in reversed(H.block_order)]
    got = element_table(KINDS["FermiU1"], H.names, H.edges, blocks)
    pool = sampling.element_pool(H)
    assert [list(v) for v in got.values()] == [list(v) for v in pool.values()]


def test_tv_table_has_the_tv_structure():
    from oracle import sampling
    from tnsp_amd.elements import tv_table

    for t, V in [(1.0, 1.0), (1.0, 0.0), (0.0, 1.0)]:
        pool = sampling.element_pool(sampling.tv_hamiltonian(t, V)) if (t or V) else {}
        got = tv_table(t, V)
        assert {k: list(v) for k, v in got.items()} == {k: list(v) for k, v in pool.items()}


def test_moves_draw_targets_and_counts():
    from tnsp_amd.elements import Moves

    mv = Moves({(0, 1): {(0, 1): 0.5, (1, 0): -1.0}, (1, 0): {(0, 1): -1.0}, (1, 1): {(1, 1): 1.2}})
    assert mv.multi
    na = np.array([0, 0, 1, 1, 0], dtype=np.int8)
    nb = np.array([1, 1, 0, 1, 0], dtype=np.int8)
    v = np.array([0.49, 0.5, 0.99, 0.3, 0.1])
    active, ta, tb, n, ns = mv.draw(na, nb, v)
    assert active.tolist() == [True, True, True, True, False]
    assert list(zip(ta[:4], tb[:4])) == [(0, 1), (1, 0), (0, 1), (1, 1)]
    assert n[:4].tolist() == [2, 2, 1, 1]
    assert ns[:4].tolist() == [2, 1, 2, 1]
    with pytest.raises(ValueError):
        Moves({(0, 1): {(1, 0): 1.0}})  # the target has no elements: n' = 0


def test_sweep_uncut_several_targets_follows_exact_metropolis():
    """Oracle sweep with the tV + mu n_1 table, uncut (3x3): each bond draws
    its target by floor(u' n), the diagonal reads the amplitude itself, and
    the acceptance is |psi(s')/psi(s)|^2 n / n' on exact amplitudes."""
    from oracle import environment as E, peps, sampling

    L1, L2, N = 3, 3, 4
    _, _, sites = small_lattice(L1, L2, N, seed=8)
    cut = 10_000
    table = sampling.element_pool(_mu_operator(1.0, 0.7, 0.5))
    (cfg,) = configs(L1, L2, N, 1, seed=9)
    bonds = sampling.bond_order(L1, L2)
    rng = np.random.default_rng(10)
    u, tu = rng.random(len(bonds)), rng.random(len(bonds))
    final, trace, last = E.sweep(sites, cfg, cut, u, table=table, target_uniforms=tu)
    cur = cfg.copy()
    psi = abs(peps.psi(sites, cur, cut)[0])
    factors = set()
    for k, (a, b) in enumerate(bonds):
        before, wss, acc = trace[k]
        assert abs(before - psi) <= 1e-10 * psi
        key = (int(cur[a]), int(cur[b]))
        if key not in table:
            assert wss is None
            continue
        outs = list(table[key])
        target = outs[min(int(tu[k] * len(outs)), len(outs) - 1)]
        new = cur.copy()
        new[a], new[b] = target
        p_new = abs(peps.psi(sites, new, cut)[0])
        assert abs(wss - p_new) <= 1e-10 * p_new
        f = len(outs) / len(table[target])
        factors.add(f)
        assert acc == (u[k] < (p_new / psi) ** 2 * f)
        if acc:
            cur, psi = new, p_new
    assert np.array_equal(final, cur)
    assert abs(last - psi) <= 1e-10 * psi
    # the draws exercised n / n' of 2 and of 1/2
    assert {2.0, 0.5} <= factors
