"""GPU sweep sampling with cached environments (tnsp_amd/environment.py)
against the oracle sweep (oracle/environment.py) on the same uniforms.

At a truncating cut every replace value depends on which environments were
used, so equal decisions and equal carried |psi| after a sweep show that the
GPU sampler builds the same environments as tetragono's lazy graph and
contracts them in the same way (1e-10 relative on |psi|; decisions exact)."""
import numpy as np
import pytest

from test_environment import configs
from test_sampler import small_lattice

pytestmark = pytest.mark.gpu


def _run(L1, L2, N, cut, chains, seed, diagonal, sweeps=1, lanes=1, speculate=False, table=None):
    from oracle import environment as E
    from tnsp_amd import peps as pp
    from tnsp_amd.engine import AmplitudeEngine
    from tnsp_amd.environment import EnvironmentSampler

    spec, storages, sites = small_lattice(L1, L2, N, seed=seed)
    eng = AmplitudeEngine(pp.Lattice(spec, storages), chunk=64, lanes=1)
    start = np.stack(configs(L1, L2, N, chains, seed=seed + 1))
    smp = EnvironmentSampler(eng, start, seed=seed + 2, cut=cut, diagonal=diagonal, lanes=lanes, speculate=speculate,
                             hamiltonian=table)
    gen = np.random.Generator(np.random.PCG64(seed + 2))
    # target draws: the sampler's second generator, seeded with (seed, 1)
    gen_t = np.random.Generator(np.random.PCG64(np.random.SeedSequence((seed + 2, 1))))
    nb = len(smp.bonds)
    cur = start.astype(np.int64)
    for _ in range(sweeps):
        smp.sweep()
        u = np.stack([gen.random(chains) for _ in range(nb)])  # (bonds, chains), the sampler's draw order
        tu = gen_t.random((nb, chains)) if table is not None and smp.moves.multi else None
        want, last = [], []
        for b in range(chains):
            if table is not None:
                cfg, trace, ws = E.sweep(sites, cur[b], cut, u[:, b], table=table,
                                         target_uniforms=None if tu is None else tu[:, b])
            else:
                cfg, trace, ws = E.sweep(sites, cur[b], cut, u[:, b], V=1.0 if diagonal == "hole" else 0.0)
            want.append(cfg)
            last.append(ws)
        cur = np.stack(want)
        assert np.array_equal(smp.configs, cur)
        np.testing.assert_allclose(smp.psi, last, rtol=1e-10, atol=0)
    return smp


def test_environment_sweep_matches_oracle_truncated():
    """4x4, N = 8, D = 4, boundary cut 4 (truncating), doubly occupied bonds
    refresh psi through hole(())."""
    smp = _run(4, 4, 8, 4, chains=3, seed=20, diagonal="hole")
    assert smp.accepted > 0 and smp.refreshed > 0


def test_environment_sweep_speculative_matches_oracle():
    """The same with the sweep run speculatively and replayed."""
    smp = _run(4, 4, 8, 4, chains=3, seed=20, diagonal="hole", speculate=True)
    assert smp.accepted > 0 and smp.refreshed > 0


def test_environment_sweep_speculation_sweeps_again():
    """Six chains at boundary cut 4: hole(()) and the carried psi differ
    enough that one chain's speculative decisions disagree with the replay
    (the oracle shows chain trajectories that differ between the two
    rules); that chain is swept again step by step, and every chain still
    takes the oracle's decisions and ends with its psi."""
    smp = _run(4, 4, 8, 4, chains=6, seed=20, diagonal="hole", speculate=True)
    assert smp.respun > 0 and smp.refreshed > 0


def test_environment_sweep_element_table_with_several_targets():
    """A bond operator whose table has two targets for (0, 1) (tV plus
    mu n_1: the hop and the diagonal) and one for (1, 0): targets drawn from
    the second uniform, acceptance with n / n' = 2 and 1/2, diagonal targets
    through hole(()); the sampler takes the oracle's decisions at a
    truncating cut, speculative and step by step."""
    from oracle import sampling
    from test_elements import _mu_operator

    table = {k: {o: float(v.data[0]) for o, v in d.items()} for k, d in
             sampling.element_pool(_mu_operator(1.0, 0.7, 0.5)).items()}
    smp = _run(4, 4, 8, 4, chains=4, seed=21, diagonal="hole", table=table)
    assert smp.accepted > 0 and smp.refreshed > 0
    smp = _run(4, 4, 8, 4, chains=4, seed=21, diagonal="hole", table=table, speculate=True)
    assert smp.accepted > 0


def test_tensor_element_of_a_gpu_operator():
    """tensor_element of a drop-in Tensor equals the table of the same
    operator's host blocks (and the oracle's pool)."""
    from oracle import sampling
    from helpers import to_gpu
    from test_elements import _host_table, _mu_operator
    from tnsp_amd.elements import tensor_element

    H = _mu_operator(1.0, 0.7, 0.5)
    got = tensor_element(to_gpu(H))
    want = _host_table(H)
    assert list(got) == list(want)
    assert all(list(got[k].items()) == list(want[k].items()) for k in want)
    assert list(got) == list(sampling.element_pool(H))


def test_environment_sweep_lanes_match_oracle():
    """Chains split over two lanes (own host thread and stream each) take
    the same decisions as the oracle, doubly occupied bonds refreshed."""
    smp = _run(4, 4, 8, 4, chains=3, seed=20, diagonal="hole", lanes=2)
    assert smp.accepted > 0 and smp.refreshed > 0


def test_environment_sweep_two_sweeps_uncut():
    """3x4, two sweeps without truncation; diagonal bonds skipped (the oracle
    with V = 0 has no diagonal element either)."""
    smp = _run(3, 4, 5, 10_000, chains=3, seed=30, diagonal="keep", sweeps=2)
    assert smp.accepted > 0


def test_environment_observer_energy_matches_oracle_truncated():
    """Local energy through replace over cached environments (tetragono's
    observer) against the oracle observer, 4x4 at a truncating cut."""
    from oracle import environment as E, sampling
    from tnsp_amd import peps as pp
    from tnsp_amd.engine import AmplitudeEngine
    from tnsp_amd.environment import EnvironmentSampler

    L1, L2, N, cut = 4, 4, 8, 4
    spec, storages, sites = small_lattice(L1, L2, N, seed=40)
    eng = AmplitudeEngine(pp.Lattice(spec, storages), chunk=64, lanes=1)
    start = np.stack(configs(L1, L2, N, 3, seed=41))
    smp = EnvironmentSampler(eng, start, cut=cut)
    got = smp.local_energy()
    H = sampling.tv_hamiltonian(1.0, 1.0)
    want = [E.local_energy(sites, c, cut, H) for c in start]
    np.testing.assert_allclose(got, want, rtol=1e-9, atol=1e-9)
    smp.lanes = 2
    np.testing.assert_allclose(smp.local_energy(), want, rtol=1e-9, atol=1e-9)


def test_environment_observer_energy_uncut_equals_engine_energy():
    from tnsp_amd import peps as pp
    from tnsp_amd.engine import AmplitudeEngine
    from tnsp_amd.environment import EnvironmentSampler
    from tnsp_amd.sampler import TVSampler

    L1, L2, N = 3, 4, 5
    spec, storages, _ = small_lattice(L1, L2, N, seed=42)
    eng = AmplitudeEngine(pp.Lattice(spec, storages), chunk=64, lanes=1)
    start = np.stack(configs(L1, L2, N, 8, seed=43))
    smp = EnvironmentSampler(eng, start, cut=10_000)
    got = smp.local_energy()
    want = TVSampler(eng, start).local_energy()
    np.testing.assert_allclose(got, want, rtol=1e-9, atol=1e-9)


@pytest.mark.parametrize("lanes", [1, 2])
def test_environment_holes_match_oracle_truncated(lanes):
    """Configuration.holes of every chain from the batched row environments
    against the oracle, 4x4 at a truncating cut: same structure, values to
    1e-10 relative."""
    from oracle import environment as E, peps as opeps
    from test_gpu_ops import rel, same_structure
    from tnsp_amd import peps as pp
    from tnsp_amd.engine import AmplitudeEngine
    from tnsp_amd.environment import EnvironmentSampler

    L1, L2, N, cut = 4, 4, 8, 3
    spec, storages, sites = small_lattice(L1, L2, N, seed=44)
    eng = AmplitudeEngine(pp.Lattice(spec, storages), chunk=64, lanes=1)
    cand = configs(L1, L2, N, 12, seed=45)
    start = np.stack([c for c in cand if opeps.psi(sites, c, cut)[0] != 0][:3])
    smp = EnvironmentSampler(eng, start, cut=cut, lanes=lanes)
    holes, psi = smp.holes()
    for b, cfg in enumerate(start):
        want = E.holes(sites, cfg, cut)
        assert abs(psi[b] - E.value(E.Environments(opeps.project(sites, cfg), cut).hole())) <= 1e-10 * abs(psi[b])
        for l1 in range(L1):
            for l2 in range(L2):
                g, o = holes[l1][l2].tensor(b), want[l1][l2]
                same_structure(g, o)
                assert rel(np.asarray(g.storage), o.data) < 1e-10, (b, l1, l2)


def test_environment_energy_gradient_matches_oracle():
    """The observer's energy gradient over chains (holes, replace energies,
    trivial-metric conjugate) against the oracle, 4x4 at a truncating cut,
    with non-uniform weights."""
    from oracle import environment as E, peps as opeps, sampling
    from test_gpu_ops import rel, same_structure
    from tnsp_amd import peps as pp
    from tnsp_amd.engine import AmplitudeEngine
    from tnsp_amd.environment import EnvironmentSampler

    L1, L2, N, cut = 4, 4, 8, 3
    spec, storages, sites = small_lattice(L1, L2, N, seed=46)
    eng = AmplitudeEngine(pp.Lattice(spec, storages), chunk=64, lanes=1)
    cand = configs(L1, L2, N, 12, seed=47)
    start = np.stack([c for c in cand if opeps.psi(sites, c, cut)[0] != 0][:3])
    w = np.array([1.0, 0.5, 2.0])
    smp = EnvironmentSampler(eng, start, cut=cut, lanes=2)
    grad, energy = smp.energy_gradient(reweight=w)
    want, e_want = E.energy_gradient(sites, start, cut, sampling.tv_hamiltonian(1.0, 1.0), reweight=w)
    assert abs(energy - e_want) <= 1e-9 * max(1.0, abs(e_want))
    for l1 in range(L1):
        for l2 in range(L2):
            same_structure(grad[l1][l2], want[l1][l2])
            assert rel(np.asarray(grad[l1][l2].storage), want[l1][l2].data) < 1e-8, (l1, l2)


def test_environment_natural_gradient_matches_oracle():
    """Natural gradient by conjugate gradient over chains against the oracle
    (3 steps), 4x4 at a truncating cut."""
    from oracle import environment as E, peps as opeps, sampling
    from test_gpu_ops import rel, same_structure
    from tnsp_amd import peps as pp
    from tnsp_amd.engine import AmplitudeEngine
    from tnsp_amd.environment import EnvironmentSampler

    L1, L2, N, cut = 4, 4, 8, 3
    spec, storages, sites = small_lattice(L1, L2, N, seed=48)
    eng = AmplitudeEngine(pp.Lattice(spec, storages), chunk=64, lanes=1)
    cand = configs(L1, L2, N, 16, seed=49)
    start = np.stack([c for c in cand if opeps.psi(sites, c, cut)[0] != 0][:4])
    smp = EnvironmentSampler(eng, start, cut=cut)
    grad, energy, steps = smp.natural_gradient(step=3)
    want, e_want, steps_want, _, _, _ = E.natural_gradient(sites, start, cut, sampling.tv_hamiltonian(1.0, 1.0),
                                                           step=3)
    assert steps == steps_want == 3
    assert abs(energy - e_want) <= 1e-9 * max(1.0, abs(e_want))
    for l1 in range(L1):
        for l2 in range(L2):
            same_structure(grad[l1][l2], want[l1][l2])
            assert rel(np.asarray(grad[l1][l2].storage), want[l1][l2].data) < 1e-7, (l1, l2)


def test_gradient_step_lowers_exact_energy():
    """apply_gradient (lattice.py:936-948) with the observer's gradient over
    every configuration of a 2x3 lattice (weights |psi|^2, i.e. the exact
    expectation) lowers the exact energy by eta |g|^2 to first order; the
    engine and the sampler re-project the updated sites."""
    import itertools

    from tnsp_amd import peps as pp
    from tnsp_amd.engine import AmplitudeEngine
    from tnsp_amd.environment import EnvironmentSampler
    from tnsp_amd.sampler import TVSampler

    L1, L2, N = 2, 3, 3
    spec, storages, _ = small_lattice(L1, L2, N, seed=50)
    lat = pp.Lattice(spec, storages)
    eng = AmplitudeEngine(lat, chunk=64, lanes=1)
    allc = []
    for occ in itertools.combinations(range(L1 * L2), N):
        c = np.zeros(L1 * L2, dtype=np.int64)
        c[list(occ)] = 1
        allc.append(c.reshape(L1, L2))
    allc = np.stack(allc)
    cfgs = allc[eng.evaluate(allc) != 0]

    def exact_energy():
        obs = TVSampler(eng, cfgs)
        return float(np.sum(obs.psi ** 2 * obs.local_energy()) / np.sum(obs.psi ** 2))

    e0 = exact_energy()
    smp = EnvironmentSampler(eng, cfgs, cut=10_000)
    grad, eg = smp.energy_gradient(reweight=smp.psi ** 2)
    assert abs(eg - e0) <= 1e-9 * max(1.0, abs(e0))
    g2 = sum(float((np.asarray(g.storage) ** 2).sum()) for row in grad for g in row)
    before = np.asarray(lat.sites[0][1].storage).copy()
    g01 = np.asarray(grad[0][1].transpose(lat.sites[0][1].names).storage)
    eta = 1e-4 / max(g2, 1e-300) * abs(e0)
    lat.apply_gradient(grad, eta)
    smp.refresh_sites()
    np.testing.assert_allclose(np.asarray(lat.sites[0][1].storage), before - eta * g01, rtol=0, atol=1e-14)
    e1 = exact_energy()
    assert e1 < e0
    assert abs((e0 - e1) - eta * g2) <= 0.1 * eta * g2


def small_lattice_z2(L1, L2, N, seed):
    from oracle import peps
    from tnsp_amd import peps as pp

    spec = pp.LatticeSpec("FermiZ2", L1, L2, (2, 2), N, 1000, seed=seed)
    storages = pp.random_storages(spec)
    sd = {"kind": spec.kind, "L1": L1, "L2": L2, "N": N, "segments": list(spec.segments)}
    return spec, storages, peps.build_sites(sd, storages)


def test_environment_fermi_z2_sweep_energy_holes():
    """FermiZ2 lattice (even/odd bonds, D = 4), 4x4 at a truncating cut:
    sweep decisions, observer energy and holes against the oracle."""
    from oracle import environment as E, peps as opeps, sampling
    from test_gpu_ops import rel, same_structure
    from tnsp_amd import peps as pp
    from tnsp_amd.engine import AmplitudeEngine
    from tnsp_amd.environment import EnvironmentSampler

    L1, L2, N, cut = 4, 4, 8, 4
    spec, storages, sites = small_lattice_z2(L1, L2, N, seed=60)
    eng = AmplitudeEngine(pp.Lattice(spec, storages), chunk=64, lanes=1)
    cand = configs(L1, L2, N, 16, seed=61)
    start = np.stack([c for c in cand if opeps.psi(sites, c, cut)[0] != 0][:3])
    H = sampling.tv_hamiltonian(1.0, 1.0, kind_name="FermiZ2")
    smp = EnvironmentSampler(eng, start, cut=cut, seed=62, lanes=2)
    # observer and holes at the start configurations
    np.testing.assert_allclose(smp.local_energy(), [E.local_energy(sites, c, cut, H) for c in start],
                               rtol=1e-9, atol=1e-9)
    holes, _ = smp.holes()
    for b, cfg in enumerate(start):
        want = E.holes(sites, cfg, cut)
        for l1 in range(L1):
            for l2 in range(L2):
                same_structure(holes[l1][l2].tensor(b), want[l1][l2])
                assert rel(np.asarray(holes[l1][l2].tensor(b).storage), want[l1][l2].data) < 1e-10
    # one sweep with the hole(()) refresh against the oracle's
    gen = np.random.Generator(np.random.PCG64(62))
    smp.sweep()
    u = np.stack([gen.random(len(start)) for _ in range(len(smp.bonds))])
    for b in range(len(start)):
        cfg, _, ws = E.sweep(sites, start[b], cut, u[:, b])
        assert np.array_equal(smp.configs[b], cfg)
        assert abs(smp.psi[b] - ws) <= 1e-10 * ws
