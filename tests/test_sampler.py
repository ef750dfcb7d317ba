"""Sweep sampler and local energy (tnsp_amd/sampler.py).

Pins:
* the edge order of the amplitude tensor against the names TAT returned for
  every golden lattice;
* the closed-form local energy against the observer's contraction of
  one-element fermionic tensors (oracle/sampling.py, restating
  tetragono/tetragono/sampling_lattice/observer.py:350-379);
* the sampled distribution against |psi|^2 by brute-force enumeration of a
  small lattice (any error in the proposal or acceptance rule breaks detailed
  balance);
* the autocorrelation estimator against the closed form of an AR(1) chain.
"""
import itertools

import numpy as np
import pytest

from helpers import load_golden, load_golden_8x8

_, _, AMPS = load_golden()
AMPS8 = load_golden_8x8()


def small_lattice(L1, L2, N, seed):
    from oracle import peps
    from tnsp_amd import peps as pp

    spec = pp.LatticeSpec("FermiU1", L1, L2, (2, 1, 1), N, 1000, seed=seed)
    storages = pp.random_storages(spec)
    sd = {"kind": spec.kind, "L1": L1, "L2": L2, "N": N, "segments": list(spec.segments)}
    return spec, storages, peps.build_sites(sd, storages)


class OracleAmplitudes:
    """Stands in for the batched engine in CPU tests: psi from the oracle."""

    def __init__(self, spec, sites):
        self.spec = spec
        self.sites = sites
        self.cache = {}
        self.calls = 0

    def evaluate(self, configs):
        from oracle import peps

        out = []
        for c in np.asarray(configs):
            key = c.tobytes()
            if key not in self.cache:
                self.cache[key] = peps.psi(self.sites, c, self.spec.cut)[0]
            out.append(self.cache[key])
        self.calls += 1
        return np.array(out)


@pytest.mark.parametrize("case", list(AMPS) + list(AMPS8))
def test_hole_name_order_matches_tat(case):
    from tnsp_amd.sampler import hole_name_order

    c = AMPS[case] if case in AMPS else AMPS8[case]
    L1, L2 = c["spec"]["L1"], c["spec"]["L2"]
    for p in c["psi"]:
        if p["names"]:
            assert p["names"] == hole_name_order(L1, L2)


@pytest.mark.parametrize("L1,L2,N,seed", [(3, 3, 4, 22), (2, 3, 3, 24), (3, 2, 3, 26)])
def test_local_energy_matches_observer_contraction(L1, L2, N, seed):
    from oracle import sampling
    from tnsp_amd.sampler import TVSampler

    spec, storages, sites = small_lattice(L1, L2, N, seed)
    amps = OracleAmplitudes(spec, sites)
    from tnsp_amd import peps as pp

    cfgs = pp.sample_configurations(spec, 12, seed=seed + 1)
    psi = amps.evaluate(cfgs)
    cfgs = cfgs[psi != 0][:6]
    sampler = TVSampler(amps, cfgs, t=1.0, V=1.0)
    got = sampler.local_energy()
    H = sampling.tv_hamiltonian(1.0, 1.0)
    for c, cfg in enumerate(cfgs):
        ref = sampling.local_energy(sites, cfg, spec.cut, H)
        assert abs(got[c] - ref) <= 1e-10 * max(1.0, abs(ref)), (c, got[c], ref)


def test_local_energy_scales_with_t_and_v():
    """H = -t CC + V NN is linear in (t, V): E(t, V) = t E(1, 0) + V E(0, 1)."""
    from oracle import sampling
    from tnsp_amd.sampler import TVSampler

    spec, storages, sites = small_lattice(3, 3, 4, 22)
    amps = OracleAmplitudes(spec, sites)
    from tnsp_amd import peps as pp

    cfgs = pp.sample_configurations(spec, 8, seed=3)
    cfgs = cfgs[amps.evaluate(cfgs) != 0][:3]
    e_t = TVSampler(amps, cfgs, t=1.0, V=0.0).local_energy()
    e_v = TVSampler(amps, cfgs, t=0.0, V=1.0).local_energy()
    e = TVSampler(amps, cfgs, t=0.7, V=-1.3).local_energy()
    np.testing.assert_allclose(e, 0.7 * e_t - 1.3 * e_v, rtol=1e-12, atol=1e-12)
    H = sampling.tv_hamiltonian(0.7, -1.3)
    for c, cfg in enumerate(cfgs):
        assert abs(e[c] - sampling.local_energy(sites, cfg, spec.cut, H)) < 1e-10 * max(1, abs(e[c]))


def enumerate_distribution(amps, L1, L2, N):
    cfgs = []
    for occ in itertools.combinations(range(L1 * L2), N):
        c = np.zeros(L1 * L2, dtype=np.int8)
        c[list(occ)] = 1
        cfgs.append(c.reshape(L1, L2))
    cfgs = np.array(cfgs)
    w = np.abs(amps.evaluate(cfgs)) ** 2
    return cfgs, w / w.sum()


def test_sweep_samples_psi_squared():
    """Detailed balance: the stationary distribution of the sweep is |psi|^2."""
    from tnsp_amd.sampler import TVSampler

    spec, storages, sites = small_lattice(2, 3, 3, 24)
    amps = OracleAmplitudes(spec, sites)
    cfgs, p = enumerate_distribution(amps, 2, 3, 3)
    keys = {c.tobytes(): i for i, c in enumerate(cfgs)}
    chains = 400
    start = np.repeat(cfgs[np.argmax(p)][None], chains, axis=0)
    s = TVSampler(amps, start, seed=11)
    counts = np.zeros(len(cfgs))
    for sweep in range(60):
        s.sweep()
        if sweep >= 10:
            for c in s.configs:
                counts[keys[c.tobytes()]] += 1
    emp = counts / counts.sum()
    # total-variation distance of 20000 correlated samples
    assert 0.5 * np.abs(emp - p).sum() < 0.05
    assert s.accepted > 0


def test_autocorrelation_of_ar1():
    from tnsp_amd.sampler import integrated_autocorrelation

    gen = np.random.default_rng(0)
    phi, n, chains = 0.8, 4000, 64
    x = np.zeros((chains, n))
    e = gen.standard_normal((chains, n))
    for i in range(1, n):
        x[:, i] = phi * x[:, i - 1] + e[:, i]
    tau = integrated_autocorrelation(x)
    exact = (1 + phi) / (1 - phi)
    assert abs(tau - exact) / exact < 0.1


def test_local_energy_closed_form_fermi_z2():
    """FermiZ2 lattices (tV/parity): the closed form with an even T edge
    equals the observer contraction with the parity operators
    (common_tensor/Parity.py) on a 3x3 lattice."""
    from oracle import peps, sampling
    from tnsp_amd import peps as pp
    from tnsp_amd.sampler import TVSampler

    spec = pp.LatticeSpec("FermiZ2", 3, 3, (2, 2), 4, 1000, seed=31)
    storages = pp.random_storages(spec)
    sd = {"kind": spec.kind, "L1": 3, "L2": 3, "N": 4, "segments": list(spec.segments)}
    sites = peps.build_sites(sd, storages)
    rng = np.random.default_rng(32)
    cfgs = []
    while len(cfgs) < 3:
        c = np.zeros(9, dtype=np.int64)
        c[rng.choice(9, 4, replace=False)] = 1
        c = c.reshape(3, 3)
        if peps.psi(sites, c, spec.cut)[0] != 0:
            cfgs.append(c)
    H = sampling.tv_hamiltonian(1.0, 1.0, kind_name="FermiZ2")
    got = TVSampler(OracleAmplitudes(spec, sites), np.stack(cfgs)).local_energy()
    want = [sampling.local_energy(sites, c, spec.cut, H) for c in cfgs]
    np.testing.assert_allclose(got, want, rtol=1e-10, atol=1e-10)
