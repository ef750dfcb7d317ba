"""Sweep sampler and local energy driven by the batched GPU engine."""
import numpy as np
import pytest

from test_sampler import OracleAmplitudes, enumerate_distribution, small_lattice

pytestmark = pytest.mark.gpu


def test_local_energy_gpu_matches_oracle_amplitudes():
    from oracle import peps as opeps
    from tnsp_amd import peps as pp
    from tnsp_amd.engine import AmplitudeEngine
    from tnsp_amd.sampler import TVSampler

    spec = pp.C2
    storages = pp.random_storages(spec)
    lat = pp.Lattice(spec, storages)
    eng = AmplitudeEngine(lat, chunk=64, lanes=2)
    sd = {"kind": spec.kind, "L1": spec.L1, "L2": spec.L2, "N": spec.N, "segments": list(spec.segments)}
    amps = OracleAmplitudes(spec, opeps.build_sites(sd, storages))
    cfgs = pp.sample_configurations(spec, 3, seed=4)
    got = TVSampler(eng, cfgs).local_energy()
    ref = TVSampler(amps, cfgs).local_energy()
    np.testing.assert_allclose(got, ref, rtol=1e-9, atol=1e-9)


def test_gpu_sweep_samples_psi_squared():
    from tnsp_amd import peps as pp
    from tnsp_amd.engine import AmplitudeEngine
    from tnsp_amd.sampler import TVSampler

    spec, storages, sites = small_lattice(2, 3, 3, 24)
    eng = AmplitudeEngine(pp.Lattice(spec, storages), chunk=512, lanes=1)
    cfgs, p = enumerate_distribution(eng, 2, 3, 3)
    keys = {c.tobytes(): i for i, c in enumerate(cfgs)}
    chains = 1000
    start = np.repeat(cfgs[np.argmax(p)][None], chains, axis=0)
    s = TVSampler(eng, start, seed=5)
    counts = np.zeros(len(cfgs))
    for sweep in range(30):
        s.sweep()
        if sweep >= 5:
            for c in s.configs:
                counts[keys[c.tobytes()]] += 1
    emp = counts / counts.sum()
    assert 0.5 * np.abs(emp - p).sum() < 0.04
