"""GPU direct sampling (tnsp_amd/direct.py) against the oracle
(oracle/direct.py).

* At truncating single- and double-layer cuts, draws driven by the same
  uniforms pick the same configurations with the same probability (1e-9
  relative), FermiU1 and FermiZ2: every site hole the GPU builds equals the
  oracle's up to rounding.
* Without truncation each drawn configuration's probability is
  |psi(s)|^2 / sum |psi|^2 (the Born rule, with psi from the oracle's pinned
  amplitude).
"""
import numpy as np
import pytest

from test_direct import all_configs, z2_lattice
from test_sampler import small_lattice

pytestmark = pytest.mark.gpu


def _sampler(spec, storages, cut, cut2, seed=0):
    from tnsp_amd import peps as pp
    from tnsp_amd.direct import DirectSampler

    return DirectSampler(pp.Lattice(spec, storages), cut=cut, cut2=cut2, seed=seed)


@pytest.mark.parametrize("kind", ["FermiU1", "FermiZ2"])
def test_direct_draw_matches_oracle_truncated(kind):
    from oracle import direct as Dr

    L1, L2, cut, cut2, B = 3, 3, 3, 3, 4
    if kind == "FermiU1":
        spec, storages, sites = small_lattice(L1, L2, 4, seed=21)
    else:
        spec, storages, sites = z2_lattice(L1, L2, seed=22)
    smp = _sampler(spec, storages, cut, cut2)
    u = np.random.default_rng(5).random((B, L1 * L2))
    poss, cfg, alive = smp._draw(B, u)
    down = Dr.double_layer_down_to_up(sites, cut2)
    for b in range(B):
        p, c, _ = Dr.direct_sample(sites, cut, cut2, uniforms=u[b], down=down)
        print(f"{kind} draw {b}: gpu {poss[b]:.6e} {cfg[b].ravel()} oracle {p:.6e}", flush=True)
        assert alive[b] == (c is not None)
        if c is not None:
            assert np.array_equal(cfg[b], c)
            assert poss[b] == pytest.approx(p, rel=1e-9)


def test_direct_uncut_born_rule():
    from oracle import peps

    L1, L2, N = 2, 3, 3
    spec, storages, sites = small_lattice(L1, L2, N, seed=3)
    big = 10_000
    smp = _sampler(spec, storages, big, big, seed=9)
    w = {cfg.tobytes(): peps.psi(sites, cfg, big)[0] ** 2 for cfg in all_configs(L1, L2, "FermiU1", N)}
    Z = sum(w.values())
    poss, cfg = smp.sample(32)
    assert smp.redrawn == 0
    want = np.array([w[c.astype(np.int64).tobytes()] / Z for c in cfg])
    np.testing.assert_allclose(poss, want, rtol=1e-9, atol=0)
    assert len({c.tobytes() for c in cfg}) > 1


def test_direct_weights_and_gradient_steps_uncut():
    """Without truncation every direct sample weighs |psi|^2 / p = Z
    (instability zero); the gradient-descent driver with direct sampling
    rebuilds the double layer after each update, so the weights stay equal
    for the updated lattice too."""
    from tnsp_amd import peps as pp
    from tnsp_amd.direct import DirectSampler, instability, reweight
    from tnsp_amd.engine import AmplitudeEngine
    from tnsp_amd.environment import EnvironmentSampler
    from tnsp_amd.vmc import gradient_descent

    L1, L2, N = 2, 3, 3
    spec, storages, _ = small_lattice(L1, L2, N, seed=50)
    lat = pp.Lattice(spec, storages)
    eng = AmplitudeEngine(lat, chunk=64, lanes=1)
    big = 10_000
    dsm = DirectSampler(lat, cut=big, cut2=big, seed=4)
    poss, cfg = dsm.sample(16)
    w = reweight(eng.evaluate(cfg, cut=big), poss)
    assert np.all(w > 0) and instability(w) < 1e-7
    smp = EnvironmentSampler(eng, cfg, cut=big)
    recs = gradient_descent(smp, 2, 1e-3, direct=dsm, direct_samples=16)
    print(recs, flush=True)
    assert all(r["instability"] < 1e-7 and np.isfinite(r["energy_per_site"]) for r in recs)
