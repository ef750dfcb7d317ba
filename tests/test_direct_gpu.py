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
