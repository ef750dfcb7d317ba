"""The gradient-descent driver's host logic (tnsp_amd/vmc.py) with scripted
samplers, on the CPU: per step the samples (sweeps, or direct draws weighted
|psi|^2 / p), the observer's gradient over them, the in-place update, the
re-projection and, for direct sampling, the rebuilt double layer -- in
tetragono's order (sampling_lattice/gradient.py:340-435)."""
from types import SimpleNamespace

import numpy as np


class Lattice:
    spec = SimpleNamespace(L1=1, L2=2)

    def __init__(self, log):
        self.log = log

    def apply_gradient(self, grad, step):
        self.log.append(("apply", grad, step))


class Sampler:
    def __init__(self, log):
        self.log = log
        self.lattice = Lattice(log)
        self.cut = 7
        self.configs = np.array([[[1, 0]], [[0, 1]]], dtype=np.int8)
        self.engine = SimpleNamespace(evaluate=self.evaluate)

    def evaluate(self, cfg, cut):
        self.log.append(("psi", len(cfg), cut))
        return np.array([2.0, 0.0, 1.0])[:len(cfg)]

    def sweep(self):
        self.log.append(("sweep",))
        self.configs = self.configs[::-1].copy()

    def energy_gradient(self, configs, reweight):
        self.log.append(("grad", len(configs), None if reweight is None else reweight.tolist()))
        return "g", -1.0

    def natural_gradient(self, step, error, configs, reweight):
        self.log.append(("natural", step, error, len(configs), reweight.tolist()))
        return "ng", -2.0, 3

    def refresh_sites(self):
        self.log.append(("refresh",))


def test_sweep_steps():
    from tnsp_amd.vmc import gradient_descent

    log = []
    recs = gradient_descent(Sampler(log), 1, 0.5, sweeps=2)
    assert log == [("sweep",), ("sweep",), ("grad", 4, None), ("apply", "g", 0.5), ("refresh",)]
    assert recs[0]["samples"] == 4 and recs[0]["energy_per_site"] == -0.5 and recs[0]["cg_steps"] == 0


def test_direct_steps_weigh_and_rebuild():
    from tnsp_amd.vmc import gradient_descent

    log = []
    direct = SimpleNamespace(sample=lambda n: (np.array([0.5, 0.25, 0.25])[:n], np.zeros((n, 1, 2), np.int8)),
                             refresh_all=lambda: log.append(("double layer",)))
    recs = gradient_descent(Sampler(log), 1, 0.1, natural=True, cg_step=5, cg_error=1e-3, direct=direct,
                            direct_samples=3)
    assert log == [("psi", 3, 7), ("natural", 5, 1e-3, 3, [8.0, 0.0, 4.0]), ("apply", "ng", 0.1), ("refresh",),
                   ("double layer",)]
    w = np.array([8.0, 0.0, 4.0])  # psi^2 / p, none where psi = 0
    assert recs[0]["instability"] == np.std(w) / np.mean(w)
    assert recs[0]["cg_steps"] == 3 and recs[0]["energy_per_site"] == -1.0
