"""Variational Monte Carlo optimisation of a PEPS: tetragono's
``gradient_descent`` with sweep sampling (sampling_lattice/gradient.py:93-430,
paths relative to the TNSP repository root) for many chains at once.

Each gradient step draws ``sweeps`` samples per chain with the
cached-environment sweep (``EnvironmentSampler``), takes the observer's
gradient over all of them (plain, or natural by conjugate gradient), applies
``Lattice.apply_gradient`` and re-projects the sites.  The samples of one step
are the chains' configurations after each sweep, each with weight one (sweep
sampling draws from |psi|^2).
"""
import time

import numpy as np


def gradient_descent(sampler, grad_steps, step_size, sweeps=1, natural=False, cg_step=20, cg_error=1e-4,
                     log=None, direct=None, direct_samples=0):
    """Runs ``grad_steps`` steps; returns a list of per-step records
    (energy per site before the step, seconds, samples, and for direct
    sampling the weights' instability)."""
    from .direct import instability, reweight

    lat = sampler.lattice
    sites = lat.spec.L1 * lat.spec.L2
    out = []
    for k in range(grad_steps):
        t0 = time.time()
        w, extra = None, {}
        if direct is not None:
            poss, cfg = direct.sample(direct_samples)
            w = reweight(sampler.engine.evaluate(cfg, cut=sampler.cut), poss)
            extra["instability"] = float(instability(w))
        else:
            samples = []
            for _ in range(sweeps):
                sampler.sweep()
                samples.append(sampler.configs.copy())
            cfg = np.concatenate(samples)
        if natural:
            grad, energy, steps = sampler.natural_gradient(cg_step, cg_error, configs=cfg, reweight=w)
        else:
            grad, energy = sampler.energy_gradient(configs=cfg, reweight=w)
            steps = 0
        lat.apply_gradient(grad, step_size)
        sampler.refresh_sites()
        if direct is not None:
            direct.refresh_all()
        rec = {"step": k, "energy_per_site": energy / sites, "seconds": time.time() - t0, "samples": len(cfg),
               "cg_steps": steps, **extra}
        out.append(rec)
        if log:
            log(f"gradient step {k}: energy/site {rec['energy_per_site']:.6f}, {rec['seconds']:.1f}s")
    return out
