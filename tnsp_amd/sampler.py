"""Metropolis sweep sampling and local energy of the tV model on the GPU.

Follows tetragono's ``SweepSampling`` (paths relative to the TNSP repository
root): bonds are visited in ``_get_proper_position_order``
(tetragono/tetragono/sampling_lattice/sampling.py:156-190: horizontal bonds
row-major, then vertical bonds column-major); on each bond ``_single_term``
(:117-143) draws one target uniformly among the non-zero elements of the bond
Hamiltonian for the current occupations and accepts it with
``p = |psi(s')/psi(s)|^(2 alpha) * n / n'``.  For the tV Hamiltonian
``H = -t CC + V NN`` (tetraku/tetraku/models/tV/__init__.py:39-45) the only
off-diagonal targets are single hops between a filled and an empty site
(n = n' = 1); a doubly occupied bond only has its diagonal element, so its
"hop" leaves the configuration unchanged.

All chains advance in lockstep: per bond, the proposed configurations of
every chain whose bond can hop are evaluated in ONE call of the batched
amplitude engine (psi(s') from scratch rather than from tetragono's cached
environments, ``single_layer_auxiliaries.py:500-708``; the two agree exactly
when the boundary-MPS cut does not truncate).

Local energy (``observer.py:350-389``):

    E(s) = V sum_<ij> n_i n_j  -  t sum_<ij>, n_i != n_j  (-1)^k psi(s^ij)/psi(s)

where s^ij is s with the particle hopped across the bond and k counts the
particles strictly between the two sites in the edge order of the amplitude
tensor ``hole(())`` (the one-dimensional ``T`` edge, carrying charge -N,
counts N).  This closed form is what the observer's contraction of
one-element fermionic tensors reduces to; the CPU test-suite checks it
against that contraction.
"""
import time

import numpy as np

from .engine import AmplitudeEngine


def bond_order(L1, L2):
    """sampling.py:156-190 for nearest-neighbour bond Hamiltonians."""
    order = [((l1, l2), (l1, l2 + 1)) for l1 in range(L1) for l2 in range(L2 - 1)]
    order += [((l1, l2), (l1 + 1, l2)) for l2 in range(L2) for l1 in range(L1 - 1)]
    return order


def hole_name_order(L1, L2):
    """Edge order of the amplitude tensor returned by ``hole(())``: the
    physical edges P_{l1}_{l2}_0 column by column from the right, bottom row
    first within a column, then ``T`` and finally P_0_0_0."""
    names = [f"P_{l1}_{l2}_0" for l2 in reversed(range(L2)) for l1 in reversed(range(L1)) if (l1, l2) != (0, 0)]
    return names + ["T", "P_0_0_0"]


def _fermion_order_index(L1, L2):
    """Position of every site (row-major index) and of T in the hole order."""
    order = hole_name_order(L1, L2)
    pos = np.zeros(L1 * L2, dtype=np.int64)
    for i, n in enumerate(order):
        if n != "T":
            _, a, b, _ = n.split("_")
            pos[int(a) * L2 + int(b)] = i
    return pos, order.index("T")


class TVSampler:
    """Lockstep Metropolis sweep sampling of B chains of a FermiU1 or
    FermiZ2 PEPS with the tV Hamiltonian; ``engine`` is an
    :class:`AmplitudeEngine`."""

    def __init__(self, engine, configs, t=1.0, V=1.0, alpha=1.0, seed=0):
        spec = engine.spec
        if spec.kind not in ("FermiU1", "FermiZ2"):
            raise ValueError("the tV sampler needs a FermiU1 or FermiZ2 lattice")
        self.engine = engine
        self.L1, self.L2 = spec.L1, spec.L2
        self.N = spec.N
        # parity of the T edge: -N for FermiU1, even for FermiZ2
        self.t_parity = spec.N % 2 if spec.kind == "FermiU1" else 0
        self.t, self.V, self.alpha = float(t), float(V), float(alpha)
        self.configs = np.ascontiguousarray(np.asarray(configs), dtype=np.int8).copy()
        if self.configs.ndim != 3 or self.configs.shape[1:] != (self.L1, self.L2):
            raise ValueError("configs must have shape (chains, L1, L2)")
        counts = self.configs.reshape(len(self.configs), -1).sum(1)
        if spec.kind == "FermiU1" and np.any(counts != self.N):
            raise ValueError("every configuration must hold N particles")
        if spec.kind == "FermiZ2" and np.any(counts % 2):
            raise ValueError("every configuration must hold an even number of particles")
        self.gen = np.random.Generator(np.random.PCG64(seed))
        self.bonds = bond_order(self.L1, self.L2)
        self.psi = engine.evaluate(self.configs)
        if np.any(self.psi == 0):
            raise ValueError("initial configurations with psi = 0")
        self._pos, self._tpos = _fermion_order_index(self.L1, self.L2)
        self.proposals = 0
        self.accepted = 0
        self.amplitudes = 0

    def sweep(self):
        """One sweep over every bond (sampling.py:145-154)."""
        for (a, b) in self.bonds:
            self._bond(a, b)
        return self.configs

    def _bond(self, a, b):
        """``_single_term`` on bond (a, b) for every chain; returns the chains
        that tried a hop, their configurations before it, psi before and
        psi of the proposal."""
        B = len(self.configs)
        flat = self.configs.reshape(B, -1)
        ia, ib = a[0] * self.L2 + a[1], b[0] * self.L2 + b[1]
        idx = np.nonzero(flat[:, ia] != flat[:, ib])[0]
        # one uniform draw per chain for the acceptance (the target choice is
        # the single hop here)
        u = self.gen.random(B)
        if len(idx) == 0:
            return idx, flat[idx], self.psi[idx], self.psi[idx]
        before = flat[idx].copy()
        psi_old = self.psi[idx].copy()
        prop = before.copy()
        prop[:, [ia, ib]] = prop[:, [ib, ia]]
        psi_new = self.engine.evaluate(prop.reshape(-1, self.L1, self.L2))
        self.amplitudes += len(idx)
        self.proposals += len(idx)
        p = np.abs(psi_new / psi_old) ** (2 * self.alpha)
        acc = u[idx] < p
        take = idx[acc]
        flat[take] = prop[acc]
        self.psi[take] = psi_new[acc]
        self.accepted += int(acc.sum())
        return idx, before, psi_old, psi_new

    def hop_sign(self, flat, ia, ib):
        """(-1)^k, k = particles strictly between sites ia and ib in the hole
        edge order (T counts its parity: N for FermiU1, even for FermiZ2)."""
        pa, pb = self._pos[ia], self._pos[ib]
        lo, hi = min(pa, pb), max(pa, pb)
        between = (self._pos > lo) & (self._pos < hi)
        k = flat[:, between].sum(1).astype(np.int64)
        if lo < self._tpos < hi:
            k = k + self.t_parity
        return 1 - 2 * (k % 2)

    def local_energy(self, configs=None, psi=None):
        """E_loc of every chain (observer.py:350-389 with the tV elements)."""
        configs = self.configs if configs is None else configs
        psi = self.psi if psi is None else psi
        B = len(configs)
        flat = configs.reshape(B, -1).astype(np.int64)
        e = np.zeros(B)
        props, owners, coefs = [], [], []
        for (a, b) in self.bonds:
            ia, ib = a[0] * self.L2 + a[1], b[0] * self.L2 + b[1]
            e += self.V * flat[:, ia] * flat[:, ib]
            idx = np.nonzero(flat[:, ia] != flat[:, ib])[0]
            if len(idx) == 0:
                continue
            prop = flat[idx].copy()
            prop[:, [ia, ib]] = prop[:, [ib, ia]]
            props.append(prop)
            owners.append(idx)
            coefs.append(-self.t * self.hop_sign(flat[idx], ia, ib))
        if props:
            allp = np.concatenate(props)
            psi_new = self.engine.evaluate(allp.reshape(-1, self.L1, self.L2))
            self.amplitudes += len(allp)
            own = np.concatenate(owners)
            with np.errstate(divide="ignore", invalid="ignore"):
                np.add.at(e, own, np.concatenate(coefs) * psi_new / psi[own])
        return e


def integrated_autocorrelation(series, c=5.0):
    """Integrated autocorrelation time of a (chains, samples) series with
    Sokal's self-consistent window (smallest W with W >= c tau(W)); the
    autocorrelation is averaged over chains.  ESS = chains * samples / tau."""
    x = np.asarray(series, dtype=np.float64)
    x = x - x.mean(axis=1, keepdims=True)
    n = x.shape[1]
    var = np.mean(x * x)
    if var == 0 or n < 4:
        return 1.0
    f = np.fft.rfft(x, 2 * n, axis=1)
    acov = np.fft.irfft(f * np.conj(f), axis=1)[:, :n].mean(axis=0) / n
    rho = acov / acov[0]
    tau = 1.0
    for W in range(1, n):
        tau = 1.0 + 2.0 * np.sum(rho[1:W + 1])
        if W >= c * tau:
            break
    return max(tau, 1e-9)


def measure_ess(advance, energy, samples, burn_in=2, sites=1, log=None, extra=None):
    """Calls ``advance()`` ``burn_in`` + ``samples`` times, recording
    ``energy()`` (per-chain local energies) after each timed advance; returns
    ESS of the energy estimator, the sampling seconds (energy evaluation not
    included), the energy per site and tau.  ``extra()``, if given, is a
    second per-chain observable recorded alongside (its ESS is reported as
    ``ess_extra``)."""
    if samples < 1:
        raise ValueError("measure_ess needs at least one sample")
    for i in range(burn_in):
        advance()
        if log:
            log(f"burn-in {i + 1}/{burn_in}")
    energies, extras = [], []
    t_sample = 0.0
    for _ in range(samples):
        t0 = time.time()
        advance()
        t_sample += time.time() - t0
        if log:
            log(f"sample {len(energies) + 1}/{samples} drawn, {t_sample:.1f}s sampling")
        energies.append(energy())
        if extra is not None:
            extras.append(extra())
        if log:
            log(f"{len(energies)}/{samples} samples, {t_sample:.1f}s sampling")
    series = np.stack(energies, axis=1)
    # a chain whose amplitude is exactly zero at some sample (every sector of
    # one boundary-MPS cut truncated away) has no local energy there: it is
    # left out of the estimator and counted
    finite = np.all(np.isfinite(series), axis=1)
    dropped = int(len(series) - finite.sum())
    series = series[finite]
    tau = integrated_autocorrelation(series)
    ess = series.size / tau
    out = {"ess": ess, "sample_s": t_sample, "ess_per_s": ess / max(t_sample, 1e-12),
           "energy_per_site": float(series.mean() / sites), "energy_median_per_site": float(np.median(series) / sites),
           "tau": tau, "chains": series.shape[0], "chains_dropped": dropped, "samples": samples}
    if extras:
        xs = np.stack(extras, axis=1)
        tau_x = integrated_autocorrelation(xs)
        out.update({"tau_extra": tau_x, "ess_extra": xs.size / tau_x,
                    "ess_extra_per_s": xs.size / tau_x / max(t_sample, 1e-12),
                    "extra_per_site": float(xs.mean() / sites)})
    return out


def make_engine(lattice, **kw):
    return AmplitudeEngine(lattice, **kw)
