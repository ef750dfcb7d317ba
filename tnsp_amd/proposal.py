"""Learned hop proposal for Metropolis sampling of the tV PEPS.

A small convolutional network (3x3 convolutions, ``channels`` wide, ReLU,
zero padding; bf16 weights with fp32 accumulation) maps an occupation map to
logits over the 4 L1 L2 nearest-neighbour hops; proposals are drawn from the
softmax over the valid hops.  Training regresses the logit of each hop the
plain sweep sampler tried onto log|psi(s')/psi(s)|, so the trained proposal
approaches the locally balanced q(s'|s) ~ |psi(s')/psi(s)|.  Every forward,
backward and optimiser step runs in the hand-written kernels of
``csrc/proposal.hip`` through the C ABI; torch only provides the buffers.

Acceptance is Metropolis-Hastings with the exact amplitude ratio,
``min(1, |psi(s')/psi(s)|^(2 alpha) q(s|s') / q(s'|s))``, so the sampled
distribution stays |psi|^(2 alpha) whatever the network has learned; the
plain sweep (``tnsp_amd.sampler.TVSampler``) remains the fallback.
No reference counterpart: tetragono proposes bond by bond
(tetragono/tetragono/sampling_lattice/sampling.py:117-154).
"""
import ctypes
import time

import numpy as np
import torch

from . import _lib

DIRS = 4  # 0 up, 1 down, 2 left, 3 right


def _p(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else ctypes.c_void_p(0)


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


class HopProposalNet:
    def __init__(self, L1, L2, channels=32, layers=4, seed=0, device="cuda"):
        self.L1, self.L2, self.C, self.layers = L1, L2, channels, layers
        self.device = torch.device(device)
        gen = np.random.Generator(np.random.PCG64(seed))
        dims = [1] + [channels] * (layers - 1) + [DIRS]
        self.shapes = [(dims[i + 1], dims[i]) for i in range(layers)]
        self.w, self.b, self.wbf = [], [], []
        for li, (co, ci) in enumerate(self.shapes):
            scale = np.sqrt(2.0 / (ci * 9)) * (0.1 if li == layers - 1 else 1.0)
            w = torch.tensor(gen.standard_normal((co, ci, 3, 3)) * scale, dtype=torch.float32, device=self.device)
            self.w.append(w.contiguous())
            self.b.append(torch.zeros(co, dtype=torch.float32, device=self.device))
            self.wbf.append(torch.empty(co * ci * 9, dtype=torch.int16, device=self.device))
        self.params = self.w + self.b
        self.m = [torch.zeros_like(p) for p in self.params]
        self.v = [torch.zeros_like(p) for p in self.params]
        self.steps = 0
        self.sync_bf16()

    def sync_bf16(self):
        for w, wb in zip(self.w, self.wbf):
            _lib.check(_lib.lib().tns_to_bf16(_p(w), _p(wb), w.numel(), _stream()))

    @property
    def n_logits(self):
        return DIRS * self.L1 * self.L2

    def _occ(self, occ):
        if isinstance(occ, torch.Tensor):
            t = occ.to(device=self.device, dtype=torch.int8)
        else:
            t = torch.as_tensor(np.ascontiguousarray(occ, dtype=np.int8), device=self.device)
        return t.reshape(-1, self.L1, self.L2).contiguous()

    def forward(self, occ):
        """logits (B, 4 L1 L2); keeps the activations for ``backward``."""
        occ = self._occ(occ)
        B = occ.shape[0]
        H, W = self.L1, self.L2
        x = torch.empty((B, 1, H, W), dtype=torch.float32, device=self.device)
        _lib.check(_lib.lib().tns_occupations_to_input(_p(occ), _p(x), x.numel(), _stream()))
        acts = [x]
        for li, (co, ci) in enumerate(self.shapes):
            y = torch.empty((B, co, H, W), dtype=torch.float32, device=self.device)
            relu = int(li + 1 < self.layers)
            _lib.check(_lib.lib().tns_conv3x3_forward(_p(acts[-1]), _p(self.wbf[li]), _p(self.b[li]), _p(y), B, ci, co,
                                                      H, W, relu, _stream()))
            acts.append(y)
        self._acts = acts
        return acts[-1].reshape(B, -1)

    def backward(self, dlogits):
        """Parameter gradients from d loss / d logits (after ``forward``)."""
        acts = self._acts
        B = acts[0].shape[0]
        H, W = self.L1, self.L2
        dy = dlogits.reshape(acts[-1].shape).contiguous()
        grads_w, grads_b = [None] * self.layers, [None] * self.layers
        for li in reversed(range(self.layers)):
            co, ci = self.shapes[li]
            relu_y = acts[li + 1] if li + 1 < self.layers else None
            ws = ctypes.c_int64()
            _lib.check(_lib.lib().tns_conv3x3_backward_weight_workspace(B, ci, co, ctypes.byref(ws)))
            work = torch.empty(max(1, ws.value // 4), dtype=torch.float32, device=self.device)
            dw = torch.empty((co, ci, 3, 3), dtype=torch.float32, device=self.device)
            db = torch.empty(co, dtype=torch.float32, device=self.device)
            _lib.check(_lib.lib().tns_conv3x3_backward_weight(_p(dy), _p(relu_y), _p(acts[li]), _p(dw), _p(db), B, ci,
                                                              co, H, W, _p(work), _stream()))
            grads_w[li], grads_b[li] = dw, db
            if li > 0:
                dx = torch.empty_like(acts[li])
                _lib.check(_lib.lib().tns_conv3x3_backward_data(_p(dy), _p(relu_y), _p(self.wbf[li]), _p(dx), B, ci,
                                                                co, H, W, _stream()))
                dy = dx
        self.grads = grads_w + grads_b
        return self.grads

    def regression_step(self, occ, hops, targets, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8):
        """One Adam step on mean_b (logit[b, hop_b] - target_b)^2; returns the loss."""
        logits = self.forward(occ)
        B = logits.shape[0]
        hops = torch.as_tensor(np.asarray(hops, dtype=np.int32), device=self.device)
        targets = torch.as_tensor(np.asarray(targets, dtype=np.float32), device=self.device)
        dlog = torch.empty_like(logits)
        loss = torch.zeros(B, dtype=torch.float32, device=self.device)
        _lib.check(_lib.lib().tns_hop_regression(_p(logits), _p(hops), _p(targets), _p(dlog), _p(loss), B,
                                                 self.n_logits, _stream()))
        self.backward(dlog)
        self.steps += 1
        for li, (p, m, v, g) in enumerate(zip(self.params, self.m, self.v, self.grads)):
            wbf = self.wbf[li] if li < self.layers else None
            _lib.check(_lib.lib().tns_adam_step(_p(p), _p(m), _p(v), _p(g), _p(wbf), p.numel(), lr, beta1, beta2, eps,
                                                self.steps, _stream()))
        return float(loss.mean())

    def sample(self, occ, uniforms):
        """Draw one hop per configuration; returns (hops, log q) on the device."""
        occ = self._occ(occ)
        logits = self.forward(occ)
        B = occ.shape[0]
        u = torch.as_tensor(np.asarray(uniforms, dtype=np.float32), device=self.device)
        hops = torch.empty(B, dtype=torch.int32, device=self.device)
        logq = torch.empty(B, dtype=torch.float32, device=self.device)
        _lib.check(_lib.lib().tns_hop_sample(_p(logits), _p(occ), _p(u), _p(hops), _p(logq), B, self.L1, self.L2,
                                             _stream()))
        return hops, logq

    def log_prob(self, occ, hops):
        occ = self._occ(occ)
        logits = self.forward(occ)
        B = occ.shape[0]
        hops = torch.as_tensor(hops, dtype=torch.int32, device=self.device).contiguous()
        logq = torch.empty(B, dtype=torch.float32, device=self.device)
        _lib.check(_lib.lib().tns_hop_log_prob(_p(logits), _p(occ), _p(hops), _p(logq), B, self.L1, self.L2,
                                               _stream()))
        return logq


def apply_hops(occ, hops, L1, L2):
    """Configurations after the hops (numpy) and the reverse hop indices."""
    occ = np.array(occ, dtype=np.int8, copy=True).reshape(-1, L1 * L2)
    hops = np.asarray(hops, dtype=np.int64)
    HW = L1 * L2
    d, s = np.divmod(hops, HW)
    i, j = np.divmod(s, L2)
    ti = i + (d == 1) - (d == 0)
    tj = j + (d == 3) - (d == 2)
    t = ti * L2 + tj
    rows = np.arange(len(occ))
    occ[rows, s] = 0
    occ[rows, t] = 1
    rev = (d ^ 1) * HW + t
    return occ.reshape(-1, L1, L2), rev


def bond_hop(flat, ia, ib, L2, horizontal):
    """Hop index of the particle crossing bond (ia, ib) (ia left of / above ib)."""
    HW = flat.shape[1]
    from_a = flat[:, ia] == 1
    if horizontal:
        return np.where(from_a, 3 * HW + ia, 2 * HW + ib)
    return np.where(from_a, 1 * HW + ia, 0 * HW + ib)


class LearnedSampler:
    """Metropolis-Hastings chains driven by a HopProposalNet."""

    def __init__(self, engine, net, configs, alpha=1.0, seed=0):
        self.engine, self.net = engine, net
        self.L1, self.L2 = engine.spec.L1, engine.spec.L2
        self.alpha = float(alpha)
        self.configs = np.ascontiguousarray(np.asarray(configs), dtype=np.int8).copy()
        self.gen = np.random.Generator(np.random.PCG64(seed))
        self.psi = engine.evaluate(self.configs)
        if np.any(self.psi == 0):
            raise ValueError("initial configurations with psi = 0")
        self.proposals = 0
        self.accepted = 0
        self.amplitudes = 0

    def step(self):
        B = len(self.configs)
        hops, logq_f = self.net.sample(self.configs, self.gen.random(B))
        hops_h = hops.cpu().numpy()
        ok = hops_h >= 0
        new, rev = apply_hops(self.configs, np.where(ok, hops_h, 0), self.L1, self.L2)
        logq_r = self.net.log_prob(new, torch.as_tensor(rev.astype(np.int32)))
        idx = np.nonzero(ok)[0]
        psi_new = np.zeros(B)
        if len(idx):
            psi_new[idx] = self.engine.evaluate(new[idx])
            self.amplitudes += len(idx)
        with np.errstate(divide="ignore"):
            log_a = (2 * self.alpha * (np.log(np.abs(psi_new)) - np.log(np.abs(self.psi))) +
                     logq_r.cpu().numpy().astype(np.float64) - logq_f.cpu().numpy().astype(np.float64))
        u = self.gen.random(B)
        acc = ok & (np.log(np.maximum(u, 1e-300)) < log_a) & (psi_new != 0)
        # what the step decided on, for callers that check the acceptance rule
        self.last = {"hops": hops_h, "valid": ok, "proposed": new, "psi_new": psi_new, "log_a": log_a,
                     "uniforms": u, "accepted": acc}
        self.configs[acc] = new[acc]
        self.psi[acc] = psi_new[acc]
        self.proposals += int(ok.sum())
        self.accepted += int(acc.sum())
        return self.configs


def collect_training_data(sampler, sweeps, max_samples=None, log=None):
    """Runs plain sweeps and records every tried hop: (configuration before,
    hop index, log|psi(s')/psi(s)|).  ``sampler`` is a TVSampler or an
    EnvironmentSampler."""
    occs, hops, targets = [], [], []
    L1, L2 = sampler.L1, sampler.L2
    count = 0

    def record(a, b, before, psi_old, psi_new):
        nonlocal count
        ia, ib = a[0] * L2 + a[1], b[0] * L2 + b[1]
        with np.errstate(divide="ignore", invalid="ignore"):
            t = np.clip(np.log(np.abs(psi_new) / np.abs(psi_old)), -12.0, 12.0)
        occs.append(before.reshape(-1, L1, L2))
        hops.append(bond_hop(before, ia, ib, L2, a[0] == b[0]))
        targets.append(np.nan_to_num(t, nan=0.0))
        count += len(before)

    for _ in range(sweeps):
        if hasattr(sampler, "recorder"):
            sampler.recorder = record
            try:
                sampler.sweep()
            finally:
                sampler.recorder = None
        else:
            for (a, b) in sampler.bonds:
                idx, before, psi_old, psi_new = sampler._bond(a, b)
                if len(idx):
                    record(a, b, before, psi_old, psi_new)
        if log:
            log(f"training data: sweep done, {count} hops")
        if max_samples and count >= max_samples:
            break
    return np.concatenate(occs), np.concatenate(hops).astype(np.int32), np.concatenate(targets).astype(np.float32)


def train(net, occs, hops, targets, epochs=2, batch=1024, lr=2e-3, seed=0, log=None):
    gen = np.random.Generator(np.random.PCG64(seed))
    n = len(hops)
    losses = []
    t0 = time.time()
    for ep in range(epochs):
        perm = gen.permutation(n)
        for s in range(0, n - batch + 1, batch):
            sel = perm[s:s + batch]
            losses.append(net.regression_step(occs[sel], hops[sel], targets[sel], lr=lr))
        if log:
            log(f"epoch {ep}: loss {np.mean(losses[-max(1, n // batch):]):.4f} ({time.time() - t0:.1f}s)")
    return losses
