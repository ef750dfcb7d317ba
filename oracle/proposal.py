"""CPU oracle for the hop-proposal network (no reference counterpart).

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

The network the GPU kernels implement (tnsp_amd/csrc/proposal.hip), in
float64 numpy: 3x3 convolutions with zero padding (NCHW), ReLU between
layers, weights rounded to bf16 (round to nearest even) before use, the
masked softmax over the 4 H W nearest-neighbour hops (hop = d H W + site,
d = 0 up, 1 down, 2 left, 3 right; valid when the site is filled and its
neighbour exists and is empty), the squared-error loss of one hop logit per
configuration, and their gradients.  The gradients are pinned by finite
differences in the test-suite.
"""
import numpy as np


def bf16_round(x):
    """float32 -> nearest bf16 (ties to even), returned as float32 values."""
    u = np.asarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def conv_forward(x, w, b):
    """y[b, o, i, j] = b[o] + sum_{c, ky, kx} w[o, c, ky, kx] x[b, c, i+ky-1, j+kx-1]."""
    B, C, H, W = x.shape
    xp = np.zeros((B, C, H + 2, W + 2))
    xp[:, :, 1:-1, 1:-1] = x
    y = np.zeros((B, w.shape[0], H, W)) + b[None, :, None, None]
    for ky in range(3):
        for kx in range(3):
            y += np.einsum("oc,bcij->boij", w[:, :, ky, kx], xp[:, :, ky:ky + H, kx:kx + W])
    return y


def conv_backward(x, w, dy):
    """(dx, dw, db) of conv_forward."""
    B, C, H, W = x.shape
    xp = np.zeros((B, C, H + 2, W + 2))
    xp[:, :, 1:-1, 1:-1] = x
    dxp = np.zeros_like(xp)
    dw = np.zeros(w.shape)
    for ky in range(3):
        for kx in range(3):
            dw[:, :, ky, kx] = np.einsum("boij,bcij->oc", dy, xp[:, :, ky:ky + H, kx:kx + W])
            dxp[:, :, ky:ky + H, kx:kx + W] += np.einsum("oc,boij->bcij", w[:, :, ky, kx], dy)
    return dxp[:, :, 1:-1, 1:-1], dw, dy.sum(axis=(0, 2, 3))


def forward(occ, weights, biases):
    """Logits (B, 4 H W) and the per-layer activations."""
    x = np.asarray(occ, dtype=np.float64)[:, None]
    acts = [x]
    for li, (w, b) in enumerate(zip(weights, biases)):
        y = conv_forward(acts[-1], bf16_round(w).astype(np.float64), b)
        if li + 1 < len(weights):
            y = np.maximum(y, 0)
        acts.append(y)
    return acts[-1].reshape(len(x), -1), acts


def backward(acts, weights, dlogits):
    """Gradients (dW list, db list) of a loss with d loss / d logits."""
    dy = dlogits.reshape(acts[-1].shape)
    dws, dbs = [None] * len(weights), [None] * len(weights)
    for li in reversed(range(len(weights))):
        if li + 1 < len(weights):
            dy = dy * (acts[li + 1] > 0)
        dx, dw, db = conv_backward(acts[li], bf16_round(weights[li]).astype(np.float64), dy)
        dws[li], dbs[li] = dw, db
        dy = dx
    return dws, dbs


def valid_hops(occ):
    """Boolean (B, 4 H W) mask of the valid hops and their target sites."""
    occ = np.asarray(occ)
    B, H, W = occ.shape
    mask = np.zeros((B, 4, H, W), dtype=bool)
    f = occ == 1
    e = occ == 0
    mask[:, 0, 1:, :] = f[:, 1:, :] & e[:, :-1, :]
    mask[:, 1, :-1, :] = f[:, :-1, :] & e[:, 1:, :]
    mask[:, 2, :, 1:] = f[:, :, 1:] & e[:, :, :-1]
    mask[:, 3, :, :-1] = f[:, :, :-1] & e[:, :, 1:]
    return mask.reshape(B, -1)


def log_softmax_hops(logits, occ):
    m = valid_hops(occ)
    lg = np.where(m, logits, -np.inf)
    mx = lg.max(axis=1, keepdims=True)
    return lg - mx - np.log(np.exp(lg - mx).sum(axis=1, keepdims=True))


def apply_hop(occ, hop):
    """Configuration after the hop and the reverse hop index."""
    occ = np.array(occ, copy=True)
    B, H, W = occ.shape
    HW = H * W
    rev = np.zeros(B, dtype=np.int64)
    for b in range(B):
        d, s = divmod(int(hop[b]), HW)
        i, j = divmod(s, W)
        ti, tj = i + (d == 1) - (d == 0), j + (d == 3) - (d == 2)
        occ[b, i, j], occ[b, ti, tj] = 0, 1
        rev[b] = (d ^ 1) * HW + ti * W + tj
    return occ, rev


def regression_grad(logits, hops, targets):
    """loss = mean_b (logit[b, hop_b] - target_b)^2 and its logit gradient."""
    B = len(hops)
    d = logits[np.arange(B), hops] - targets
    g = np.zeros_like(logits)
    g[np.arange(B), hops] = 2 * d / B
    return float(np.mean(d * d)), g
