"""Hop-proposal network kernels on the GPU against the CPU oracle, and the
learned Metropolis-Hastings sampler's stationary distribution."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def make_net(L1, L2, C, layers, seed=0):
    import torch

    from tnsp_amd.proposal import HopProposalNet

    net = HopProposalNet(L1, L2, channels=C, layers=layers, seed=seed)
    gen = np.random.default_rng(seed + 1)
    for b in net.b:
        b.copy_(torch.as_tensor(gen.standard_normal(b.numel()) * 0.1, dtype=torch.float32))
    for w in net.w:
        w.mul_(4.0)  # larger logits exercise the softmax
    net.sync_bf16()
    return net


def host_params(net):
    return [w.cpu().numpy().astype(np.float64) for w in net.w], [b.cpu().numpy().astype(np.float64) for b in net.b]


@pytest.mark.parametrize("L1,L2,C,layers", [(8, 8, 32, 4), (4, 5, 8, 3)])
def test_forward_and_gradients_match_oracle(L1, L2, C, layers):
    import torch

    from oracle import proposal as op

    net = make_net(L1, L2, C, layers)
    gen = np.random.default_rng(2)
    occ = (gen.random((70, L1, L2)) < 0.5).astype(np.int8)
    ws, bs = host_params(net)
    ref, acts = op.forward(occ, ws, bs)
    got = net.forward(occ).cpu().numpy()
    scale = np.max(np.abs(ref))
    assert np.max(np.abs(got - ref)) <= 1e-5 * scale
    hops = np.array([gen.choice(np.nonzero(m)[0]) for m in op.valid_hops(occ)])
    targets = gen.standard_normal(len(occ)).astype(np.float32)
    _, g = op.regression_grad(ref, hops, targets.astype(np.float64))
    dws, dbs = op.backward(acts, ws, g)
    net.forward(occ)
    dl = torch.as_tensor(g, dtype=torch.float32, device="cuda")
    grads = net.backward(dl)
    for li in range(layers):
        gw = grads[li].cpu().numpy()
        gb = grads[layers + li].cpu().numpy()
        assert np.max(np.abs(gw - dws[li])) <= 1e-4 * max(np.max(np.abs(dws[li])), 1e-12), li
        assert np.max(np.abs(gb - dbs[li])) <= 1e-4 * max(np.max(np.abs(dbs[li])), 1e-12), li


def test_hop_sampling_and_log_prob_match_oracle():
    import torch

    from oracle import proposal as op

    L1, L2 = 6, 7
    net = make_net(L1, L2, 8, 3)
    gen = np.random.default_rng(3)
    occ = (gen.random((200, L1, L2)) < 0.5).astype(np.int8)
    occ[0] = 0  # no valid hop
    u = gen.random(len(occ)).astype(np.float32)
    hops, logq = net.sample(occ, u)
    hops, logq = hops.cpu().numpy(), logq.cpu().numpy()
    ws, bs = host_params(net)
    logits, _ = op.forward(occ, ws, bs)
    lsm = op.log_softmax_hops(logits, occ)
    assert hops[0] == -1
    for b in range(1, len(occ)):
        p = np.exp(lsm[b])
        c = np.cumsum(np.where(np.isfinite(lsm[b]), p, 0.0))
        # the float32 inverse CDF may differ only within rounding of a boundary
        h = hops[b]
        assert np.isfinite(lsm[b, h])
        lo = c[h] - p[h]
        assert lo - 1e-5 <= u[b] <= c[h] + 1e-5, (b, h, u[b], lo, c[h])
        assert abs(logq[b] - lsm[b, h]) < 1e-4
    lp = net.log_prob(occ[1:], torch.as_tensor(hops[1:])).cpu().numpy()
    np.testing.assert_allclose(lp, logq[1:], atol=1e-5)


def test_adam_step_matches_numpy():
    import torch

    L1, L2 = 4, 4
    net = make_net(L1, L2, 4, 2)
    gen = np.random.default_rng(4)
    occ = (gen.random((16, L1, L2)) < 0.5).astype(np.int8)
    from oracle import proposal as op

    hops = np.array([gen.choice(np.nonzero(m)[0]) if m.any() else 0 for m in op.valid_hops(occ)])
    targets = gen.standard_normal(len(occ)).astype(np.float32)
    w0 = [p.clone() for p in net.params]
    net.regression_step(occ, hops, targets, lr=1e-2)
    for p0, p, g in zip(w0, net.params, net.grads):
        g = g.double()
        m = 0.1 * g
        v = 0.001 * g * g
        ref = p0.double() - 1e-2 * (m / 0.1) / (torch.sqrt(v / 0.001) + 1e-8)
        assert torch.max(torch.abs(p.double() - ref)) < 1e-5


def test_learned_sampler_samples_psi_squared():
    """Metropolis-Hastings with the network's proposal keeps |psi|^2 (here a
    briefly trained network, so the proposal is far from uniform)."""
    from test_sampler import enumerate_distribution, small_lattice
    from tnsp_amd import peps as pp
    from tnsp_amd.engine import AmplitudeEngine
    from tnsp_amd.proposal import LearnedSampler, collect_training_data, train
    from tnsp_amd.sampler import TVSampler

    spec, storages, _ = small_lattice(2, 3, 3, 24)
    eng = AmplitudeEngine(pp.Lattice(spec, storages), chunk=512)
    cfgs, p = enumerate_distribution(eng, 2, 3, 3)
    keys = {c.tobytes(): i for i, c in enumerate(cfgs)}
    start = np.repeat(cfgs[np.argmax(p)][None], 256, axis=0)
    plain = TVSampler(eng, start, seed=1)
    occs, hops, targets = collect_training_data(plain, sweeps=4)
    net = make_net(2, 3, 8, 3, seed=5)
    losses = train(net, occs, hops, targets, epochs=20, batch=128, lr=3e-3)
    assert losses[-1] < losses[0]
    # chains start in |psi|^2 itself: a correct Metropolis-Hastings step keeps
    # them there whatever the proposal's mixing speed
    gen = np.random.default_rng(9)
    s = LearnedSampler(eng, net, cfgs[gen.choice(len(cfgs), 2000, p=p)], seed=2)
    counts = np.zeros(len(cfgs))
    for step in range(40):
        s.step()
        for c in s.configs:
            counts[keys[c.tobytes()]] += 1
    assert s.accepted > 0.05 * s.proposals
    emp = counts / counts.sum()
    assert 0.5 * np.abs(emp - p).sum() < 0.04


def test_learned_sampler_acceptance_rule_c4():
    """Each Metropolis-Hastings decision on the 8x8 D=8 Dc=64 PEPS follows
    min(1, |psi(s')/psi(s)|^2 q(s|s') / q(s'|s)), with psi(s') recomputed by
    a separate engine and both proposal probabilities rescored by the
    network (the reverse hop's and the forward hop's log q)."""
    from tnsp_amd import peps as pp
    from tnsp_amd.engine import AmplitudeEngine
    from tnsp_amd.proposal import HopProposalNet, LearnedSampler, apply_hops

    spec = pp.C4
    lat = pp.Lattice(spec)
    eng = AmplitudeEngine(lat, chunk=64, lanes=2)
    cand = pp.sample_configurations(spec, 96, seed=31)
    start = cand[eng.evaluate(cand) != 0][:48]
    net = HopProposalNet(spec.L1, spec.L2, channels=32, layers=4, seed=6)
    s = LearnedSampler(eng, net, start, seed=8)
    check = AmplitudeEngine(lat, chunk=16, lanes=1)
    moved = 0
    for step in range(3):
        before, psi_before = s.configs.copy(), s.psi.copy()
        s.step()
        rec = s.last
        ok = rec["valid"]
        assert ok.all()  # half filling: every configuration has a valid hop
        hops = rec["hops"]
        new, rev = apply_hops(before, hops, spec.L1, spec.L2)
        assert np.array_equal(new, rec["proposed"])
        psi_new = check.evaluate(new)
        np.testing.assert_allclose(rec["psi_new"], psi_new, rtol=1e-10, atol=0)
        lf = net.log_prob(before, torch.as_tensor(hops.astype(np.int32))).cpu().numpy().astype(np.float64)
        lr = net.log_prob(new, torch.as_tensor(rev.astype(np.int32))).cpu().numpy().astype(np.float64)
        nz = psi_new != 0
        expect = 2 * (np.log(np.abs(psi_new[nz])) - np.log(np.abs(psi_before[nz]))) + lr[nz] - lf[nz]
        np.testing.assert_allclose(rec["log_a"][nz], expect, rtol=0, atol=1e-5)
        acc = nz & (np.log(np.maximum(rec["uniforms"], 1e-300)) < rec["log_a"])
        assert np.array_equal(acc, rec["accepted"])
        after = np.where(acc[:, None, None], new, before)
        assert np.array_equal(s.configs, after)
        np.testing.assert_array_equal(s.psi, np.where(acc, psi_new, psi_before))
        moved += int(acc.sum())
    assert moved > 0
