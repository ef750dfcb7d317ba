"""Hop-proposal network: the CPU oracle's gradients against finite
differences, and the host-side hop bookkeeping against the oracle."""
import numpy as np


def small_net(gen, L1=4, L2=5, C=6, layers=3):
    dims = [1] + [C] * (layers - 1) + [4]
    ws = [gen.standard_normal((dims[i + 1], dims[i], 3, 3)) * 0.4 for i in range(layers)]
    bs = [gen.standard_normal(dims[i + 1]) * 0.1 for i in range(layers)]
    occ = (gen.random((3, L1, L2)) < 0.5).astype(np.int8)
    return ws, bs, occ


def test_oracle_gradients_match_finite_differences():
    from oracle import proposal as op

    gen = np.random.default_rng(0)
    ws, bs, occ = small_net(gen)
    # bf16 rounding is piecewise constant: differentiate at bf16 points with
    # steps that stay inside one rounding cell, i.e. w.r.t. the rounded
    # weights themselves (the straight-through gradient the kernels compute)
    ws = [op.bf16_round(w).astype(np.float64) for w in ws]
    logits, acts = op.forward(occ, ws, bs)
    hops = gen.integers(0, logits.shape[1], len(occ))
    targets = gen.standard_normal(len(occ))
    loss, g = op.regression_grad(logits, hops, targets)
    dws, dbs = op.backward(acts, ws, g)

    def loss_at(ws2, bs2):
        x = np.asarray(occ, dtype=np.float64)[:, None]
        for li, (w, b) in enumerate(zip(ws2, bs2)):
            x = op.conv_forward(x, w, b)
            if li + 1 < len(ws2):
                x = np.maximum(x, 0)
        lg = x.reshape(len(occ), -1)
        return np.mean((lg[np.arange(len(occ)), hops] - targets) ** 2)

    eps = 1e-6
    for li in range(len(ws)):
        for _ in range(6):
            idx = tuple(gen.integers(0, s) for s in ws[li].shape)
            wp = [w.copy() for w in ws]
            wm = [w.copy() for w in ws]
            wp[li][idx] += eps
            wm[li][idx] -= eps
            fd = (loss_at(wp, bs) - loss_at(wm, bs)) / (2 * eps)
            assert abs(fd - dws[li][idx]) < 1e-6 + 1e-5 * abs(fd), (li, idx, fd, dws[li][idx])
        j = int(gen.integers(0, len(bs[li])))
        bp = [b.copy() for b in bs]
        bm = [b.copy() for b in bs]
        bp[li][j] += eps
        bm[li][j] -= eps
        fd = (loss_at(ws, bp) - loss_at(ws, bm)) / (2 * eps)
        assert abs(fd - dbs[li][j]) < 1e-6 + 1e-5 * abs(fd)


def test_bf16_round_ties_to_even():
    from oracle import proposal as op

    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -9, -2.5], dtype=np.float32)
    # 1 + 2^-8 is halfway between 1 and 1 + 2^-7: ties to the even mantissa (1)
    np.testing.assert_array_equal(op.bf16_round(x), np.array([1.0, 1.0, 1.0 + 2 ** -6, 1.0, -2.5], dtype=np.float32))


def test_apply_hops_and_bond_hops_match_oracle():
    from oracle import proposal as op
    from tnsp_amd.proposal import apply_hops, bond_hop

    gen = np.random.default_rng(1)
    L1, L2 = 4, 5
    occ = (gen.random((40, L1, L2)) < 0.5).astype(np.int8)
    mask = op.valid_hops(occ)
    hops = np.array([gen.choice(np.nonzero(m)[0]) for m in mask])
    new, rev = apply_hops(occ, hops, L1, L2)
    new_o, rev_o = op.apply_hop(occ, hops)
    np.testing.assert_array_equal(new, new_o)
    np.testing.assert_array_equal(rev, rev_o)
    # the reverse hop is valid on the new configuration and undoes the hop
    assert op.valid_hops(new)[np.arange(len(rev)), rev].all()
    back, _ = apply_hops(new, rev, L1, L2)
    np.testing.assert_array_equal(back, occ)
    # bond hops: the particle crossing a bond, in both orientations
    flat = occ.reshape(len(occ), -1)
    for (ia, ib, horiz) in [(6, 7, True), (2, 7, False)]:
        sel = flat[:, ia] != flat[:, ib]
        h = bond_hop(flat[sel], ia, ib, L2, horiz)
        assert mask[np.nonzero(sel)[0], h].all()
