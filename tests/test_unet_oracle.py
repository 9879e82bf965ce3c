"""CPU checks of the U-Net oracle (oracle/unet_oracle.py): its convolution
against a direct loop, its maxpool routing, and its gradients against central
finite differences of its own FP64 loss (no BF16 rounding); and the
network's input stacking in the warm-start host code."""
import numpy as np

import unet_oracle as U


def test_conv_matches_direct_loop():
    rng = np.random.default_rng(0)
    x = rng.normal(size=(2, 5, 6, 3))
    w = rng.normal(size=(9, 4, 3))
    b = rng.normal(size=4)
    y = U.conv(x, w, b)
    ref = np.zeros((2, 5, 6, 4))
    for n in range(2):
        for i in range(5):
            for j in range(6):
                s = b.copy()
                for t in range(9):
                    ii, jj = i + t // 3 - 1, j + t % 3 - 1
                    if 0 <= ii < 5 and 0 <= jj < 6:
                        s += w[t] @ x[n, ii, jj]
                ref[n, i, j] = s
    assert np.allclose(y, ref)
    # the input-gradient and weight-gradient formulas are the adjoints
    dy = rng.normal(size=(2, 5, 6, 4))
    dx = U.conv_dgrad(dy, w)
    dw, db = U.conv_wgrad(x, dy)
    lhs = np.sum(dy * (U.conv(x, w, np.zeros(4))))
    assert np.isclose(lhs, np.sum(dx * x))
    assert np.isclose(lhs, np.sum(dw * w))
    assert np.allclose(db, dy.sum((0, 1, 2)))


def test_maxpool_routes_to_first_maximum():
    x = np.zeros((1, 2, 2, 1))
    x[0, 0, 1, 0] = x[0, 1, 0, 0] = 3.0
    p, arg = U.maxpool(x)
    assert p[0, 0, 0, 0] == 3.0 and arg[0, 0, 0, 0] == 1
    g = U.maxpool_bwd(np.ones((1, 1, 1, 1)), arg, x.shape)
    assert g[0, 0, 1, 0] == 1.0 and g.sum() == 1.0


def test_oracle_gradients_match_finite_differences():
    rng = np.random.default_rng(1)
    L, base, B, H, W = 2, 8, 2, 8, 12
    p = U.init_params(L, base, rng, final_scale=0.5).astype(np.float64)
    inp = rng.uniform(0.0, 1.0, (B, 3, H, W))
    tgt = rng.uniform(0.0, 1.0, (B, H, W))
    _, loss, g = U.forward_backward(p, inp, tgt, L, base, bf16_round=False)
    idx = rng.choice(p.size, 60, replace=False)
    idx = np.concatenate([idx, [p.size - 1, p.size - 2]])
    for i in idx:
        if U.layers(L, base)[0][0] == 8 and i < 9 * base * 8 and (i % 8) >= 3:
            continue   # padding channels of the input layer: zero inputs
        h = 1e-6
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        lp = U.forward_backward(pp, inp, tgt, L, base, bf16_round=False)[1]
        lm = U.forward_backward(pm, inp, tgt, L, base, bf16_round=False)[1]
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 + 1e-4 * abs(fd), (i, fd, g[i])


def test_slice_stacks_repeat_the_end_slices():
    """The network input of slice k is (k-1, k, k+1), the end slices
    repeated (host logic of the learned warm start)."""
    from ctbb_icd_amd.warmstart import slice_stacks
    vol = np.arange(4 * 2 * 3, dtype=np.float32).reshape(4, 2, 3)
    st = slice_stacks(vol)
    assert st.shape == (4, 3, 2, 3)
    for k in range(4):
        assert np.array_equal(st[k, 1], vol[k])
        assert np.array_equal(st[k, 0], vol[max(k - 1, 0)])
        assert np.array_equal(st[k, 2], vol[min(k + 1, 3)])
