"""GPU U-Net (BF16 MFMA convolutions, include/ctbb_unet.h) against the numpy
oracle that rounds to BF16 at the same points.  Tolerances: outputs and
losses within 2e-2 relative, gradients within 8e-2 relative norm per
parameter tensor (FP32 accumulation order differs, so a few BF16 roundings
of stored activation gradients flip) and cosine > 0.999 overall."""
import numpy as np
import pytest

import unet_oracle as U

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu(ct):
    ct.init(0)
    return ct


def _tensor_slices(L, base):
    out, o = [], 0
    for cin, cout, _ in U.layers(L, base):
        out.append(("w", cin, cout, slice(o, o + 9 * cin * cout)))
        o += 9 * cin * cout
        out.append(("b", cin, cout, slice(o, o + cout)))
        o += cout
    out.append(("fw", base, 1, slice(o, o + base)))
    out.append(("fb", 1, 1, slice(o + base, o + base + 1)))
    return out


@pytest.mark.parametrize("L,base,B,H,W", [(3, 8, 2, 16, 24), (4, 32, 1, 32, 40),
                                          (4, 32, 2, 64, 64)])
def test_forward_and_gradients_match_oracle(gpu, L, base, B, H, W):
    print(f"[unet] L={L} base={base} batch={B} {H}x{W}", flush=True)
    rng = np.random.default_rng(5)
    net = gpu.UNet(L, base, B, H, W, seed=3)
    p = U.init_params(L, base, rng, final_scale=0.5)
    net.set_params(p)
    assert np.array_equal(net.params(), p)
    inp = rng.uniform(0, 1, (B, 3, H, W)).astype(np.float32)
    tgt = rng.uniform(0, 1, (B, H, W)).astype(np.float32)
    loss_g, g_gpu = net.grad(inp, tgt, want_grads=True)
    out_g = net.infer(inp)
    out_o, loss_o, g_o = U.forward_backward(p, inp, tgt, L, base, bf16_round=True)
    assert np.abs(out_g - out_o).max() <= 2e-2 * np.abs(out_o).max()
    assert abs(loss_g - loss_o) <= 2e-2 * loss_o
    worst = 0.0
    for kind, cin, cout, sl in _tensor_slices(L, base):
        a, b = g_gpu[sl].astype(np.float64), g_o[sl]
        nb = np.linalg.norm(b)
        if nb == 0:
            assert np.linalg.norm(a) == 0
            continue
        worst = max(worst, np.linalg.norm(a - b) / nb)
    cos = g_gpu @ g_o / (np.linalg.norm(g_gpu) * np.linalg.norm(g_o))
    print(f"[unet] worst tensor rel err {worst:.3e}, cosine {cos:.6f}", flush=True)
    assert worst < 8e-2
    assert cos > 0.999


def test_adam_step_matches_numpy(gpu):
    L, base, B, H, W = 2, 8, 2, 8, 8
    rng = np.random.default_rng(6)
    net = gpu.UNet(L, base, B, H, W, seed=4)
    p0 = net.params().astype(np.float64)
    inp = rng.uniform(0, 1, (B, 3, H, W)).astype(np.float32)
    tgt = rng.uniform(0, 1, (B, H, W)).astype(np.float32)
    m = np.zeros_like(p0)
    v = np.zeros_like(p0)
    p = p0.copy()
    for t in range(1, 4):
        net.set_params(p.astype(np.float32)) if t == 1 else None
        _, g = net.grad(inp, tgt, want_grads=True)
        net.adam_step(1e-2, 0.9, 0.999, 1e-8)
        g = g.astype(np.float64)
        m = 0.9 * m + 0.1 * g
        v = 0.999 * v + 0.001 * g * g
        p = p - 1e-2 * (m / (1 - 0.9 ** t)) / (np.sqrt(v / (1 - 0.999 ** t)) + 1e-8)
        got = net.params().astype(np.float64)
        assert np.abs(got - p).max() <= 1e-5 * max(1.0, np.abs(p).max())
        p = got


def test_training_reduces_the_loss(gpu):
    # denoising: three noisy copies of a smooth image -> the image
    L, base, B, H, W = 3, 16, 4, 32, 32
    rng = np.random.default_rng(7)
    net = gpu.UNet(L, base, B, H, W, seed=5)
    yy, xx = np.mgrid[0:H, 0:W] / H
    losses = []
    for step in range(60):
        cx, cy = rng.uniform(0.3, 0.7, 2)
        img = np.exp(-((xx - cx) ** 2 + (yy - cy) ** 2) / 0.05)
        tgt = np.repeat(img[None], B, 0).astype(np.float32)
        inp = (tgt[:, None] + rng.normal(0, 0.2, (B, 3, H, W))).astype(np.float32)
        losses.append(net.grad(inp, tgt))
        net.adam_step(2e-3)
    first, last = np.mean(losses[:5]), np.mean(losses[-5:])
    print(f"[unet] loss {first:.4e} -> {last:.4e}", flush=True)
    assert last < 0.5 * first
