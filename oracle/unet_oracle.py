"""numpy restatement of the warm-start U-Net (include/ctbb_unet.h).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the GPU
network's forward pass and gradients.  Same parameter order and layouts as
the C ABI.  With bf16=True the values the GPU stores in BF16 (input,
convolution weights, activations, activation gradients) are rounded to BF16
at the same points; sums are FP64.  Pins (tests/test_unet_oracle.py): the
convolution against a direct loop, the max-pool routing, and the gradients
of every parameter tensor against central finite differences of its own
FP64 loss.
"""
from __future__ import annotations

import numpy as np


def bf16(a):
    """Round to the nearest BF16 (ties to even), returned as float64."""
    a = np.asarray(a, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def layers(levels, base):
    """(cin, cout, level) of every 3x3 convolution, in parameter order."""
    ch = [base << l for l in range(levels)]
    out = []
    for l in range(levels):
        out.append((8 if l == 0 else ch[l - 1], ch[l], l))
        out.append((ch[l], ch[l], l))
    for d in range(levels - 2, -1, -1):
        out.append((ch[d] + ch[d + 1], ch[d], d))
        out.append((ch[d], ch[d], d))
    return out


def unpack(params, levels, base):
    params = np.asarray(params, np.float64)
    convs, o = [], 0
    for cin, cout, lvl in layers(levels, base):
        w = params[o:o + 9 * cout * cin].reshape(9, cout, cin)
        o += 9 * cout * cin
        b = params[o:o + cout]
        o += cout
        convs.append((w, b, lvl))
    fw = params[o:o + base]
    fb = params[o + base]
    assert o + base + 1 == params.size
    return convs, fw, fb


TAPS = [(t // 3 - 1, t % 3 - 1) for t in range(9)]


def _shift(x, dy, dx):
    """y[n, i, j] = x[n, i + dy, j + dx] (zero outside)."""
    B, H, W, C = x.shape
    y = np.zeros_like(x)
    ys, ye = max(0, -dy), min(H, H - dy)
    xs, xe = max(0, -dx), min(W, W - dx)
    y[:, ys:ye, xs:xe] = x[:, ys + dy:ye + dy, xs + dx:xe + dx]
    return y


def conv(x, w, b):
    out = np.zeros(x.shape[:3] + (w.shape[1],))
    for t, (dy, dx) in enumerate(TAPS):
        out += _shift(x, dy, dx) @ w[t].T
    return out + b


def conv_dgrad(dy_, w):
    dx = np.zeros(dy_.shape[:3] + (w.shape[2],))
    for t, (dy, dxo) in enumerate(TAPS):
        dx += _shift(dy_ @ w[t], -dy, -dxo)
    return dx


def conv_wgrad(x, dy_):
    cout, cin = dy_.shape[-1], x.shape[-1]
    dw = np.zeros((9, cout, cin))
    d2 = dy_.reshape(-1, cout)
    for t, (dy, dx) in enumerate(TAPS):
        dw[t] = d2.T @ _shift(x, dy, dx).reshape(-1, cin)
    return dw, d2.sum(0)


def maxpool(x):
    B, H, W, C = x.shape
    v = x.reshape(B, H // 2, 2, W // 2, 2, C).transpose(0, 1, 3, 2, 4, 5).reshape(B, H // 2, W // 2, 4, C)
    return v.max(3), v.argmax(3)


def maxpool_bwd(d, arg, shape):
    B, H, W, C = shape
    g = np.zeros((B, H // 2, W // 2, 4, C))
    np.put_along_axis(g, arg[:, :, :, None, :], d[:, :, :, None, :], axis=3)
    return g.reshape(B, H // 2, W // 2, 2, 2, C).transpose(0, 1, 3, 2, 4, 5).reshape(B, H, W, C)


def up(x):
    return x.repeat(2, axis=1).repeat(2, axis=2)


def up_bwd(d):
    B, H, W, C = d.shape
    return d.reshape(B, H // 2, 2, W // 2, 2, C).sum((2, 4))


def forward_backward(params, inp, target, levels, base, bf16_round=True):
    """inp [B][3][H][W], target [B][H][W] -> (output [B][H][W], loss, grads)."""
    R = bf16 if bf16_round else (lambda a: np.asarray(a, np.float64))
    convs, fw, fb = unpack(params, levels, base)
    L = levels
    inp = np.asarray(inp, np.float64)
    B, _, H, W = inp.shape
    x8 = np.zeros((B, H, W, 8))
    x8[..., :3] = R(inp.transpose(0, 2, 3, 1))
    centre = np.asarray(inp[:, 1], np.float64)
    wq = [(R(w), b, lvl) for w, b, lvl in convs]
    acts = []          # (input, output) of each conv
    pools = []
    cur = x8
    for l in range(L):
        for s in range(2):
            w, b, _ = wq[2 * l + s]
            o = R(np.maximum(conv(cur, w, b), 0.0))
            acts.append((cur, o))
            cur = o
        if l < L - 1:
            p, arg = maxpool(cur)
            pools.append((p, arg, cur.shape))
            cur = p
    for pos in range(L - 1):
        lvl = L - 2 - pos
        skip = acts[2 * lvl + 1][1]
        cat = np.concatenate([skip, up(cur)], axis=-1)
        for s in range(2):
            w, b, _ = wq[2 * L + 2 * pos + s]
            inn = cat if s == 0 else cur
            o = R(np.maximum(conv(inn, w, b), 0.0))
            acts.append((inn, o))
            cur = o
    f = cur
    out = centre + fb + f @ fw
    d = out - np.asarray(target, np.float64)
    P = d.size
    loss = np.mean(d * d)
    g = 2.0 * d / P
    grads = [None] * (len(convs) * 2 + 2)
    grads[-2] = np.einsum("bhw,bhwc->c", g, f)
    grads[-1] = np.array([g.sum()])
    dcur = R(np.where(f > 0, g[..., None] * fw, 0.0))
    dskip = [None] * (L - 1)
    for pos in range(L - 2, -1, -1):
        lvl = L - 2 - pos
        ia, ib = 2 * L + 2 * pos, 2 * L + 2 * pos + 1
        # conv b
        inn, o = acts[ib]
        dw, db = conv_wgrad(inn, dcur)
        grads[2 * ib], grads[2 * ib + 1] = dw, db
        da = R(conv_dgrad(dcur, wq[ib][0]))
        da = np.where(acts[ia][1] > 0, da, 0.0)
        inn, o = acts[ia]
        dw, db = conv_wgrad(inn, da)
        grads[2 * ia], grads[2 * ia + 1] = dw, db
        dcat = R(conv_dgrad(da, wq[ia][0]))
        cs = acts[2 * lvl + 1][1].shape[-1]
        dskip[lvl] = dcat[..., :cs]
        dprod = R(up_bwd(dcat[..., cs:]))
        prod_out = acts[2 * (L - 1) + 1][1] if pos == 0 else acts[2 * L + 2 * (pos - 1) + 1][1]
        dcur = np.where(prod_out > 0, dprod, 0.0)
    dpool = None
    for l in range(L - 1, -1, -1):
        ia, ib = 2 * l, 2 * l + 1
        if l < L - 1:
            p, arg, shape = pools[l]
            db_ = R(R(maxpool_bwd(dpool, arg, shape)) + dskip[l])
            dcur = np.where(acts[ib][1] > 0, db_, 0.0)
        inn, o = acts[ib]
        dw, db = conv_wgrad(inn, dcur)
        grads[2 * ib], grads[2 * ib + 1] = dw, db
        da = R(conv_dgrad(dcur, wq[ib][0]))
        da = np.where(acts[ia][1] > 0, da, 0.0)
        inn, o = acts[ia]
        dw, db = conv_wgrad(inn, da)
        grads[2 * ia], grads[2 * ia + 1] = dw, db
        if l > 0:
            dpool = R(conv_dgrad(da, wq[ia][0]))
    flat = np.concatenate([np.ravel(a) for a in grads])
    return out, loss, flat


def init_params(levels, base, rng, final_scale=0.1):
    """Random parameters in the C-ABI order (tests only)."""
    parts = []
    for cin, cout, _ in layers(levels, base):
        w = rng.normal(0, np.sqrt(2.0 / (9 * cin)), (9, cout, cin))
        if cin == 8:
            w[:, :, 3:] = 0.0
        parts += [w.ravel(), rng.normal(0, 0.05, cout)]
    parts += [rng.normal(0, final_scale, base), rng.normal(0, 0.01, 1)]
    return np.concatenate(parts).astype(np.float32)
