"""CPU oracle of the learned pressure guess (U-Net) -- TEST INFRASTRUCTURE ONLY.

Only tests/ may import this module; nothing in scihpc_amd/ uses it.  SciHPC
has no learned component, so there is no reference file to cite: the network
is defined by the build (4 levels, 16/32/64/128 channels, two 3x3
convolutions + ReLU per level, 2x2 average pooling, nearest upsampling with
the skip concatenated after the upsampled channels, 1x1 head) and this file
restates it in numpy, forward and backward, in float64.

With ``emulate=True`` every tensor the GPU stores in BF16 is rounded to
BF16 (round to nearest even) at the same place, so the GPU result differs
from the oracle only by FP32 accumulation order.

Pinning (tests/test_unet_oracle.py): the 3x3 convolution against a
brute-force loop and torch.nn.functional.conv2d; every parameter gradient
against central finite differences; the whole forward/backward against
torch CPU autograd of the same network in float64.

Parameter layout (identical to the library's): for each of the 14 3x3
layers, w [cout][3][3][cin] then b [cout]; then the head's 16 weights and
1 bias.  Activations are NHWC.
"""
from __future__ import annotations

import numpy as np

# (cin, cout, level) of the 14 3x3 convolutions, in execution order of the
# forward pass: encoder levels 0..3, then decoder levels 2, 1, 0
SPEC = [(16, 16, 0), (16, 16, 0), (16, 32, 1), (32, 32, 1), (32, 64, 2), (64, 64, 2), (64, 128, 3),
        (128, 128, 3), (192, 64, 2), (64, 64, 2), (96, 32, 1), (32, 32, 1), (48, 16, 0), (16, 16, 0)]
HEAD_C = 16


def layout():
    offs, o = [], 0
    for cin, cout, _ in SPEC:
        offs.append(o)
        o += cout * 9 * cin + cout
    return offs, o, o + HEAD_C + 1  # layer offsets, head offset, total


def bf16(a):
    """Round float values to BF16 (nearest even), returned as float64."""
    f = np.ascontiguousarray(a, np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(params):
    offs, head, _ = layout()
    p = np.asarray(params, np.float64)
    layers = []
    for (cin, cout, _), o in zip(SPEC, offs):
        w = p[o:o + cout * 9 * cin].reshape(cout, 3, 3, cin)
        b = p[o + cout * 9 * cin:o + cout * 9 * cin + cout]
        layers.append((w, b))
    return layers, p[head:head + HEAD_C], p[head + HEAD_C]


def _shift(x, dy, dx):
    """s[n, y, x] = x[n, y + dy, x + dx] with zeros outside."""
    n, h, w, c = x.shape
    out = np.zeros_like(x)
    ys, yd = (slice(dy, h), slice(0, h - dy)) if dy >= 0 else (slice(0, h + dy), slice(-dy, h))
    xs, xd = (slice(dx, w), slice(0, w - dx)) if dx >= 0 else (slice(0, w + dx), slice(-dx, w))
    out[:, yd, xd, :] = x[:, ys, xs, :]
    return out


def conv3x3(x, w, b=None):
    """y[n,y,x,o] = sum_{ty,tx,c} x[n, y+ty-1, x+tx-1, c] w[o, ty, tx, c] (+ b[o]), zero padding."""
    n, h, wd, _ = x.shape
    y = np.zeros((n, h, wd, w.shape[0]))
    for ty in range(3):
        for tx in range(3):
            y += _shift(x, ty - 1, tx - 1) @ w[:, ty, tx, :].T
    if b is not None:
        y += b
    return y


def conv3x3_bwd(x, gz, w):
    """(dx, dw, db) of conv3x3 for the gradient gz of its output."""
    dw = np.zeros_like(w)
    dx = np.zeros_like(x)
    for ty in range(3):
        for tx in range(3):
            xs = _shift(x, ty - 1, tx - 1)
            dw[:, ty, tx, :] = np.einsum("nhwo,nhwc->oc", gz, xs)
            dx += _shift(gz @ w[:, ty, tx, :], 1 - ty, 1 - tx)
    return dx, dw, gz.sum(axis=(0, 1, 2))


def pool2(x):
    n, h, w, c = x.shape
    return x.reshape(n, h // 2, 2, w // 2, 2, c).mean(axis=(2, 4))


def up2(x):
    return x.repeat(2, axis=1).repeat(2, axis=2)


def forward(params, inp, emulate=True):
    """inp [n, H, W, 2] features -> (out [n, H, W], cache)."""
    R = bf16 if emulate else (lambda a: np.asarray(a, np.float64))
    layers, hw, hb = unpack(params)
    if emulate:
        layers = [(bf16(w), b) for w, b in layers]
    n, H, W, _ = inp.shape
    x0 = np.zeros((n, H, W, 16))
    x0[..., :2] = R(inp)
    acts = {"in0": x0}

    def conv(i, x):
        w, b = layers[i]
        return R(np.maximum(conv3x3(x, w, b), 0.0))

    t0a = conv(0, x0)
    s0 = conv(1, t0a)
    p1 = R(pool2(s0))
    t1a = conv(2, p1)
    s1 = conv(3, t1a)
    p2 = R(pool2(s1))
    t2a = conv(4, p2)
    s2 = conv(5, t2a)
    p3 = R(pool2(s2))
    t3a = conv(6, p3)
    t3b = conv(7, t3a)
    cat2 = np.concatenate([up2(t3b), s2], axis=3)
    t2c = conv(8, cat2)
    t2d = conv(9, t2c)
    cat1 = np.concatenate([up2(t2d), s1], axis=3)
    t1c = conv(10, cat1)
    t1d = conv(11, t1c)
    cat0 = np.concatenate([up2(t1d), s0], axis=3)
    t0c = conv(12, cat0)
    t0d = conv(13, t0c)
    out = t0d @ hw + hb
    acts.update(t0a=t0a, s0=s0, p1=p1, t1a=t1a, s1=s1, p2=p2, t2a=t2a, s2=s2, p3=p3, t3a=t3a, t3b=t3b,
                cat2=cat2, t2c=t2c, t2d=t2d, cat1=cat1, t1c=t1c, t1d=t1d, cat0=cat0, t0c=t0c, t0d=t0d)
    return out, (layers, hw, hb, acts)


def energy(v, faces):
    """Per image sum over faces c_f (v_i - v_j)^2; faces [n, H, W, 2] hold the
    coefficient toward x+1 and y+1 (zero on walls)."""
    dx = v[:, :, :-1] - v[:, :, 1:]
    dy = v[:, :-1, :] - v[:, 1:, :]
    return np.sum(faces[:, :, :-1, 0] * dx * dx, axis=(1, 2)) + np.sum(faces[:, :-1, :, 1] * dy * dy, axis=(1, 2))


def energy_grad(e, faces):
    """d/de of sum_f c_f (e_i - e_j)^2 per image (a weighted graph Laplacian)."""
    g = np.zeros_like(e)
    cx, cy = faces[:, :, :-1, 0], faces[:, :-1, :, 1]
    dx = cx * (e[:, :, :-1] - e[:, :, 1:])
    dy = cy * (e[:, :-1, :] - e[:, 1:, :])
    g[:, :, :-1] += dx
    g[:, :, 1:] -= dx
    g[:, :-1, :] += dy
    g[:, 1:, :] -= dy
    return 2.0 * g


def loss_and_grads(params, inp, target, emulate=True, faces=None):
    """Loss over the batch and its gradient w.r.t. params: the mean squared
    error, or with faces the mean over images of E(out - target) / E(target)."""
    R = bf16 if emulate else (lambda a: np.asarray(a, np.float64))
    out, (layers, hw, hb, a) = forward(params, inp, emulate)
    npix = out.size
    d = out - target
    if faces is None:
        loss = float(np.sum(d * d) / npix)
        dout = 2.0 * d / npix
    else:
        n = out.shape[0]
        et = energy(target, faces)
        loss = float(np.sum(energy(d, faces) / et) / n)
        dout = energy_grad(d, faces) / (et[:, None, None] * n)
    if emulate:
        dout = dout.astype(np.float32).astype(np.float64)
    offs, head, total = layout()
    g = np.zeros(total)
    g[head:head + HEAD_C] = np.einsum("nhw,nhwc->c", dout, a["t0d"])
    g[head + HEAD_C] = dout.sum()
    gy = R(dout[..., None] * hw)

    def back(i, x, y, gy, want_dx=True):
        w, _ = layers[i]
        gz = gy * (y > 0)
        dx, dw, db = conv3x3_bwd(x, gz, w)
        cin, cout, _ = SPEC[i]
        o = offs[i]
        g[o:o + cout * 9 * cin] = dw.ravel()
        g[o + cout * 9 * cin:o + cout * 9 * cin + cout] = db
        return R(dx) if want_dx else None

    def up_bwd(gf):
        n, h, w, c = gf.shape
        return R(gf.reshape(n, h // 2, 2, w // 2, 2, c).sum(axis=(2, 4)))

    g_t0c = back(13, a["t0c"], a["t0d"], gy)
    g_cat0 = back(12, a["cat0"], a["t0c"], g_t0c)
    g_t1d = up_bwd(g_cat0[..., :32])
    g_t1c = back(11, a["t1c"], a["t1d"], g_t1d)
    g_cat1 = back(10, a["cat1"], a["t1c"], g_t1c)
    g_t2d = up_bwd(g_cat1[..., :64])
    g_t2c = back(9, a["t2c"], a["t2d"], g_t2d)
    g_cat2 = back(8, a["cat2"], a["t2c"], g_t2c)
    g_t3b = up_bwd(g_cat2[..., :128])
    g_t3a = back(7, a["t3a"], a["t3b"], g_t3b)
    g_p3 = back(6, a["p3"], a["t3a"], g_t3a)
    g_s2 = R(g_cat2[..., 128:] + 0.25 * up2(g_p3))
    g_t2a = back(5, a["t2a"], a["s2"], g_s2)
    g_p2 = back(4, a["p2"], a["t2a"], g_t2a)
    g_s1 = R(g_cat1[..., 64:] + 0.25 * up2(g_p2))
    g_t1a = back(3, a["t1a"], a["s1"], g_s1)
    g_p1 = back(2, a["p1"], a["t1a"], g_t1a)
    g_s0 = R(g_cat0[..., 32:] + 0.25 * up2(g_p1))
    g_t0a = back(1, a["t0a"], a["s0"], g_s0)
    back(0, a["in0"], a["t0a"], g_t0a, want_dx=False)
    return loss, g, out


def init_params(seed=0):
    """Random parameters of the same family the library draws (He-normal)."""
    rng = np.random.default_rng(seed)
    offs, head, total = layout()
    p = np.zeros(total)
    for i, ((cin, cout, _), o) in enumerate(zip(SPEC, offs)):
        fan = 9 * (2 if i == 0 else cin)
        w = rng.standard_normal((cout, 3, 3, cin)) * np.sqrt(2.0 / fan)
        if i == 0:
            w[..., 2:] = 0.0
        p[o:o + cout * 9 * cin] = w.ravel()
        p[o + cout * 9 * cin:o + cout * 9 * cin + cout] = 0.05 * rng.standard_normal(cout)
    p[head:head + HEAD_C] = 0.25 * rng.standard_normal(HEAD_C)
    p[head + HEAD_C] = 0.1
    return p
