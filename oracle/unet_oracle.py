"""CPU oracle of the learned pressure guess (U-Net) -- TEST INFRASTRUCTURE ONLY.

Only tests/ may import this module; nothing in scihpc_amd/ uses it.  SciHPC
has no learned component, so there is no reference file to cite: the network
is defined by the build (4 levels, 16/32/64/128 channels, two 3x3 (2-D) or
3x3x3 (3-D) convolutions + ReLU per level, 2x2(x2) average pooling, nearest
upsampling with the skip concatenated after the upsampled channels, 1x1
head) and this file restates it in numpy, forward and backward, in float64.
The dimension follows the input: [n, H, W, 2] or [n, D, H, W, 2].

With ``emulate=True`` every tensor the GPU stores in BF16 is rounded to
BF16 (round to nearest even) at the same place, so the GPU result differs
from the oracle only by FP32 accumulation order.

Pinning (tests/test_unet_oracle.py): the 3x3 and 3x3x3 convolutions against
brute-force loops and torch.nn.functional.conv2d/conv3d (forward and
backward); every parameter gradient against central finite differences;
the whole forward/backward against torch CPU autograd of the same network
in float64 (2-D with the mean squared error, 3-D with the energy loss).

Parameter layout (identical to the library's): for each of the 14
convolutions, w [cout][3][3](3)[cin] then b [cout]; then the head's 16
weights and 1 bias.  Activations are N(D)HWC.
"""
from __future__ import annotations

import itertools

import numpy as np

# (cin, cout, level) of the 14 3x3 convolutions, in execution order of the
# forward pass: encoder levels 0..3, then decoder levels 2, 1, 0
SPEC = [(16, 16, 0), (16, 16, 0), (16, 32, 1), (32, 32, 1), (32, 64, 2), (64, 64, 2), (64, 128, 3),
        (128, 128, 3), (192, 64, 2), (64, 64, 2), (96, 32, 1), (32, 32, 1), (48, 16, 0), (16, 16, 0)]
HEAD_C = 16


def taps(ndim=2):
    return 3 ** ndim


def layout(ndim=2):
    offs, o = [], 0
    for cin, cout, _ in SPEC:
        offs.append(o)
        o += cout * taps(ndim) * cin + cout
    return offs, o, o + HEAD_C + 1  # layer offsets, head offset, total


def bf16(a):
    """Round float values to BF16 (nearest even), returned as float64."""
    f = np.ascontiguousarray(a, np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(params, ndim=2):
    """[(w [cout, 3, 3(, 3), cin], b [cout])], head weights, head bias."""
    offs, head, _ = layout(ndim)
    p = np.asarray(params, np.float64)
    layers = []
    for (cin, cout, _), o in zip(SPEC, offs):
        nw = cout * taps(ndim) * cin
        layers.append((p[o:o + nw].reshape((cout,) + (3,) * ndim + (cin,)), p[o + nw:o + nw + cout]))
    return layers, p[head:head + HEAD_C], p[head + HEAD_C]


def _shift(x, *d):
    """s[n, ..., i, ...] = x[n, ..., i + d, ...] over the spatial axes 1..ndim,
    zeros outside."""
    out = np.zeros_like(x)
    src, dst = [slice(None)], [slice(None)]
    for ax, dd in enumerate(d):
        m = x.shape[1 + ax]
        if dd >= 0:
            src.append(slice(dd, m))
            dst.append(slice(0, m - dd))
        else:
            src.append(slice(0, m + dd))
            dst.append(slice(-dd, m))
    out[tuple(dst)] = x[tuple(src)]
    return out


def _tap_offsets(ndim):
    return list(itertools.product(range(3), repeat=ndim))


def conv3x3(x, w, b=None):
    """y[n, p, o] = sum_{t, c} x[n, p + t - 1, c] w[o, t, c] (+ b[o]) over the
    3^ndim taps t (z, y, x order), zero padding; x [n, (D,) H, W, c]."""
    ndim = x.ndim - 2
    y = np.zeros(x.shape[:-1] + (w.shape[0],))
    for t in _tap_offsets(ndim):
        y += _shift(x, *[v - 1 for v in t]) @ w[(slice(None),) + t + (slice(None),)].T
    if b is not None:
        y += b
    return y


def conv3x3_bwd(x, gz, w):
    """(dx, dw, db) of conv3x3 for the gradient gz of its output."""
    ndim = x.ndim - 2
    dw = np.zeros_like(w)
    dx = np.zeros_like(x)
    sp = "dhw"[3 - ndim:]
    for t in _tap_offsets(ndim):
        xs = _shift(x, *[v - 1 for v in t])
        sl = (slice(None),) + t + (slice(None),)
        dw[sl] = np.einsum(f"n{sp}o,n{sp}c->oc", gz, xs)
        dx += _shift(gz @ w[sl], *[1 - v for v in t])
    return dx, dw, gz.sum(axis=tuple(range(ndim + 1)))


def pool2(x):
    """Mean over each 2 x 2 (x 2) block of children."""
    ndim = x.ndim - 2
    shape = [x.shape[0]]
    for ax in range(ndim):
        shape += [x.shape[1 + ax] // 2, 2]
    return x.reshape(shape + [x.shape[-1]]).mean(axis=tuple(2 + 2 * a for a in range(ndim)))


def up2(x):
    """Nearest upsampling by 2 along every spatial axis."""
    for ax in range(x.ndim - 2):
        x = x.repeat(2, axis=1 + ax)
    return x


def up2_bwd(g):
    """Gradient of up2: the sum over each block of children."""
    return pool2(g) * 2 ** (g.ndim - 2)


def forward(params, inp, emulate=True):
    """inp [n, (D,) H, W, 2] features -> (out [n, (D,) H, W], cache)."""
    R = bf16 if emulate else (lambda a: np.asarray(a, np.float64))
    ndim = inp.ndim - 2
    layers, hw, hb = unpack(params, ndim)
    if emulate:
        layers = [(bf16(w), b) for w, b in layers]
    x0 = np.zeros(inp.shape[:-1] + (16,))
    x0[..., :2] = R(inp)
    acts = {"in0": x0}

    def conv(i, x):
        w, b = layers[i]
        return R(np.maximum(conv3x3(x, w, b), 0.0))

    cat = lambda a, b: np.concatenate([a, b], axis=-1)  # noqa: E731
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
    cat2 = cat(up2(t3b), s2)
    t2c = conv(8, cat2)
    t2d = conv(9, t2c)
    cat1 = cat(up2(t2d), s1)
    t1c = conv(10, cat1)
    t1d = conv(11, t1c)
    cat0 = cat(up2(t1d), s0)
    t0c = conv(12, cat0)
    t0d = conv(13, t0c)
    out = t0d @ hw + hb
    acts.update(t0a=t0a, s0=s0, p1=p1, t1a=t1a, s1=s1, p2=p2, t2a=t2a, s2=s2, p3=p3, t3a=t3a, t3b=t3b,
                cat2=cat2, t2c=t2c, t2d=t2d, cat1=cat1, t1c=t1c, t1d=t1d, cat0=cat0, t0c=t0c, t0d=t0d)
    return out, (layers, hw, hb, acts)


def _face_axes(ndim):
    """Face component a (x, y, z order) -> array axis of v [n, (D,) H, W]."""
    return [ndim - a for a in range(ndim)]


def _diff(v, ax):
    m = v.shape[ax]
    lo = [slice(None)] * v.ndim
    hi = [slice(None)] * v.ndim
    lo[ax], hi[ax] = slice(0, m - 1), slice(1, m)
    return tuple(lo), tuple(hi)


def energy(v, faces):
    """Per image sum over faces c_f (v_i - v_j)^2; faces [n, (D,) H, W, ndim]
    hold the coefficient toward x+1, y+1 (, z+1) (zero on walls)."""
    ndim = v.ndim - 1
    tot = np.zeros(v.shape[0])
    for a, ax in enumerate(_face_axes(ndim)):
        lo, hi = _diff(v, ax)
        d = v[lo] - v[hi]
        tot += np.sum(faces[..., a][lo] * d * d, axis=tuple(range(1, v.ndim)))
    return tot


def energy_grad(e, faces):
    """d/de of sum_f c_f (e_i - e_j)^2 per image (a weighted graph Laplacian)."""
    ndim = e.ndim - 1
    g = np.zeros_like(e)
    for a, ax in enumerate(_face_axes(ndim)):
        lo, hi = _diff(e, ax)
        d = faces[..., a][lo] * (e[lo] - e[hi])
        g[lo] += d
        g[hi] -= d
    return 2.0 * g


def _lap_diag(faces, ndim):
    """(A v as a function, D): A v = sum over a cell's faces c_f (v_p - v_nb)
    (half the energy gradient), D = the sum of a cell's faces."""
    ones_grad = lambda v: 0.5 * energy_grad(v, faces)  # noqa: E731
    dsum = np.zeros(faces.shape[:-1])
    for a, ax in enumerate(_face_axes(ndim)):
        lo, hi = _diff(dsum, ax)
        dsum[lo] += faces[..., a][lo]
        dsum[hi] += faces[..., a][lo]
    return ones_grad, dsum


def jacobi(v, faces, omega, transpose=False):
    """One damped-Jacobi sweep of the energy form's operator on an error
    (M v = v - omega D^-1 A v) or, transposed, on a gradient
    (M^T g = g - omega A (D^-1 g)); cells with D = 0 are left unchanged."""
    ndim = v.ndim - 1
    A, dsum = _lap_diag(faces, ndim)
    inv = np.where(dsum > 0, 1.0 / np.where(dsum > 0, dsum, 1.0), 0.0)
    if transpose:
        return v - omega * A(v * inv)
    return v - omega * A(v) * inv


def loss_and_grads(params, inp, target, emulate=True, faces=None, smooth=0, omega=0.8):
    """Loss over the batch and its gradient w.r.t. params: the mean squared
    error, or with faces the mean over images of E(M^smooth (out - target)) /
    E(target), M a damped-Jacobi sweep (smooth = 0: the plain energy norm)."""
    R = bf16 if emulate else (lambda a: np.asarray(a, np.float64))
    ndim = inp.ndim - 2
    out, (layers, hw, hb, a) = forward(params, inp, emulate)
    npix = out.size
    d = out - target
    n = out.shape[0]
    if faces is None:
        loss = float(np.sum(d * d) / npix)
        dout = 2.0 * d / npix
    else:
        et = energy(target, faces)
        e = d
        for _ in range(smooth):
            e = jacobi(e, faces, omega)
        loss = float(np.sum(energy(e, faces) / et) / n)
        dout = energy_grad(e, faces) / (et.reshape((n,) + (1,) * ndim) * n)
        for _ in range(smooth):
            dout = jacobi(dout, faces, omega, transpose=True)
    if emulate:
        dout = dout.astype(np.float32).astype(np.float64)
    offs, head, total = layout(ndim)
    T = taps(ndim)
    g = np.zeros(total)
    g[head:head + HEAD_C] = np.tensordot(dout, a["t0d"], axes=(list(range(ndim + 1)), list(range(ndim + 1))))
    g[head + HEAD_C] = dout.sum()
    gy = R(dout[..., None] * hw)

    def back(i, x, y, gy, want_dx=True):
        w, _ = layers[i]
        gz = gy * (y > 0)
        dx, dw, db = conv3x3_bwd(x, gz, w)
        cin, cout, _ = SPEC[i]
        o = offs[i]
        g[o:o + cout * T * cin] = dw.ravel()
        g[o + cout * T * cin:o + cout * T * cin + cout] = db
        return R(dx) if want_dx else None

    frac = 1.0 / 2 ** ndim  # pooling weight of each child
    g_t0c = back(13, a["t0c"], a["t0d"], gy)
    g_cat0 = back(12, a["cat0"], a["t0c"], g_t0c)
    g_t1d = R(up2_bwd(g_cat0[..., :32]))
    g_t1c = back(11, a["t1c"], a["t1d"], g_t1d)
    g_cat1 = back(10, a["cat1"], a["t1c"], g_t1c)
    g_t2d = R(up2_bwd(g_cat1[..., :64]))
    g_t2c = back(9, a["t2c"], a["t2d"], g_t2d)
    g_cat2 = back(8, a["cat2"], a["t2c"], g_t2c)
    g_t3b = R(up2_bwd(g_cat2[..., :128]))
    g_t3a = back(7, a["t3a"], a["t3b"], g_t3b)
    g_p3 = back(6, a["p3"], a["t3a"], g_t3a)
    g_s2 = R(g_cat2[..., 128:] + frac * up2(g_p3))
    g_t2a = back(5, a["t2a"], a["s2"], g_s2)
    g_p2 = back(4, a["p2"], a["t2a"], g_t2a)
    g_s1 = R(g_cat1[..., 64:] + frac * up2(g_p2))
    g_t1a = back(3, a["t1a"], a["s1"], g_s1)
    g_p1 = back(2, a["p1"], a["t1a"], g_t1a)
    g_s0 = R(g_cat0[..., 32:] + frac * up2(g_p1))
    g_t0a = back(1, a["t0a"], a["s0"], g_s0)
    back(0, a["in0"], a["t0a"], g_t0a, want_dx=False)
    return loss, g, out


def init_params(seed=0, ndim=2):
    """Random parameters of the same family the library draws (He-normal)."""
    rng = np.random.default_rng(seed)
    offs, head, total = layout(ndim)
    T = taps(ndim)
    p = np.zeros(total)
    for i, ((cin, cout, _), o) in enumerate(zip(SPEC, offs)):
        fan = T * (2 if i == 0 else cin)
        w = rng.standard_normal((cout, T, cin)) * np.sqrt(2.0 / fan)
        if i == 0:
            w[..., 2:] = 0.0
        p[o:o + cout * T * cin] = w.ravel()
        p[o + cout * T * cin:o + cout * T * cin + cout] = 0.05 * rng.standard_normal(cout)
    p[head:head + HEAD_C] = 0.25 * rng.standard_normal(HEAD_C)
    p[head + HEAD_C] = 0.1
    return p
