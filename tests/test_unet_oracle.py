"""Pins the U-Net oracle (oracle/unet_oracle.py) on the CPU.

The learned guess has no SciHPC counterpart, so its oracle is pinned by
mathematics: a brute-force convolution loop, torch.nn.functional.conv2d,
central finite differences of the loss for parameters of every layer, and
torch CPU autograd of the same network (all float64).
"""
import numpy as np
import pytest

from oracle import unet_oracle as U


def _rand(shape, seed):
    return np.random.default_rng(seed).standard_normal(shape)


def test_bf16_rounding():
    x = np.array([1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, -2.5, 1.0 + 2.0 ** -7 + 2.0 ** -9])
    got = U.bf16(x)
    # ties go to even; 3*2^-8 rounds up to 2^-6 mantissa step
    assert got.tolist() == [1.0, 1.0, 1.0 + 2.0 ** -6, -2.5, 1.0 + 2.0 ** -7]


def test_conv_bruteforce():
    x = _rand((2, 5, 6, 16), 1)
    w = _rand((16, 3, 3, 16), 2)
    b = _rand(16, 3)
    y = U.conv3x3(x, w, b)
    ref = np.zeros_like(y)
    for n in range(2):
        for i in range(5):
            for j in range(6):
                acc = b.copy()
                for ty in range(3):
                    for tx in range(3):
                        yy, xx = i + ty - 1, j + tx - 1
                        if 0 <= yy < 5 and 0 <= xx < 6:
                            acc += w[:, ty, tx, :] @ x[n, yy, xx, :]
                ref[n, i, j] = acc
    np.testing.assert_allclose(y, ref, rtol=1e-12, atol=1e-12)


def test_conv_matches_torch():
    torch = pytest.importorskip("torch")
    x = _rand((2, 8, 12, 32), 4)
    w = _rand((16, 3, 3, 32), 5)
    b = _rand(16, 6)
    y = U.conv3x3(x, w, b)
    t = torch.nn.functional.conv2d(torch.from_numpy(x).permute(0, 3, 1, 2),
                                   torch.from_numpy(w).permute(0, 3, 1, 2), torch.from_numpy(b), padding=1)
    np.testing.assert_allclose(y, t.permute(0, 2, 3, 1).numpy(), rtol=1e-11, atol=1e-11)


def test_conv_backward_matches_torch():
    torch = pytest.importorskip("torch")
    x = _rand((1, 6, 8, 16), 7)
    w = _rand((32, 3, 3, 16), 8)
    gz = _rand((1, 6, 8, 32), 9)
    dx, dw, db = U.conv3x3_bwd(x, gz, w)
    xt = torch.from_numpy(x).permute(0, 3, 1, 2).requires_grad_(True)
    wt = torch.from_numpy(w).permute(0, 3, 1, 2).contiguous().requires_grad_(True)
    bt = torch.zeros(32, dtype=torch.float64, requires_grad=True)
    y = torch.nn.functional.conv2d(xt, wt, bt, padding=1)
    y.backward(torch.from_numpy(gz).permute(0, 3, 1, 2))
    np.testing.assert_allclose(dx, xt.grad.permute(0, 2, 3, 1).numpy(), rtol=1e-11, atol=1e-11)
    np.testing.assert_allclose(dw, wt.grad.permute(0, 2, 3, 1).numpy(), rtol=1e-11, atol=1e-11)
    np.testing.assert_allclose(db, bt.grad.numpy(), rtol=1e-11, atol=1e-11)


def _problem(seed=0, n=1, H=8, W=8):
    p = U.init_params(seed)
    inp = _rand((n, H, W, 2), seed + 100)
    tgt = _rand((n, H, W), seed + 200)
    return p, inp, tgt


def test_gradients_finite_differences():
    p, inp, tgt = _problem(seed=11)
    loss, g, _ = U.loss_and_grads(p, inp, tgt, emulate=False)
    offs, head, total = U.layout()
    rng = np.random.default_rng(5)
    idx = [head + c for c in (0, 7, U.HEAD_C)]  # head weights and bias
    for (cin, cout, _), o in zip(U.SPEC, offs):
        nw = cout * 9 * cin
        ci = rng.integers(0, min(cin, 2) if o == 0 else cin)
        idx += [o + ((rng.integers(cout) * 9 + rng.integers(9)) * cin + ci), o + nw + rng.integers(cout)]
    eps = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += eps
        pm[i] -= eps
        lp = U.loss_and_grads(pp, inp, tgt, emulate=False)[0]
        lm = U.loss_and_grads(pm, inp, tgt, emulate=False)[0]
        fd = (lp - lm) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(g[i])) + 1e-9, (i, fd, g[i])


def _torch_unet(torch, params, inp):
    layers, hw, hb = U.unpack(params)
    W = [(torch.tensor(w).permute(0, 3, 1, 2).contiguous().requires_grad_(True),
          torch.tensor(b).requires_grad_(True)) for w, b in layers]
    hwt = torch.tensor(hw).requires_grad_(True)
    hbt = torch.tensor(float(hb), dtype=torch.float64).requires_grad_(True)
    F = torch.nn.functional
    x = torch.zeros(inp.shape[0], 16, inp.shape[1], inp.shape[2], dtype=torch.float64)
    x[:, :2] = torch.tensor(inp).permute(0, 3, 1, 2)

    def c(i, t):
        return F.relu(F.conv2d(t, W[i][0], W[i][1], padding=1))

    def up(t):
        return t.repeat_interleave(2, 2).repeat_interleave(2, 3)

    s0 = c(1, c(0, x))
    s1 = c(3, c(2, F.avg_pool2d(s0, 2)))
    s2 = c(5, c(4, F.avg_pool2d(s1, 2)))
    b3 = c(7, c(6, F.avg_pool2d(s2, 2)))
    d2 = c(9, c(8, torch.cat([up(b3), s2], 1)))
    d1 = c(11, c(10, torch.cat([up(d2), s1], 1)))
    d0 = c(13, c(12, torch.cat([up(d1), s0], 1)))
    out = torch.einsum("nchw,c->nhw", d0, hwt) + hbt
    return out, W, hwt, hbt


def test_network_matches_torch_autograd():
    torch = pytest.importorskip("torch")
    p, inp, tgt = _problem(seed=3, n=2, H=16, W=8)
    loss, g, out = U.loss_and_grads(p, inp, tgt, emulate=False)
    t_out, W, hwt, hbt = _torch_unet(torch, p, inp)
    t_loss = ((t_out - torch.tensor(tgt)) ** 2).mean()
    t_loss.backward()
    np.testing.assert_allclose(out, t_out.detach().numpy(), rtol=1e-11, atol=1e-12)
    assert abs(loss - t_loss.item()) <= 1e-12 * abs(loss)
    offs, head, total = U.layout()
    tg = np.zeros(total)
    for (cin, cout, _), o, (wt, bt) in zip(U.SPEC, offs, W):
        tg[o:o + cout * 9 * cin] = wt.grad.permute(0, 2, 3, 1).numpy().ravel()
        tg[o + cout * 9 * cin:o + cout * 9 * cin + cout] = bt.grad.numpy()
    tg[head:head + U.HEAD_C] = hwt.grad.numpy()
    tg[head + U.HEAD_C] = hbt.grad.item()
    np.testing.assert_allclose(g, tg, rtol=1e-9, atol=1e-12 * np.abs(tg).max())


def test_emulation_close_to_exact():
    """BF16 storage perturbs the output by about BF16 resolution, not more."""
    p, inp, tgt = _problem(seed=4, n=1, H=16, W=16)
    out_e, _ = U.forward(p, inp, emulate=True)
    out_x, _ = U.forward(p, inp, emulate=False)
    rel = np.linalg.norm(out_e - out_x) / np.linalg.norm(out_x)
    assert 1e-5 < rel < 5e-2


def _faces(n, H, W, seed):
    rng = np.random.default_rng(seed)
    f = rng.uniform(0.001, 1.0, (n, H, W, 2))
    f[:, :, -1, 0] = 0.0
    f[:, -1, :, 1] = 0.0
    return f


def test_energy_loss_gradient():
    """Energy-norm loss: closed form on a 2x1 grid, constants carry no energy,
    and the network gradient matches finite differences."""
    f = np.zeros((1, 1, 2, 2))
    f[0, 0, 0, 0] = 3.0
    v = np.array([[[1.0, -1.0]]])
    assert U.energy(v, f)[0] == 12.0
    np.testing.assert_allclose(U.energy_grad(v, f), [[[12.0, -12.0]]])
    assert U.energy(np.full((1, 4, 4), 2.5), _faces(1, 4, 4, 0))[0] == 0.0
    p, inp, tgt = _problem(seed=21, n=2, H=8, W=8)
    faces = _faces(2, 8, 8, 22)
    loss, g, _ = U.loss_and_grads(p, inp, tgt, emulate=False, faces=faces)
    offs, head, total = U.layout()
    eps = 1e-6
    for i in [head + 3, head + U.HEAD_C, offs[13] + 17, offs[7] + 1000, offs[0] + 1]:
        pp, pm = p.copy(), p.copy()
        pp[i] += eps
        pm[i] -= eps
        fd = (U.loss_and_grads(pp, inp, tgt, emulate=False, faces=faces)[0]
              - U.loss_and_grads(pm, inp, tgt, emulate=False, faces=faces)[0]) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(g[i])) + 1e-9, (i, fd, g[i])


# ------------------------------------------------------------------ 3-D
def test_conv3d_bruteforce():
    x = _rand((1, 3, 4, 5, 16), 31)
    w = _rand((16, 3, 3, 3, 16), 32)
    b = _rand(16, 33)
    y = U.conv3x3(x, w, b)
    ref = np.zeros_like(y)
    for k in range(3):
        for i in range(4):
            for j in range(5):
                acc = b.copy()
                for tz in range(3):
                    for ty in range(3):
                        for tx in range(3):
                            zz, yy, xx = k + tz - 1, i + ty - 1, j + tx - 1
                            if 0 <= zz < 3 and 0 <= yy < 4 and 0 <= xx < 5:
                                acc += w[:, tz, ty, tx, :] @ x[0, zz, yy, xx, :]
                ref[0, k, i, j] = acc
    np.testing.assert_allclose(y, ref, rtol=1e-12, atol=1e-12)


def test_conv3d_forward_backward_match_torch():
    torch = pytest.importorskip("torch")
    x = _rand((2, 4, 6, 5, 16), 34)
    w = _rand((32, 3, 3, 3, 16), 35)
    b = _rand(32, 36)
    gz = _rand((2, 4, 6, 5, 32), 37)
    y = U.conv3x3(x, w, b)
    dx, dw, db = U.conv3x3_bwd(x, gz, w)
    xt = torch.from_numpy(x).permute(0, 4, 1, 2, 3).requires_grad_(True)
    wt = torch.from_numpy(w).permute(0, 4, 1, 2, 3).contiguous().requires_grad_(True)
    bt = torch.from_numpy(b).requires_grad_(True)
    yt = torch.nn.functional.conv3d(xt, wt, bt, padding=1)
    np.testing.assert_allclose(y, yt.detach().permute(0, 2, 3, 4, 1).numpy(), rtol=1e-11, atol=1e-11)
    yt.backward(torch.from_numpy(gz).permute(0, 4, 1, 2, 3))
    np.testing.assert_allclose(dx, xt.grad.permute(0, 2, 3, 4, 1).numpy(), rtol=1e-11, atol=1e-11)
    np.testing.assert_allclose(dw, wt.grad.permute(0, 2, 3, 4, 1).numpy(), rtol=1e-11, atol=1e-11)
    np.testing.assert_allclose(db, bt.grad.numpy(), rtol=1e-11, atol=1e-11)


def _faces3(n, D, H, W, seed):
    f = np.random.default_rng(seed).uniform(0.001, 1.0, (n, D, H, W, 3))
    f[:, :, :, -1, 0] = 0.0
    f[:, :, -1, :, 1] = 0.0
    f[:, -1, :, :, 2] = 0.0
    return f


def test_network3d_matches_torch_autograd():
    """The 3-D network (27-tap convolutions, 2x2x2 pooling) with the energy
    loss against torch CPU autograd in float64."""
    torch = pytest.importorskip("torch")
    n, D, H, W = 2, 8, 8, 16
    p = U.init_params(41, ndim=3)
    inp = _rand((n, D, H, W, 2), 42)
    tgt = _rand((n, D, H, W), 43)
    faces = _faces3(n, D, H, W, 44)
    loss, g, out = U.loss_and_grads(p, inp, tgt, emulate=False, faces=faces)
    layers, hw, hb = U.unpack(p, 3)
    Wt = [(torch.tensor(w).permute(0, 4, 1, 2, 3).contiguous().requires_grad_(True),
           torch.tensor(b).requires_grad_(True)) for w, b in layers]
    hwt = torch.tensor(hw).requires_grad_(True)
    hbt = torch.tensor(float(hb), dtype=torch.float64).requires_grad_(True)
    F = torch.nn.functional
    x = torch.zeros(n, 16, D, H, W, dtype=torch.float64)
    x[:, :2] = torch.tensor(inp).permute(0, 4, 1, 2, 3)

    def c(i, t):
        return F.relu(F.conv3d(t, Wt[i][0], Wt[i][1], padding=1))

    def up(t):
        return t.repeat_interleave(2, 2).repeat_interleave(2, 3).repeat_interleave(2, 4)

    s0 = c(1, c(0, x))
    s1 = c(3, c(2, F.avg_pool3d(s0, 2)))
    s2 = c(5, c(4, F.avg_pool3d(s1, 2)))
    b3 = c(7, c(6, F.avg_pool3d(s2, 2)))
    d2 = c(9, c(8, torch.cat([up(b3), s2], 1)))
    d1 = c(11, c(10, torch.cat([up(d2), s1], 1)))
    d0 = c(13, c(12, torch.cat([up(d1), s0], 1)))
    t_out = torch.einsum("ncdhw,c->ndhw", d0, hwt) + hbt
    ft = torch.tensor(faces)
    tt = torch.tensor(tgt)

    def en(v):
        ex = (v[:, :, :, :-1] - v[:, :, :, 1:]) ** 2 * ft[:, :, :, :-1, 0]
        ey = (v[:, :, :-1, :] - v[:, :, 1:, :]) ** 2 * ft[:, :, :-1, :, 1]
        ez = (v[:, :-1] - v[:, 1:]) ** 2 * ft[:, :-1, :, :, 2]
        return ex.sum((1, 2, 3)) + ey.sum((1, 2, 3)) + ez.sum((1, 2, 3))

    t_loss = (en(t_out - tt) / en(tt)).mean()
    t_loss.backward()
    np.testing.assert_allclose(out, t_out.detach().numpy(), rtol=1e-11, atol=1e-12)
    assert abs(loss - t_loss.item()) <= 1e-11 * abs(loss)
    offs, head, total = U.layout(3)
    tg = np.zeros(total)
    for (cin, cout, _), o, (wt, bt) in zip(U.SPEC, offs, Wt):
        tg[o:o + cout * 27 * cin] = wt.grad.permute(0, 2, 3, 4, 1).numpy().ravel()
        tg[o + cout * 27 * cin:o + cout * 27 * cin + cout] = bt.grad.numpy()
    tg[head:head + U.HEAD_C] = hwt.grad.numpy()
    tg[head + U.HEAD_C] = hbt.grad.item()
    np.testing.assert_allclose(g, tg, rtol=1e-9, atol=1e-12 * np.abs(tg).max())


def test_energy3d_closed_form():
    """3-D energy on a 2x1x1 column: one z face; constants carry no energy."""
    f = np.zeros((1, 2, 1, 1, 3))
    f[0, 0, 0, 0, 2] = 2.0
    v = np.array([[[[3.0]], [[1.0]]]])
    assert U.energy(v, f)[0] == 8.0
    np.testing.assert_allclose(U.energy_grad(v, f).ravel(), [8.0, -8.0])
    assert U.energy(np.full((1, 4, 4, 4), -1.5), _faces3(1, 4, 4, 4, 1))[0] == 0.0


def test_smoothed_energy_loss():
    """The smoothed energy loss E(M^S e) / E(t): the Jacobi sweep M is the
    textbook damped Jacobi of the face Laplacian (checked against a dense
    matrix), M^T is its transpose, it removes the highest-frequency mode
    of a uniform grid faster than a smooth one, and the network gradient
    matches finite differences (2-D and 3-D)."""
    rng = np.random.default_rng(3)
    f = _faces(1, 4, 5, 8)
    n = 20
    A = np.zeros((n, n))
    for c in range(n):
        e = np.zeros(n)
        e[c] = 1.0
        A[:, c] = (0.5 * U.energy_grad(e.reshape(1, 4, 5), f)).ravel()
    Dm = np.diag(np.diag(A))
    M = np.eye(n) - 0.7 * np.linalg.solve(Dm, A)
    v = rng.standard_normal((1, 4, 5))
    np.testing.assert_allclose(U.jacobi(v, f, 0.7).ravel(), M @ v.ravel(), rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(U.jacobi(v, f, 0.7, transpose=True).ravel(), M.T @ v.ravel(), rtol=1e-12, atol=1e-12)
    ones = np.ones((1, 8, 8, 2))
    ones[:, :, -1, 0] = 0.0
    ones[:, -1, :, 1] = 0.0
    y, x = np.meshgrid(np.arange(8), np.arange(8), indexing="ij")
    rough = ((-1.0) ** (x + y))[None]
    smooth = np.cos(np.pi * (x + 0.5) / 8)[None]
    r_rough = np.linalg.norm(U.jacobi(rough, ones, 0.8)) / np.linalg.norm(rough)
    r_smooth = np.linalg.norm(U.jacobi(smooth, ones, 0.8)) / np.linalg.norm(smooth)
    assert r_rough < 0.7 and r_smooth > 0.9, (r_rough, r_smooth)
    for ndim, (n_, shape) in ((2, (2, (8, 8))), (3, (1, (8, 8, 8)))):
        p = U.init_params(31, ndim=ndim)
        inp = _rand((n_,) + shape + (2,), 32)
        tgt = _rand((n_,) + shape, 33)
        fc = _faces(n_, *shape, 34) if ndim == 2 else _faces3(n_, *shape, 34)
        loss, g, _ = U.loss_and_grads(p, inp, tgt, emulate=False, faces=fc, smooth=2, omega=0.8)
        offs, head, _ = U.layout(ndim)
        eps = 1e-6
        for i in [head + 2, head + U.HEAD_C, offs[13] + 11, offs[6] + 500, offs[0] + 1]:
            pp, pm = p.copy(), p.copy()
            pp[i] += eps
            pm[i] -= eps
            fd = (U.loss_and_grads(pp, inp, tgt, emulate=False, faces=fc, smooth=2, omega=0.8)[0]
                  - U.loss_and_grads(pm, inp, tgt, emulate=False, faces=fc, smooth=2, omega=0.8)[0]) / (2 * eps)
            assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(g[i])) + 1e-9, (ndim, i, fd, g[i])
