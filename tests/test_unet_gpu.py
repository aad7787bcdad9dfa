"""GPU parity of the learned pressure guess (U-Net kernels through the C ABI)
against the CPU oracle (oracle/unet_oracle.py), and its use in the solve.

Bars: the oracle rounds every BF16-stored tensor where the GPU does, so the
two differ only by FP32 accumulation order, which flips an occasional BF16
rounding: single convolution within 1 BF16 ulp (2^-7 relative) of the
rounded oracle value; whole-network output within 1e-2 and every layer's
gradient within 3e-2 relative L2.  The solve with a learned guess must
reach the same FP64 relative residual as without it, and a guess that does
not lower the FP64 residual must be rejected.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from oracle import unet_oracle as U
from scihpc_amd import PREC_FP64, PREC_MIXED, Multigrid, ScihpcError, UNet
from scihpc_amd import _capi
from scihpc_amd import learned as LRN
from scihpc_amd import workloads as W


def rel(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / max(np.linalg.norm(np.ravel(b)), 1e-300))


def _bf16_bits(a):
    f = np.ascontiguousarray(a, np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)
    return u


def _bits_to_f(u):
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


CONV_CASES = [(1, 8, 64, 16, 16), (2, 5, 40, 48, 16), (1, 7, 100, 32, 64), (1, 4, 16, 192, 64),
              (1, 3, 130, 128, 128), (2, 6, 24, 16, 48), (1, 1, 9, 64, 96), (1, 4, 12, 16, 192)]
# 64 pixel tiles or more (no output-channel split) with all output blocks in
# one workgroup over an LDS input tile, the wavefronts splitting them:
# 2, 4, 6, 8 and 12 blocks, ragged and full 64-pixel segments, and a tile
# past the default 64 KiB of dynamic LDS (192 input channels)
CONV_CASES += [(1, 16, 200, 64, 32), (1, 32, 130, 32, 64), (2, 16, 100, 64, 96), (1, 64, 64, 64, 128),
               (1, 64, 70, 48, 192), (1, 32, 66, 192, 192)]


@pytest.mark.parametrize("case", CONV_CASES, ids=lambda c: "x".join(map(str, c)))
def test_conv3x3_forward(case):
    n, h, w, cin, cout = case
    rng = np.random.default_rng(sum(case))
    xb = _bf16_bits(rng.standard_normal((n, h, w, cin)))
    wb = _bf16_bits(rng.standard_normal((cout, 3, 3, cin)) / np.sqrt(9 * cin))
    b = rng.standard_normal(cout).astype(np.float32)
    y = np.zeros((n, h, w, cout), np.uint16)
    L = _capi.lib()
    for relu in (0, 1):
        print(f"conv3x3 {case} relu={relu}", flush=True)
        _capi.check(L.scihpc_conv3x3_forward(xb.ravel(), n, h, w, cin, wb.ravel(), b, cout, relu, y.reshape(-1)))
        got = _bits_to_f(y)
        ref = U.conv3x3(_bits_to_f(xb), _bits_to_f(wb), b.astype(np.float64))
        if relu:
            ref = np.maximum(ref, 0.0)
        ref_b = U.bf16(ref)
        tol = 2.0 ** -7 * np.abs(ref) + 1e-6 * np.abs(ref).max()
        assert np.all(np.abs(got - ref_b) <= tol), np.abs(got - ref_b).max()


def _net_problem(n=2, H=16, W=32, seed=0):
    rng = np.random.default_rng(seed)
    p = U.init_params(seed).astype(np.float32)
    feat = rng.standard_normal((n, H * W, 2)).astype(np.float32)
    tgt = rng.standard_normal((n, H * W)).astype(np.float32)
    return p, feat, tgt


@pytest.mark.parametrize("shape", [(2, 16, 32), (1, 24, 16), (3, 8, 72)], ids=lambda s: "x".join(map(str, s)))
def test_unet_forward_backward(shape):
    n, H, Wd = shape
    p, feat, tgt = _net_problem(n, H, Wd, seed=H + Wd)
    print(f"unet fwd/bwd {shape}", flush=True)
    net = UNet(H, Wd, n)
    net.set_params(p)
    net.out_scale = 1.0
    net.set_samples(feat, tgt)
    out = net.forward(np.arange(n))
    loss = net.backward()
    g = net.grads()
    inp = feat.reshape(n, H, Wd, 2).astype(np.float64)
    l_ref, g_ref, out_ref = U.loss_and_grads(p.astype(np.float64), inp, tgt.reshape(n, H, Wd).astype(np.float64))
    assert rel(out, out_ref) < 1e-2
    assert abs(loss - l_ref) < 1e-2 * l_ref
    offs, head, total = U.layout()
    assert g.size == total
    for i, ((cin, cout, _), o) in enumerate(zip(U.SPEC, offs)):
        nw = cout * 9 * cin
        sl = slice(o, o + nw + cout)
        if i == 0:  # only the two real input channels carry gradient
            wg = g[o:o + nw].reshape(cout, 9, cin)
            assert np.all(wg[:, :, 2:] == 0.0)
        assert rel(g[sl], g_ref[sl]) < 3e-2, (i, rel(g[sl], g_ref[sl]))
    assert rel(g[head:], g_ref[head:]) < 3e-2
    # BF16 storage changes the exact-arithmetic gradient only a little
    _, g_exact, _ = U.loss_and_grads(p.astype(np.float64), inp, tgt.reshape(n, H, Wd).astype(np.float64),
                                     emulate=False)
    assert rel(g, g_exact) < 1e-1


def test_unet_energy_loss_gradients():
    n, H, Wd = 2, 16, 24
    p, feat, tgt = _net_problem(n, H, Wd, seed=9)
    rng = np.random.default_rng(10)
    faces = rng.uniform(0.001, 1.0, (n, H, Wd, 2)).astype(np.float32)
    faces[:, :, -1, 0] = 0.0
    faces[:, -1, :, 1] = 0.0
    net = UNet(H, Wd, n)
    net.set_params(p)
    net.out_scale = 1.0
    net.set_samples(feat, tgt, faces.reshape(n, H * Wd, 2))
    net.forward(np.arange(n))
    loss = net.backward()
    g = net.grads()
    inp = feat.reshape(n, H, Wd, 2).astype(np.float64)
    l_ref, g_ref, _ = U.loss_and_grads(p.astype(np.float64), inp, tgt.reshape(n, H, Wd).astype(np.float64),
                                       faces=faces.astype(np.float64))
    assert abs(loss - l_ref) < 1e-2 * l_ref
    offs, head, _ = U.layout()
    for i, ((cin, cout, _), o) in enumerate(zip(U.SPEC, offs)):
        sl = slice(o, o + cout * 9 * cin + cout)
        assert rel(g[sl], g_ref[sl]) < 3e-2, (i, rel(g[sl], g_ref[sl]))
    assert rel(g[head:], g_ref[head:]) < 3e-2
    # mirrored faces: an x (y) mirror maps the face between x and x+1 to W-2-x
    # (H-2-y); the loss of a mirrored batch equals the oracle's on mirrored data
    for fl, mirror, fmir in [
        (1, lambda a: a[:, :, ::-1], lambda f: np.stack([np.concatenate([f[:, :, -2::-1, 0], f[:, :, -1:, 0]], 2),
                                                         f[:, :, ::-1, 1]], -1)),
        (2, lambda a: a[:, ::-1, :], lambda f: np.stack([f[:, ::-1, :, 0],
                                                         np.concatenate([f[:, -2::-1, :, 1], f[:, -1:, :, 1]], 1)],
                                                        -1))]:
        net.forward([0], [fl])
        lf = net.backward()
        mi = np.ascontiguousarray(mirror(inp[:1]))
        mt = np.ascontiguousarray(mirror(tgt.reshape(n, H, Wd)[:1].astype(np.float64)))
        mf = np.ascontiguousarray(fmir(faces[:1].astype(np.float64)))
        l_m, _, _ = U.loss_and_grads(p.astype(np.float64), mi, mt, faces=mf)
        assert abs(lf - l_m) < 1e-2 * l_m, (fl, lf, l_m)


def test_unet_input_flips():
    n, H, Wd = 1, 16, 24
    p, feat, tgt = _net_problem(n, H, Wd, seed=5)
    net = UNet(H, Wd, 1)
    net.set_params(p)
    net.set_samples(feat, tgt)
    inp = feat.reshape(1, H, Wd, 2).astype(np.float64)
    for fl, mirror in [(1, lambda a: a[:, :, ::-1]), (2, lambda a: a[:, ::-1, :]), (3, lambda a: a[:, ::-1, ::-1])]:
        out = net.forward([0], [fl]).reshape(1, H, Wd)
        ref, _ = U.forward(p.astype(np.float64), np.ascontiguousarray(mirror(inp)))
        assert rel(out, ref) < 1e-2, fl


def test_unet_training_reduces_loss():
    """Adam on the GPU fits a smooth target field: loss falls by >= 10x."""
    H, Wd, n = 32, 32, 4
    rng = np.random.default_rng(3)
    y, x = np.meshgrid(np.linspace(0, 1, H), np.linspace(0, 1, Wd), indexing="ij")
    feats, tgts = [], []
    for k in range(n):
        a, b = rng.uniform(0.5, 2.0, 2)
        f0 = np.sin(np.pi * a * x) * np.cos(np.pi * b * y)
        f1 = np.cos(np.pi * b * x)
        feats.append(np.stack([f0, f1], -1).reshape(H * Wd, 2))
        tgts.append((0.5 * f0 + 0.3 * f1 * f0).reshape(H * Wd))
    feat = np.array(feats, np.float32)
    tgt = np.array(tgts, np.float32)
    net = UNet(H, Wd, 4, seed=2)
    hist = LRN.train(net, feat, tgt, steps=300, batch=4, lr=2e-3, log_every=50)
    assert hist[-1][1] < 0.1 * hist[0][1], hist


def test_learned_guess_fallback_on_bad_network():
    """A network whose guess does not lower the FP64 residual (NaN weights)
    is rejected and the solve equals the plain solve."""
    w = W.bubble_2d(64, 128, seed=11)
    mg = Multigrid(2, 64, 128, 1, 1 / 64, 2 / 128, 1.0, "neumann")
    mg.set_density(w["rho"])
    mg.set_rhs(w["rhs"])
    plain = mg.solve(rel_tol=1e-8, precision=PREC_MIXED, accel=1)
    x_plain = mg.sol(0)
    H, Wd = mg.shapes[1][1], mg.shapes[1][0]
    net = UNet(H, Wd, 1)
    p = net.params()
    p[:] = np.nan
    net.set_params(p)
    mg.attach_learned(net, 1)
    st = mg.solve(rel_tol=1e-8, precision=PREC_MIXED, accel=1, use_learned=3)
    assert st["converged"] and st["learned_rejected"] == 1 and st["learned_accepted"] == 0
    assert st["cycles"] == plain["cycles"]
    assert rel(mg.sol(0), x_plain) < 1e-12
    # a zero network (guess 0) is rejected as well: it cannot lower ||b||
    net.set_params(np.zeros_like(p))
    st = mg.solve(rel_tol=1e-8, precision=PREC_FP64, accel=0, use_learned=3)
    assert st["converged"] and st["learned_rejected"] == 1
    mg.attach_learned(None)
    st = mg.solve(rel_tol=1e-8, precision=PREC_FP64, accel=0, use_learned=True)
    assert st["learned_rejected"] == 0 and st["learned_accepted"] == 0


def test_learned_correction_judged_and_fallback():
    """learned_correction: the network's coarse-level correction of the FP64
    defect after the first flexible-CG iteration.  A correction the FP64
    residual rejects (NaN weights) leaves the iteration untouched, so the
    result is the plain solve's; a forced one (2) moves the iterate and
    restarts the recurrence, and the solve still reaches the tolerance; a
    judged one from random weights is counted either way.  Other solve
    modes refuse the option."""
    w = W.bubble_2d(64, 128, seed=17)
    mg = Multigrid(2, 64, 128, 1, 1 / 64, 2 / 128, 1.0, "neumann")
    mg.set_density(w["rho"])
    mg.set_rhs(w["rhs"])
    plain = mg.solve(rel_tol=1e-8, precision=PREC_MIXED, accel=1)
    x_plain = mg.sol(0)
    H, Wd = mg.shapes[1][1], mg.shapes[1][0]
    net = UNet(H, Wd, 1)
    n_par = net.num_params
    net.set_params(np.full(n_par, np.nan, np.float32))
    mg.attach_learned(net, 1)
    st = mg.solve(rel_tol=1e-8, precision=PREC_MIXED, accel=1, learned_correction=1)
    assert st["converged"] and st["correction_rejected"] == 1 and st["correction_accepted"] == 0
    assert st["cycles"] == plain["cycles"]
    np.testing.assert_array_equal(mg.sol(0), x_plain)
    rng = np.random.default_rng(5)
    net.set_params((0.05 * rng.standard_normal(n_par)).astype(np.float32))
    st = mg.solve(rel_tol=1e-8, precision=PREC_MIXED, accel=1, learned_correction=2)
    assert st["converged"] and st["rel_residual"] <= 1e-8
    assert st["correction_accepted"] == 0 and st["correction_rejected"] == 0
    st = mg.solve(rel_tol=1e-8, precision=PREC_MIXED, accel=1, learned_correction=1)
    assert st["converged"] and st["rel_residual"] <= 1e-8
    assert st["correction_accepted"] + st["correction_rejected"] == 1
    if st["correction_rejected"]:
        assert st["correction_residual"] >= 1.0 or st["correction_residual"] < 0
        assert st["cycles"] == plain["cycles"]
    else:
        assert 0.0 <= st["correction_residual"] < 1.0
    # with the guess overlapping the first iterations as well: the correction
    # waits until the guess is judged, and both verdicts are counted
    st = mg.solve(rel_tol=1e-8, precision=PREC_MIXED, accel=1, use_learned=3, learned_correction=1)
    assert st["converged"] and st["rel_residual"] <= 1e-8
    assert st["learned_accepted"] + st["learned_rejected"] == 1
    assert st["correction_accepted"] + st["correction_rejected"] <= 1
    with pytest.raises(ScihpcError):
        mg.solve(rel_tol=1e-8, precision=PREC_FP64, accel=1, learned_correction=1)
    with pytest.raises(ScihpcError):
        mg.solve(rel_tol=1e-8, precision=PREC_MIXED, accel=0, learned_correction=1)
    mg.attach_learned(None)
    st = mg.solve(rel_tol=1e-8, precision=PREC_MIXED, accel=1, learned_correction=1)
    assert st["correction_accepted"] == 0 and st["correction_rejected"] == 0


def test_learned_guess_backoff_after_rejections():
    """use_learned=1: after the k-th rejection in a row the next 2^k - 1
    solves skip the network (the result is the plain solve's); attaching a
    network again resets the back-off."""
    w = W.bubble_2d(64, 128, seed=13)
    mg = Multigrid(2, 64, 128, 1, 1 / 64, 2 / 128, 1.0, "neumann")
    mg.set_density(w["rho"])
    mg.set_rhs(w["rhs"])
    plain = mg.solve(rel_tol=1e-8, precision=PREC_MIXED, accel=1)
    x_plain = mg.sol(0)
    H, Wd = mg.shapes[1][1], mg.shapes[1][0]
    net = UNet(H, Wd, 1)
    net.set_params(np.zeros(net.num_params, np.float32))  # guess 0: always rejected
    mg.attach_learned(net, 1)
    tried = []
    for s in range(12):
        st = mg.solve(rel_tol=1e-8, precision=PREC_MIXED, accel=1, use_learned=1)
        assert st["converged"] and st["cycles"] == plain["cycles"]
        assert rel(mg.sol(0), x_plain) < 1e-12
        assert st["learned_rejected"] + st["learned_skipped"] == 1
        if st["learned_rejected"]:
            tried.append(s)
    assert tried == [0, 2, 6], tried
    mg.attach_learned(net, 1)
    st = mg.solve(rel_tol=1e-8, precision=PREC_MIXED, accel=1, use_learned=1)
    assert st["learned_rejected"] == 1 and st["learned_skipped"] == 0
    st = mg.solve(rel_tol=1e-8, precision=PREC_MIXED, accel=1, use_learned=3)
    assert st["learned_rejected"] == 1 and st["learned_skipped"] == 0


def test_learned_features_match_block_means():
    w = W.bubble_2d(64, 128, seed=12)
    mg = Multigrid(2, 64, 128, 1, 1 / 64, 2 / 128, 1.0, "neumann")
    mg.set_density(w["rho"])
    mg.set_rhs(w["rhs"])
    mg.solve(rel_tol=1e-10, precision=PREC_FP64, accel=1)
    p = mg.sol(0)
    for level in (1, 2):
        f, t, c = mg.learned_features(level, with_target=True, with_faces=True)
        bs = 2 ** level
        b = w["rhs"].reshape(128, 64)
        sigma = np.sqrt(np.mean(b ** 2))

        def bm(a):
            return a.reshape(128 // bs, bs, 64 // bs, bs).mean(axis=(1, 3)).ravel()

        np.testing.assert_allclose(f[:, 0], bm(b) / sigma, rtol=1e-5, atol=1e-6)
        np.testing.assert_allclose(f[:, 1], bm(np.log(w["rho"].reshape(128, 64))) / np.log(1000.0), rtol=1e-5,
                                   atol=1e-6)
        np.testing.assert_allclose(t, bm(p.reshape(128, 64)) / sigma, rtol=1e-5,
                                   atol=1e-6 * np.abs(t).max())
        # faces x h^2: level 0 is 1 / (arithmetic mean rho) (init_ppe_proj_coefficients,
        # projection_method.cpp:404-465); coarse faces average the fine faces they cover
        rho = w["rho"].reshape(128, 64)
        fx0 = np.zeros((128, 64))
        fy0 = np.zeros((128, 64))
        fx0[:, :-1] = 1.0 / (0.5 * (rho[:, :-1] + rho[:, 1:]))
        fy0[:-1, :] = 1.0 / (0.5 * (rho[:-1, :] + rho[1:, :]))
        ex = fx0[:, bs - 1::bs].reshape(128 // bs, bs, 64 // bs).mean(axis=1)
        ey = fy0[bs - 1::bs, :].reshape(128 // bs, 64 // bs, bs).mean(axis=2)
        ex[:, -1] = 0.0
        ey[-1, :] = 0.0
        np.testing.assert_allclose(c[:, 0], ex.ravel(), rtol=1e-6, atol=1e-9)
        np.testing.assert_allclose(c[:, 1], ey.ravel(), rtol=1e-6, atol=1e-9)


def test_learned_guess_trained_gains():
    """Ablation on held-out bubbles with the shipped weights (trained on the
    GPU by python -m scihpc_amd.learned on solver-written samples): the
    network's guess is far closer to the solution than the zero guess in the
    A(rho) energy norm, and whether the FP64 judgement keeps it or not the
    learned solve reaches 1e-8 with at most one more FCG iteration than the
    plain one (a kept guess replaces the first iteration from p = 0)."""
    net, meta = LRN.load_net()
    mg = Multigrid(2, 512, 1024, 1, 1 / 512, 2 / 1024, 1.0, "neumann")
    mg.attach_learned(net, meta["level"])
    errs = []
    for cfg in LRN.bubble_configs(3, LRN.EVAL_SEED):
        w = LRN.make_problem(cfg)
        mg.set_density(w["rho"])
        mg.set_rhs(w["rhs"])
        plain = mg.solve(rel_tol=1e-8, precision=PREC_MIXED, accel=1)
        mg.solve(rel_tol=1e-10, precision=PREC_FP64, accel=1)
        xs = mg.sol(0)
        mg.solve(rel_tol=1e-8, max_cycles=0, precision=PREC_MIXED, accel=1, use_learned=2)
        e = mg.sol(0) - xs
        e -= e.mean()
        err = np.sqrt((e @ mg.apply(e)) / (xs @ mg.apply(xs)))  # both energies are negative
        errs.append(err)
        st = mg.solve(rel_tol=1e-8, precision=PREC_MIXED, accel=1, use_learned=3)
        print(f"r={cfg['radius']:.3f} energy err {err:.3f} guess residual {st['learned_residual']:.3e} "
              f"accepted {st['learned_accepted']} cycles plain {plain['cycles']} learned {st['cycles']}", flush=True)
        assert st["converged"] and st["rel_residual"] <= 1e-8
        assert st["learned_accepted"] + st["learned_rejected"] == 1
        assert st["cycles"] <= plain["cycles"] + 1
    assert max(errs) < 0.6, errs


# ------------------------------------------------------------------ 3-D
def _grads_close(g, g_ref, g_exact, ndim):
    """Per-layer gradient bar of the 3-D and smoothed-loss tests.  The GPU
    differs from the BF16-emulating oracle only by FP32 accumulation order,
    which flips occasional BF16 roundings of stored activations and
    gradients; each flip perturbs the backward chain as much as BF16
    storage itself does.  So a layer's bar is twice what BF16 storage moves
    that layer's gradient from exact arithmetic (the oracle with
    emulate=False: 2-25 % on these random problems, most where the
    smoothing's low-pass gradient meets high-frequency activations), and
    never below 3e-2 (the same for the head)."""
    offs, head, _ = U.layout(ndim)
    T = 3 ** ndim
    for i, ((cin, cout, _), o) in enumerate(zip(U.SPEC, offs)):
        sl = slice(o, o + cout * T * cin + cout)
        noise = rel(g_ref[sl], g_exact[sl])
        assert rel(g[sl], g_ref[sl]) < max(3e-2, 2.0 * noise), (i, rel(g[sl], g_ref[sl]), noise)
    noise = rel(g_ref[head:], g_exact[head:])
    assert rel(g[head:], g_ref[head:]) < max(3e-2, 2.0 * noise), (rel(g[head:], g_ref[head:]), noise)

CONV3_CASES = [(1, 4, 5, 64, 16, 16), (2, 3, 4, 40, 32, 48), (1, 2, 3, 70, 192, 64), (1, 4, 4, 8, 128, 128),
               (1, 3, 2, 16, 48, 192)]
# the 3-D LDS-tile path with the wavefronts splitting 2, 4, 6, 8 and 12
# output blocks (48 input channels: a tile past 64 KiB of dynamic LDS)
CONV3_CASES += [(1, 4, 16, 70, 32, 32), (1, 4, 8, 100, 48, 64), (2, 4, 8, 40, 32, 96), (1, 8, 8, 64, 64, 128),
                (1, 4, 16, 20, 16, 192)]


@pytest.mark.parametrize("case", CONV3_CASES, ids=lambda c: "x".join(map(str, c)))
def test_conv3x3x3_forward(case):
    n, d, h, w, cin, cout = case
    rng = np.random.default_rng(sum(case))
    xb = _bf16_bits(rng.standard_normal((n, d, h, w, cin)))
    wb = _bf16_bits(rng.standard_normal((cout, 3, 3, 3, cin)) / np.sqrt(27 * cin))
    b = rng.standard_normal(cout).astype(np.float32)
    y = np.zeros((n, d, h, w, cout), np.uint16)
    L = _capi.lib()
    for relu in (0, 1):
        print(f"conv3x3x3 {case} relu={relu}", flush=True)
        _capi.check(L.scihpc_conv3x3x3_forward(xb.ravel(), n, d, h, w, cin, wb.ravel(), b, cout, relu,
                                               y.reshape(-1)))
        got = _bits_to_f(y)
        ref = U.conv3x3(_bits_to_f(xb), _bits_to_f(wb), b.astype(np.float64))
        if relu:
            ref = np.maximum(ref, 0.0)
        ref_b = U.bf16(ref)
        tol = 2.0 ** -7 * np.abs(ref) + 1e-6 * np.abs(ref).max()
        assert np.all(np.abs(got - ref_b) <= tol), np.abs(got - ref_b).max()


def _faces3(n, D, H, Wd, seed):
    f = np.random.default_rng(seed).uniform(0.001, 1.0, (n, D, H, Wd, 3)).astype(np.float32)
    f[:, :, :, -1, 0] = 0.0
    f[:, :, -1, :, 1] = 0.0
    f[:, -1, :, :, 2] = 0.0
    return f


@pytest.mark.parametrize("shape", [(2, 8, 8, 16), (1, 16, 8, 8)], ids=lambda s: "x".join(map(str, s)))
def test_unet3d_forward_backward(shape):
    """The 3-D network (27-tap convolutions, 2x2x2 pooling): output, loss and
    every layer's gradient against the oracle, for the mean squared error
    and the energy-norm loss."""
    n, D, H, Wd = shape
    rng = np.random.default_rng(D + H + Wd)
    p = U.init_params(7, ndim=3).astype(np.float32)
    pix = D * H * Wd
    feat = rng.standard_normal((n, pix, 2)).astype(np.float32)
    tgt = rng.standard_normal((n, pix)).astype(np.float32)
    faces = _faces3(n, D, H, Wd, 3)
    print(f"unet3d fwd/bwd {shape}", flush=True)
    net = UNet(H, Wd, n, depth=D)
    assert net.num_params == U.layout(3)[2]
    net.set_params(p)
    net.out_scale = 1.0
    inp = feat.reshape(n, D, H, Wd, 2).astype(np.float64)
    t64 = tgt.reshape(n, D, H, Wd).astype(np.float64)
    for fc in (None, faces):
        net.set_samples(feat, tgt, None if fc is None else fc.reshape(n, pix, 3))
        out = net.forward(np.arange(n))
        loss = net.backward()
        g = net.grads()
        f64 = None if fc is None else fc.astype(np.float64)
        l_ref, g_ref, out_ref = U.loss_and_grads(p.astype(np.float64), inp, t64, faces=f64)
        _, g_exact, _ = U.loss_and_grads(p.astype(np.float64), inp, t64, faces=f64, emulate=False)
        assert rel(out, out_ref) < 1e-2
        assert abs(loss - l_ref) < 1e-2 * l_ref
        _grads_close(g, g_ref, g_exact, 3)


def test_unet3d_input_flips():
    D, H, Wd = 8, 8, 16
    rng = np.random.default_rng(8)
    p = U.init_params(9, ndim=3).astype(np.float32)
    feat = rng.standard_normal((1, D * H * Wd, 2)).astype(np.float32)
    net = UNet(H, Wd, 1, depth=D)
    net.set_params(p)
    net.set_samples(feat, np.zeros((1, D * H * Wd), np.float32))
    inp = feat.reshape(1, D, H, Wd, 2).astype(np.float64)
    for fl in range(1, 8):
        axes = [ax for bit, ax in ((1, 3), (2, 2), (4, 1)) if fl & bit]
        out = net.forward([0], [fl]).reshape(1, D, H, Wd)
        ref, _ = U.forward(p.astype(np.float64), np.ascontiguousarray(np.flip(inp, axes)))
        assert rel(out, ref) < 1e-2, fl


def test_unet3d_training_reduces_loss():
    """Adam on the GPU fits a smooth 3-D target: loss falls by >= 10x."""
    D = H = Wd = 8
    n = 2
    z, y, x = np.meshgrid(*(np.linspace(0, 1, 8),) * 3, indexing="ij")
    tgt = np.stack([np.sin(np.pi * x) * np.cos(np.pi * y) * z, np.cos(np.pi * z) * y]).reshape(n, -1)
    feat = np.stack([np.stack([np.sin(np.pi * x) * z, y], -1), np.stack([np.cos(np.pi * z), y * x], -1)])
    net = UNet(H, Wd, n, seed=4, depth=D)
    net.out_scale = float(np.sqrt(np.mean(tgt ** 2)))
    net.set_samples(feat.reshape(n, -1, 2).astype(np.float32), tgt.astype(np.float32))
    first = net.train_step([0, 1], 1e-3)
    for _ in range(300):
        last = net.train_step([0, 1], 1e-3, want_loss=False)
    last = net.train_step([0, 1], 1e-4)
    print(f"3-D training loss {first:.3e} -> {last:.3e}", flush=True)
    assert last < 0.1 * first


def test_learned3d_features_and_fallback():
    """3-D multigrid hook: features, targets and faces on a coarse level are
    block means of the level-0 fields (faces: the coarse operator's, scaled
    by h^2), and a 3-D network that cannot lower the FP64 residual is
    rejected with the plain result kept."""
    n = 32
    w = W.poisson_3d(n, variable=True)
    h = 1.0 / n
    mg = Multigrid(3, n, n, n, h, h, h, "neumann")
    mg.set_density(w["rho"])
    mg.set_rhs(w["rhs"])
    mg.solve(rel_tol=1e-10, precision=PREC_FP64, accel=1)
    p = mg.sol(0)
    level, bs = 2, 4
    f, t, c = mg.learned_features(level, with_target=True, with_faces=True)
    b = w["rhs"].reshape(n, n, n)
    sigma = np.sqrt(np.mean(b ** 2))

    def bm(a):
        m = n // bs
        return a.reshape(m, bs, m, bs, m, bs).mean(axis=(1, 3, 5)).ravel()

    np.testing.assert_allclose(f[:, 0], bm(b) / sigma, rtol=1e-5, atol=1e-6)
    np.testing.assert_allclose(f[:, 1], bm(np.log(w["rho"].reshape(n, n, n))) / np.log(1000.0), rtol=1e-5, atol=1e-6)
    np.testing.assert_allclose(t, bm(p.reshape(n, n, n)) / sigma, rtol=1e-5, atol=1e-6 * np.abs(t).max())
    m = n // bs
    cz = c[:, 2].reshape(m, m, m)
    assert np.all(cz[-1] == 0.0) and np.all(cz[:-1] > 0.0)
    assert np.all(c[:, 0].reshape(m, m, m)[:, :, -1] == 0.0)
    # the plane's 1000:1 jump: faces inside the light layer are ~1000x those in the heavy one
    assert cz[0].mean() / cz[-2].mean() > 100 or cz[-2].mean() / cz[0].mean() > 100
    plain = mg.solve(rel_tol=1e-8, precision=PREC_MIXED, accel=1)
    x_plain = mg.sol(0)
    net = UNet(m, m, 1, depth=m)
    net.set_params(np.full(net.num_params, np.nan, np.float32))
    mg.attach_learned(net, level)
    st = mg.solve(rel_tol=1e-8, precision=PREC_MIXED, accel=1, use_learned=3)
    assert st["converged"] and st["learned_rejected"] == 1 and st["learned_accepted"] == 0
    assert st["cycles"] == plain["cycles"]
    assert rel(mg.sol(0), x_plain) < 1e-12
    net2 = UNet(m, m, 1)  # a 2-D network does not fit a 3-D solver
    with pytest.raises(ScihpcError):
        mg.attach_learned(net2, level)


@pytest.mark.parametrize("ndim", [2, 3])
def test_unet_smoothed_energy_loss(ndim):
    """Energy loss through 2 damped-Jacobi sweeps of the error (and their
    transposes in the gradient) against the oracle."""
    shape = (16, 24) if ndim == 2 else (8, 8, 16)
    n = 2
    pix = int(np.prod(shape))
    rng = np.random.default_rng(50 + ndim)
    p = U.init_params(51, ndim=ndim).astype(np.float32)
    feat = rng.standard_normal((n, pix, 2)).astype(np.float32)
    tgt = rng.standard_normal((n, pix)).astype(np.float32)
    fc = rng.uniform(0.001, 1.0, (n,) + shape + (ndim,)).astype(np.float32)
    for a in range(ndim):  # walls: the last cell along axis a (x, y, z order) has no +a face
        idx = [slice(None)] * (ndim + 2)
        idx[ndim - a] = -1
        idx[-1] = a
        fc[tuple(idx)] = 0.0
    net = UNet(shape[-2], shape[-1], n, depth=shape[0] if ndim == 3 else None)
    net.set_params(p)
    net.out_scale = 1.0
    net.set_samples(feat, tgt, fc.reshape(n, pix, ndim))
    net.set_loss_smoothing(2, 0.8)
    net.forward(np.arange(n))
    loss = net.backward()
    g = net.grads()
    inp = feat.reshape((n,) + shape + (2,)).astype(np.float64)
    l_ref, g_ref, _ = U.loss_and_grads(p.astype(np.float64), inp, tgt.reshape((n,) + shape).astype(np.float64),
                                       faces=fc.astype(np.float64), smooth=2, omega=0.8)
    l_plain, _, _ = U.loss_and_grads(p.astype(np.float64), inp, tgt.reshape((n,) + shape).astype(np.float64),
                                     faces=fc.astype(np.float64))
    _, g_exact, _ = U.loss_and_grads(p.astype(np.float64), inp, tgt.reshape((n,) + shape).astype(np.float64),
                                     faces=fc.astype(np.float64), smooth=2, omega=0.8, emulate=False)
    print(f"smoothed loss {loss:.4e} oracle {l_ref:.4e} (unsmoothed {l_plain:.4e})", flush=True)
    assert abs(loss - l_ref) < 1e-2 * l_ref
    assert l_ref < l_plain  # the sweeps remove error energy
    _grads_close(g, g_ref, g_exact, ndim)
    with pytest.raises(ScihpcError):
        net.set_loss_smoothing(-1, 0.8)
