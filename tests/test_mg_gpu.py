"""GPU parity of the multigrid (HIP through the C ABI) against the oracle.

Bars: integer/shape work exact; FP64 operator and residual within 1e-13
relative (FMA contraction on the GPU is the only difference); iterate-level
V-cycle parity within 1e-10; converged pressure within 1e-10 relative L2 of
the oracle's dense solve; the mixed FP32/FP64 path reaches the requested
FP64 relative residual.
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from scihpc_amd import PREC_FP64, PREC_MIXED, Multigrid, ScihpcError
from scihpc_amd import workloads as W


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


CASES = [
    # ndim, (nx, ny, nz), spacing, bc, variable
    (2, (48, 96, 1), (1 / 48, 2 / 96, 1.0), "neumann", False),
    (2, (64, 64, 1), (1 / 64, 1 / 64, 1.0), "dirichlet", False),
    (2, (40, 80, 1), (1 / 40, 2 / 80, 1.0), "neumann", True),
    (3, (16, 16, 16), (1 / 16, 1 / 16, 1 / 16), "neumann", False),
    (3, (24, 24, 24), (1 / 24, 1 / 24, 1 / 24), "dirichlet", False),
    (3, (16, 16, 32), (1 / 16, 1 / 16, 1 / 16), "neumann", True),
    (2, (54, 54, 1), (1 / 54, 1 / 54, 1.0), "neumann", False),  # ratio-3 coarsening
    # z-streaming 3-D sweeps on levels 0 and 1 (density and face coefficients)
    (3, (32, 32, 32), (1 / 32, 1 / 32, 1 / 32), "neumann", True),
    (3, (32, 48, 32), (1 / 32, 1 / 32, 1 / 32), "dirichlet", False),
]


def make_pair(oracle, case, seed=0):
    ndim, (nx, ny, nz), (dx, dy, dz), bc, var = case
    g = Multigrid(ndim, nx, ny, nz, dx, dy, dz, bc)
    o = oracle.mg(ndim, nx, ny, nz, dx, dy, dz, bc)
    rho = None
    if var:
        rng = np.random.default_rng(seed + 100)
        # two-phase density with a 1000:1 jump on a random blob
        n = nx * ny * nz
        idx = np.arange(n)
        x = (idx % nx + 0.5) / nx
        y = ((idx // nx) % ny + 0.5) / ny
        phi = 0.3 - np.hypot(x - rng.uniform(0.3, 0.7), y - rng.uniform(0.3, 0.7))
        rho = W.density(phi, 1.5 / nx, 1000.0)
        g.set_density(rho)
        o.set_rho(rho)
    return g, o, rho


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"{c[0]}d-{c[1]}-{c[3]}-{'var' if c[4] else 'const'}")
def test_levels_and_operator(oracle, case):
    g, o, _ = make_pair(oracle, case)
    assert g.level_num == o.L
    for l in range(g.level_num):
        assert g.shapes[l] == o.shapes[l][:4]
    rng = np.random.default_rng(1)
    for l in range(g.level_num):
        x = rng.standard_normal(g.n[l])
        y_g = g.apply(x, l)
        y_o = o.apply(l, x)
        assert rel(y_g, y_o) < 1e-13, l


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"{c[0]}d-{c[1]}-{c[3]}-{'var' if c[4] else 'const'}")
def test_residual_parity(oracle, case):
    g, o, _ = make_pair(oracle, case)
    rng = np.random.default_rng(2)
    n = g.n[0]
    b, x = rng.standard_normal(n), rng.standard_normal(n)
    g.set_rhs(b)
    g.set_sol(x)
    o.vec(0, "rhs")[:] = b
    o.vec(0, "sol")[:] = x
    rg = g.residual(0)
    ro = o.residual(0)
    assert abs(rg - ro) / ro < 1e-13
    assert rel(g.res(0), o.vec(0, "res")) < 1e-13


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"{c[0]}d-{c[1]}-{c[3]}-{'var' if c[4] else 'const'}")
def test_vcycle_iterates_match_oracle(oracle, case):
    g, o, _ = make_pair(oracle, case)
    rng = np.random.default_rng(3)
    b = rng.standard_normal(g.n[0])
    if case[3] == "neumann":
        b -= b.mean()
    g.set_rhs(b)
    g.set_sol(np.zeros_like(b))
    o.vec(0, "rhs")[:] = b
    o.vec(0, "sol")[:] = 0
    for _ in range(3):
        g.v_cycle(2, 2)
        oracle.lib.orc_mg_vcycle_rb(o.h, 2, 2)
        assert rel(g.sol(0), o.vec(0, "sol")) < 1e-10


DENSE_CASES = [
    (2, (32, 64, 1), (1 / 32, 2 / 64, 1.0), "neumann", False),
    (2, (40, 40, 1), (1 / 40, 1 / 40, 1.0), "dirichlet", False),
    (2, (32, 48, 1), (1 / 32, 1.5 / 48, 1.0), "neumann", True),
    (3, (8, 8, 32), (1 / 8, 1 / 8, 1 / 8), "neumann", True),
]


@pytest.mark.parametrize("precision", [PREC_FP64, PREC_MIXED], ids=["fp64", "mixed"])
@pytest.mark.parametrize("case", DENSE_CASES, ids=["2d-neu", "2d-dir", "2d-var", "3d-var"])
def test_converged_pressure_vs_dense(oracle, case, precision):
    g, o, _ = make_pair(oracle, case)
    rng = np.random.default_rng(4)
    b = rng.standard_normal(g.n[0])
    if case[3] == "neumann":
        b -= b.mean()
    g.set_rhs(b)
    x_ref = o.dense_solve(b)
    tol = 1e-13 if precision == PREC_FP64 else 1e-8
    st = g.solve(rel_tol=tol, max_cycles=400, precision=precision)
    assert st["converged"], st
    assert st["rel_residual"] <= tol
    x = g.sol(0)
    if case[3] == "neumann":
        assert abs(x.mean()) < 1e-12 * np.abs(x).max()
    # error vs the FP64 dense oracle: bounded by cond(A) * residual
    err = rel(x, x_ref)
    if precision == PREC_FP64:
        assert err < 1e-10, err
    else:
        assert err < 1e-5, err


def test_c1_analytic_first_order(oracle):
    """2-D Dirichlet Poisson, sin(pi x) sin(pi y): the reference operator is
    first order at the wall (ghost centre = 0), relative L2 error 9.29e-3 at
    256^2 and halving with h."""
    errs = []
    for n in (128, 256):
        w = W.poisson_dirichlet_2d(n)
        g = Multigrid(2, n, n, 1, w["spacing"][0], w["spacing"][1], 1.0, "dirichlet")
        g.set_rhs(w["rhs"])
        st = g.solve(rel_tol=1e-10, max_cycles=100, precision=PREC_FP64)
        assert st["converged"]
        x = g.sol(0)
        errs.append(rel(x, w["exact"]))
    assert abs(errs[1] - 9.29e-3) < 0.05e-3, errs
    assert 1.8 < errs[0] / errs[1] < 2.2, errs


def test_reference_faithful_vcycle(oracle):
    """GPU steepest-descent V(10,10) (multigrid.cpp:43-55) follows the
    oracle's iterates, which follow the compiled reference."""
    n = 72
    g = Multigrid(2, n, n, 1, 1.0 / n, 1.0 / n, 1.0, "dirichlet")
    o = oracle.mg(2, n, n, 1, 1.0 / n, 1.0 / n, 1.0, "dirichlet")
    rng = np.random.default_rng(5)
    b = rng.standard_normal(n * n)
    g.set_rhs(b)
    g.set_sol(np.zeros(n * n))
    o.vec(0, "rhs")[:] = b
    for _ in range(2):
        g.v_cycle_ref()
        oracle.lib.orc_mg_vcycle_ref(o.h)
        assert rel(g.sol(0), o.vec(0, "sol")) < 1e-9
        assert abs(g.residual(0) - o.residual(0)) / o.residual(0) < 1e-8


def test_reference_faithful_neumann_fullcycle(oracle):
    nx, ny = 32, 64
    g = Multigrid(2, nx, ny, 1, 1.0 / nx, 2.0 / ny, 1.0, "neumann")
    o = oracle.mg(2, nx, ny, 1, 1.0 / nx, 2.0 / ny, 1.0, "neumann")
    rng = np.random.default_rng(6)
    b = rng.standard_normal(nx * ny)
    b -= b.mean()
    g.set_rhs(b)
    g.set_sol(np.zeros(nx * ny))
    o.vec(0, "rhs")[:] = b
    g.full_cycle_ref()
    oracle.lib.orc_mg_fullcycle_ref(o.h)
    assert rel(g.sol(0), o.vec(0, "sol")) < 1e-8


def test_bubble_c2_shape_mixed_and_fp64(oracle):
    """The 512x1024 bubble workload: plain FP64 multigrid, mixed FP32/FP64
    and both with flexible CG reach 1e-8 and agree within the conditioning
    bound.  Against the oracle's own red-black multigrid both sides are
    driven to 1e-10 (the FP64 floor of this 1000:1 operator: |p| ~ 10 and
    coefficients spanning 1e-3..1 x 1/h^2 stall the relative residual near
    1e-10)."""
    w = W.bubble_2d()
    nx, ny, _ = w["shape"]
    g = Multigrid(2, nx, ny, 1, w["spacing"][0], w["spacing"][1], 1.0, "neumann")
    g.set_density(w["rho"])
    g.set_rhs(w["rhs"])
    sols = {}
    for prec in (PREC_FP64, PREC_MIXED):
        for acc in (0, 1):
            st = g.solve(rel_tol=1e-8, max_cycles=200, precision=prec, accel=acc)
            assert st["converged"] and st["rel_residual"] <= 1e-8, (prec, acc, st)
            sols[(prec, acc)] = g.sol(0)
    ref = sols[(PREC_FP64, 0)]
    for k, x in sols.items():
        assert rel(x, ref) < 1e-5, k
    assert g.solve(rel_tol=1e-8, precision=PREC_MIXED, accel=1)["cycles"] < \
        g.solve(rel_tol=1e-8, precision=PREC_MIXED, accel=0)["cycles"]
    o = oracle.mg(2, nx, ny, 1, w["spacing"][0], w["spacing"][1], 1.0, "neumann")
    o.set_rho(w["rho"])
    o.vec(0, "rhs")[:] = w["rhs"]
    o.solve_rb(1e-10, 100)
    s = g.solve(rel_tol=1e-10, max_cycles=100, precision=PREC_FP64, accel=1)
    assert s["converged"]
    assert rel(g.sol(0), o.vec(0, "sol")) < 1e-6


@pytest.mark.parametrize("precision", [PREC_FP64, PREC_MIXED], ids=["fp64", "mixed"])
@pytest.mark.parametrize("case", DENSE_CASES, ids=["2d-neu", "2d-dir", "2d-var", "3d-var"])
def test_flexible_cg_vs_dense(oracle, case, precision):
    g, o, _ = make_pair(oracle, case)
    rng = np.random.default_rng(7)
    b = rng.standard_normal(g.n[0])
    if case[3] == "neumann":
        b -= b.mean()
    g.set_rhs(b)
    x_ref = o.dense_solve(b)
    tol = 1e-12 if precision == PREC_FP64 else 1e-8
    st = g.solve(rel_tol=tol, max_cycles=200, precision=precision, accel=1)
    assert st["converged"], st
    err = rel(g.sol(0), x_ref)
    assert err < (1e-9 if precision == PREC_FP64 else 1e-5), err
    # warm start from the converged answer: zero iterations
    st2 = g.solve(rel_tol=max(tol, 10 * st["rel_residual"]), precision=precision, accel=1, zero_guess=False)
    assert st2["cycles"] == 0


def test_errors_are_status_codes():
    with pytest.raises(ScihpcError) as e:
        Multigrid(4, 8, 8, 8)
    assert e.value.code == 2
    g = Multigrid(2, 64, 64, 1, 1 / 64, 1 / 64, 1.0, "dirichlet")
    with pytest.raises(ScihpcError):
        g.set_density(np.ones(64 * 64))  # A(rho) needs wall faces
    g.set_rhs(np.random.default_rng(0).standard_normal(64 * 64))
    st = g.solve(rel_tol=1e-12, max_cycles=1)
    assert not st["converged"] and st["cycles"] == 1


@pytest.mark.parametrize("case", [c for c in CASES if c[0] == 3],
                         ids=lambda c: f"{c[1]}-{c[3]}-{'var' if c[4] else 'const'}")
@pytest.mark.parametrize("precision", [PREC_FP64, PREC_MIXED], ids=["fp64", "mixed"])
def test_z3_sweeps_match_colour_kernels(oracle, case, precision, monkeypatch):
    """The z-streaming 3-D sweep is the same red-black Gauss-Seidel as the
    two colour kernels: equal V-cycle iterates (FP64 to rounding) and equal
    solve results."""
    rng = np.random.default_rng(7)
    sols, stats = [], []
    for flag in ("1", "0"):
        monkeypatch.setenv("SCIHPC_MG_TILED3D", flag)
        g, _, _ = make_pair(oracle, case)
        b = rng.standard_normal(g.n[0]) if not sols else b
        if case[3] == "neumann":
            b = b - b.mean()
        g.set_rhs(b)
        g.set_sol(np.zeros_like(b))
        for _ in range(2):
            g.v_cycle(2, 2)
        it = g.sol(0)
        st = g.solve(rel_tol=1e-8, max_cycles=100, precision=precision, accel=1)
        sols.append((it, g.sol(0)))
        stats.append(st)
    assert rel(sols[0][0], sols[1][0]) < 1e-12
    assert stats[0]["converged"] and stats[1]["converged"]
    assert abs(stats[0]["cycles"] - stats[1]["cycles"]) <= (0 if precision == PREC_FP64 else 1)
    assert rel(sols[0][1], sols[1][1]) < (1e-10 if precision == PREC_FP64 else 1e-6)


@pytest.mark.parametrize("variable", [True, False], ids=["var", "const"])
def test_flexible_cg_3d_z_march(variable):
    """3-D levels of 2^18 cells and more run the flexible-CG direction kernel
    in its z-marching form: FP64 and mixed flexible CG reach the tolerance,
    in no more iterations than the stationary solve, and agree with it."""
    n = 64
    if variable:
        w = W.poisson_3d(n, variable=True)
        g = Multigrid(3, n, n, n, 1 / n, 1 / n, 1 / n, "neumann")
        g.set_density(w["rho"])
    else:
        w = W.poisson_3d(n, variable=False)
        g = Multigrid(3, n, n, n, 1 / n, 1 / n, 1 / n, "dirichlet")
    g.set_rhs(w["rhs"])
    print(f"z-march FCG 3-D {n}^3 variable={variable}", flush=True)
    ref = g.solve(rel_tol=1e-10, max_cycles=200, precision=PREC_FP64, accel=0)
    assert ref["converged"]
    x_ref = g.sol(0)
    for prec in (PREC_FP64, PREC_MIXED):
        st = g.solve(rel_tol=1e-10, max_cycles=200, precision=prec, accel=1)
        assert st["converged"] and st["rel_residual"] <= 1e-10, st
        assert st["cycles"] <= ref["cycles"]
        assert rel(g.sol(0), x_ref) < 1e-8


@pytest.mark.parametrize("ndim", [2, 3])
def test_timing_hooks(ndim):
    """The level-0 timing hooks run every mode, report the algorithmic bytes
    of the prolongation and residual passes, reject unknown modes, and leave
    the handle able to solve."""
    if ndim == 2:
        n = (128, 256, 1)
        w = W.bubble_2d(*n[:2])
        g = Multigrid(2, *n, 1 / n[0], 2 / n[1], 1.0, "neumann")
    else:
        n = (64, 64, 64)
        w = W.poisson_3d(64, variable=True)
        g = Multigrid(3, *n, 1 / 64, 1 / 64, 1 / 64, "neumann")
    g.set_density(w["rho"])
    g.set_rhs(w["rhs"])
    print(f"timing hooks {ndim}-D {n}", flush=True)
    assert g.solve(rel_tol=1e-8, max_cycles=100, precision=PREC_MIXED, accel=1)["converged"]
    cells = int(np.prod(n))
    coarse = int(np.prod([max(x // 2, 1) for x in n]))
    for which in range(6):
        ms, b = g.time_kernel(which, iters=2)
        assert math.isfinite(ms) and ms > 0 and b > 0
    assert g.time_kernel(4, iters=1)[1] == 16.0 * cells + 8.0 * coarse
    assert g.time_kernel(5, iters=1)[1] == (2 + ndim) * 8.0 * cells
    if ndim == 2:  # the fused FCG update + pre-smoothing, on the last mixed FCG solve's vectors
        ms, b = g.time_kernel(6, iters=2)
        assert math.isfinite(ms) and ms > 0 and b == (4 * ndim + 56) * cells + 8.0 * coarse
    else:
        with pytest.raises(ScihpcError):
            g.time_kernel(6, iters=1)
    with pytest.raises(ScihpcError):
        g.time_kernel(7, iters=1)
    st = g.solve(rel_tol=1e-8, max_cycles=100, precision=PREC_MIXED, accel=1)
    assert st["converged"], st


@pytest.mark.parametrize("precision,accel", [(PREC_FP64, 0), (PREC_MIXED, 0), (PREC_FP64, 1), (PREC_MIXED, 1)],
                         ids=["fp64-stationary", "mixed-stationary", "fp64-fcg", "mixed-fcg"])
def test_lagged_check(precision, accel):
    """lagged_check: the same answer as reading the residual every iteration
    once the count is known; after a harder solve, an easier start (a
    half-converged guess) still meets the tolerance, overshooting by the
    lag, and the count comes back down over the next solves."""
    nx, ny = 128, 256
    w = W.bubble_2d(nx, ny)
    g = Multigrid(2, nx, ny, 1, 1 / nx, 2 / ny, 1.0, "neumann")
    g.set_density(w["rho"])
    g.set_rhs(w["rhs"])
    print(f"lagged check {precision} accel={accel}", flush=True)
    kw = dict(rel_tol=1e-9, max_cycles=200, precision=precision, accel=accel)
    ref = g.solve(**kw)
    assert ref["converged"]
    x_ref = g.sol(0)
    for _ in range(3):
        st = g.solve(lagged_check=1, **kw)
        assert st["converged"] and st["rel_residual"] <= 1e-9, st
        assert st["cycles"] == ref["cycles"]
        assert rel(g.sol(0), x_ref) < 1e-12
    # an easier problem: the same b from a half-converged start
    g.solve(**{**kw, "max_cycles": ref["cycles"] // 2})
    x_half = g.sol(0)
    g.set_sol(x_half)
    want = g.solve(zero_guess=False, **kw)
    assert want["converged"] and want["cycles"] < ref["cycles"] - 2
    g.solve(lagged_check=1, **kw)  # the count of the zero start again
    counts = []
    for _ in range(ref["cycles"] - want["cycles"] + 2):
        g.set_sol(x_half)
        st = g.solve(zero_guess=False, lagged_check=1, **kw)
        assert st["converged"] and st["rel_residual"] <= 1e-9, st
        assert rel(g.sol(0), x_ref) < 1e-8
        counts.append(st["cycles"])
    assert counts[0] == ref["cycles"] - 1  # read first one iteration before the old count
    assert counts[-1] == want["cycles"]
    assert all(b <= a for a, b in zip(counts, counts[1:]))
    with pytest.raises(ScihpcError):
        g.solve(lagged_check=2, **kw)


def test_lagged_check_density_changes():
    """Lagged checks in a time loop whose density changes between solves:
    every solve meets the tolerance; a harder operator (1000:1 after 10:1)
    converges at its own count at once (the residual is read every iteration
    from one before the remembered count); an easier one overshoots at first
    and is back at its own count within a few solves."""
    nx, ny = 128, 256
    hard, easy = W.bubble_2d(nx, ny), W.bubble_2d(nx, ny, ratio=10.0)
    g = Multigrid(2, nx, ny, 1, 1 / nx, 2 / ny, 1.0, "neumann")
    kw = dict(rel_tol=1e-9, max_cycles=200, precision=PREC_MIXED, accel=1)
    own = {}
    for name, w in (("hard", hard), ("easy", easy)):
        g.set_density(w["rho"])
        g.set_rhs(w["rhs"])
        own[name] = g.solve(**kw)["cycles"]
    assert own["easy"] < own["hard"] - 2, own
    seq = ["hard"] * 3 + ["easy"] * 8 + ["hard"] * 2
    got = []
    for name in seq:
        w = hard if name == "hard" else easy
        g.set_density(w["rho"])
        g.set_rhs(w["rhs"])
        st = g.solve(lagged_check=1, **kw)
        assert st["converged"] and st["rel_residual"] <= 1e-9, (name, st)
        got.append(st["cycles"])
    print(f"own counts {own}, lagged sequence {got}", flush=True)
    assert all(c >= own[n] for c, n in zip(got, seq))
    assert got[10] == own["easy"]
    assert got[11] == own["hard"]


@pytest.mark.parametrize("lagged", [0, 1], ids=["checked", "lagged"])
def test_fused_fcg_update_matches_separate_kernels(monkeypatch, lagged):
    """The mixed 2-D FCG update fused into the next level-0 pre-smoothing
    (the default) gives the iterates of the separate update kernel: the same
    iteration counts and a bit-identical pressure, also when max_cycles stops
    the solve."""
    nx, ny = 128, 256
    w = W.bubble_2d(nx, ny)
    out = {}
    for fused in ("1", "0"):
        monkeypatch.setenv("SCIHPC_MG_FUSED_UPDATE", fused)
        print(f"fused FCG update={fused} lagged={lagged}", flush=True)
        g = Multigrid(2, nx, ny, 1, 1 / nx, 2 / ny, 1.0, "neumann")
        g.set_density(w["rho"])
        g.set_rhs(w["rhs"])
        kw = dict(rel_tol=1e-9, precision=PREC_MIXED, accel=1, lagged_check=lagged)
        sts = [g.solve(max_cycles=200, **kw) for _ in range(2)]
        x_conv = g.sol(0)
        capped = g.solve(max_cycles=3, **kw)
        out[fused] = (sts, x_conv, capped, g.sol(0))
        g.close()
    (s1, x1, c1, y1), (s0, x0, c0, y0) = out["1"], out["0"]
    for a, b in zip(s1, s0):
        assert a["converged"] and b["converged"] and a["rel_residual"] <= 1e-9
        assert a["cycles"] == b["cycles"]
        assert abs(a["rel_residual"] - b["rel_residual"]) <= 1e-12 * b["rel_residual"]
    np.testing.assert_array_equal(x1, x0)
    assert not c1["converged"] and c1["cycles"] == c0["cycles"] == 3
    np.testing.assert_array_equal(y1, y0)
