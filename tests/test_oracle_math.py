"""Pins of the oracle that do not rest on the reference's outputs: closed
forms, invariants, brute force on tiny inputs.  Each check is chosen so that
a dropped term, a wrong sign/index or a transposed operand fails it."""
import math

import numpy as np
import pytest

from oracle.pyoracle import padded_shape


def test_level_lists(oracle):
    # multigrid.cpp:8-41 coarsening rule on the configuration shapes
    assert [l[:2] for l in oracle.levels(2, 512, 1024)][-1] == (4, 8)
    assert len(oracle.levels(2, 512, 1024)) == 8
    assert len(oracle.levels(2, 1024, 512)) == 8
    assert len(oracle.levels(3, 256, 256, 256)) == 7
    assert len(oracle.levels(3, 128, 128, 128)) == 6
    assert [l[0] for l in oracle.levels(2, 144, 144)] == [144, 72, 36, 18, 9]
    assert len(oracle.levels(2, 257, 257)) == 1          # odd size never coarsens
    assert [l[3] for l in oracle.levels(2, 54, 54)] == [1, 2, 6]  # 54 -> 27 -> 9 (ratio 3)


def dense(m, level=0):
    n = m.shapes[level][4]
    return np.array([m.apply(level, e) for e in np.eye(n)]).T


@pytest.mark.parametrize("bc", ["neumann", "dirichlet"])
def test_operator_closed_form(oracle, bc):
    nx, ny, dx, dy = 6, 5, 0.2, 0.3
    m = oracle.mg(2, nx, ny, 1, dx, dy, 1.0, bc)
    A = dense(m)
    assert np.allclose(A, A.T, rtol=0, atol=1e-12)
    ax, ay = 1 / dx / dx, 1 / dy / dy
    # interior row: 5-point Laplacian
    c = 2 + nx * 2
    assert A[c, c] == pytest.approx(-2 * ax - 2 * ay)
    assert A[c, c - 1] == A[c, c + 1] == pytest.approx(ax)
    assert A[c, c - nx] == A[c, c + nx] == pytest.approx(ay)
    # corner row: Neumann drops the wall terms, Dirichlet keeps them in the diagonal
    corner = -ax - ay if bc == "neumann" else -2 * ax - 2 * ay
    assert A[0, 0] == pytest.approx(corner)
    if bc == "neumann":
        assert np.allclose(A @ np.ones(nx * ny), 0, atol=1e-10)
    else:
        assert np.all(np.linalg.eigvalsh(A) < 0)


def test_variable_operator_closed_form(oracle):
    nx, ny, dx, dy = 8, 8, 0.25, 0.5
    rng = np.random.default_rng(0)
    rho = rng.uniform(1, 1000, nx * ny)
    m = oracle.mg(2, nx, ny, 1, dx, dy, 1.0, "neumann")
    m.set_rho(rho)
    A = dense(m)
    R = rho.reshape(ny, nx)
    for j in range(ny):
        for i in range(nx):
            c = i + nx * j
            if i + 1 < nx:
                assert A[c, c + 1] == pytest.approx(2.0 / (R[j, i] + R[j, i + 1]) / dx ** 2)
            if j + 1 < ny:
                assert A[c, c + nx] == pytest.approx(2.0 / (R[j, i] + R[j + 1, i]) / dy ** 2)
    assert np.allclose(A, A.T)
    assert np.allclose(A @ np.ones(nx * ny), 0, atol=1e-9)
    # coarse level: face-averaged coefficients keep symmetry and the null space
    Ac = dense(m, 1)
    assert np.allclose(Ac, Ac.T) and np.allclose(Ac @ np.ones(Ac.shape[0]), 0, atol=1e-9)


def test_dense_solve_against_numpy(oracle):
    m = oracle.mg(2, 8, 6, 1, 0.1, 0.2, 1.0, "dirichlet")
    A = dense(m)
    b = np.random.default_rng(1).standard_normal(48)
    assert np.allclose(m.dense_solve(b), np.linalg.solve(A, b), rtol=1e-12, atol=1e-14)
    m2 = oracle.mg(2, 8, 6, 1, 0.1, 0.2, 1.0, "neumann")
    A2 = dense(m2)
    b2 = b - b.mean()
    x = m2.dense_solve(b2)
    assert abs(x.mean()) < 1e-12 and np.allclose(A2 @ x, b2, atol=1e-9)


@pytest.mark.parametrize("bc", ["neumann", "dirichlet"])
def test_rb_multigrid_converges_to_dense(oracle, bc):
    nx, ny = 32, 48
    m = oracle.mg(2, nx, ny, 1, 1 / nx, 1.5 / ny, 1.0, bc)
    rng = np.random.default_rng(2)
    rho = None
    if bc == "neumann":  # 1000:1 bubble with a smoothed interface
        i, j = np.meshgrid(np.arange(nx), np.arange(ny), indexing="xy")
        phi = 0.3 - np.hypot((i + 0.5) / nx - 0.45, (j + 0.5) * 1.5 / ny - 0.8)
        from scihpc_amd.workloads import density
        rho = density(phi.ravel(), 1.5 / nx, 1000.0)
    if rho is not None:
        m.set_rho(rho)
    b = rng.standard_normal(nx * ny)
    if bc == "neumann":
        b -= b.mean()
    x = m.dense_solve(b)
    m.vec(0, "rhs")[:] = b
    c, hist = m.solve_rb(1e-13, 300)
    assert hist[-1] <= 1e-13
    assert np.linalg.norm(m.vec(0, "sol") - x) / np.linalg.norm(x) < 1e-9


def _bubble_rho(nx, ny, ly):
    from scihpc_amd.workloads import density
    i, j = np.meshgrid(np.arange(nx), np.arange(ny), indexing="xy")
    phi = 0.3 - np.hypot((i + 0.5) / nx - 0.45, (j + 0.5) * ly / ny - 0.8)
    return density(phi.ravel(), 1.5 / nx, 1000.0)


@pytest.mark.parametrize("bc", ["neumann", "dirichlet"])
def test_fcg_converges_to_dense(oracle, bc):
    """FCG(1) around the red-black V-cycle reaches the dense solution of the
    same operator, in fewer iterations than the stationary V-cycle.  The
    tolerance stays above the FP64 floor (about 1e-13 here): below it the
    recurrence residual of FCG(1) on the singular Neumann operator grows
    again (the V-cycle's constant component enters the direction)."""
    nx, ny = 32, 48
    m = oracle.mg(2, nx, ny, 1, 1 / nx, 1.5 / ny, 1.0, bc)
    if bc == "neumann":
        m.set_rho(_bubble_rho(nx, ny, 1.5))
    b = np.random.default_rng(12).standard_normal(nx * ny)
    if bc == "neumann":
        b -= b.mean()
    x = m.dense_solve(b)
    m.vec(0, "rhs")[:] = b
    it, hist = m.solve_fcg(1e-12, 200)
    assert hist[-1] <= 1e-12 and np.all(hist[:-1] > 1e-12)
    assert np.linalg.norm(m.vec(0, "sol") - x) / np.linalg.norm(x) < 1e-10
    assert np.array_equal(m.vec(0, "rhs"), b)  # b restored
    m.vec(0, "sol")[:] = 0
    cyc, _ = m.solve_rb(1e-12, 300)
    assert it < cyc


def test_fcg_invariants(oracle):
    """After every FCG(1) step: the new direction is A-conjugate to the
    previous one, the updated residual is orthogonal to the direction (both
    with the direction's constant component removed, the step sizes' form on
    the all-Neumann operator), the recurrence residual equals b - A x, and
    the first iterate is the energy-optimal multiple of z = V(b)."""
    nx, ny = 24, 40
    m = oracle.mg(2, nx, ny, 1, 1 / nx, 2 / ny, 1.0, "neumann")
    m.set_rho(_bubble_rho(nx, ny, 2.0))
    A = dense(m)
    b = np.random.default_rng(13).standard_normal(nx * ny)
    b -= b.mean()
    # z = V(b) from the oracle's V-cycle alone
    m.vec(0, "rhs")[:] = b
    m.vec(0, "sol")[:] = 0
    oracle.lib.orc_mg_vcycle_rb(m.h, 2, 2)
    z1 = m.vec(0, "sol").copy()
    for k in range(1, 6):
        m.vec(0, "rhs")[:] = b
        it, _ = m.solve_fcg(1e-30, k)
        assert it == k
        x, r, p, q = (m.fcg_vec(w) for w in ("x", "r", "p", "q"))
        pc = p - p.mean()
        assert np.allclose(q, A @ p, rtol=0, atol=1e-9 * np.abs(q).max())
        assert abs(pc @ r) < 1e-10 * np.linalg.norm(pc) * np.linalg.norm(b)
        assert np.linalg.norm(r - (b - A @ x)) < 1e-9 * np.linalg.norm(b)
        if k == 1:
            zc = z1 - z1.mean()
            assert np.allclose(x, (zc @ b) / (zc @ (A @ z1)) * z1, rtol=1e-10, atol=1e-14)
        else:
            qo = m.fcg_vec("q_old")
            assert abs(pc @ qo) < 1e-10 * np.linalg.norm(pc) * np.linalg.norm(qo)


def test_fcg_does_not_drift_below_the_floor(oracle):
    """Asked for a residual below the FP64 floor, FCG(1) on the singular
    all-Neumann operator keeps b - A x at the floor (without the direction's
    constant component left out of the step sizes it grew back to about
    1e-6 within 20 iterations), and a requested 1e-12 is met by the true
    residual."""
    nx, ny = 32, 48
    m = oracle.mg(2, nx, ny, 1, 1 / nx, 1.5 / ny, 1.0, "neumann")
    m.set_rho(_bubble_rho(nx, ny, 1.5))
    A = dense(m)
    b = np.random.default_rng(12).standard_normal(nx * ny)
    b -= b.mean()
    m.vec(0, "rhs")[:] = b
    m.solve_fcg(1e-30, 40)
    x = m.vec(0, "sol")
    assert np.linalg.norm(b - A @ x) / np.linalg.norm(b) < 1e-12
    m.vec(0, "rhs")[:] = b
    it, _ = m.solve_fcg(1e-12, 100)
    x = m.vec(0, "sol")
    assert np.linalg.norm(b - A @ x) / np.linalg.norm(b) <= 1e-12 and it < 30


def test_rb_transfers(oracle):
    nx, ny = 16, 8
    m = oracle.mg(2, nx, ny, 1, 1.0, 1.0, 1.0, "neumann")
    # restriction of a constant residual is that constant (child average)
    m.vec(0, "rhs")[:] = 3.0
    m.vec(0, "sol")[:] = 0.0
    oracle.lib.orc_mg_restrict_rb(m.h, 0)
    assert np.allclose(m.vec(1, "rhs"), 3.0) and np.all(m.vec(1, "sol") == 0)
    # prolongation is exact for a linear coarse field on interior fine cells
    cx, cy = m.shapes[1][:2]
    I, J = np.meshgrid(np.arange(cx), np.arange(cy), indexing="xy")
    m.vec(1, "sol")[:] = (2.0 * (I + 0.5) * 2 + 5.0 * (J + 0.5) * 2).ravel()
    m.vec(0, "sol")[:] = 0.0
    oracle.lib.orc_mg_prolong_rb(m.h, 1)
    i, j = np.meshgrid(np.arange(nx), np.arange(ny), indexing="xy")
    exact = 2.0 * (i + 0.5) + 5.0 * (j + 0.5)
    f = m.vec(0, "sol").reshape(ny, nx)
    assert np.allclose(f[1:-1, 1:-1], exact[1:-1, 1:-1])


def test_gs_smoother_is_gauss_seidel(oracle):
    """One red-black sweep equals GS on the red cells then the black cells,
    brute force on the dense matrix."""
    nx, ny = 6, 4
    m = oracle.mg(2, nx, ny, 1, 1.0, 1.0, 1.0, "dirichlet")
    A = dense(m)
    rng = np.random.default_rng(3)
    b, x = rng.standard_normal(nx * ny), rng.standard_normal(nx * ny)
    m.vec(0, "rhs")[:] = b
    m.vec(0, "sol")[:] = x
    oracle.lib.orc_mg_smooth_rb(m.h, 0, 1, 0)
    y = x.copy()
    for color in (0, 1):
        for c in range(nx * ny):
            if ((c % nx) + (c // nx)) % 2 == color:
                y[c] = (b[c] - (A[c] @ y - A[c, c] * y[c])) / A[c, c]
    assert np.allclose(m.vec(0, "sol"), y, rtol=1e-13)


def test_weno_exact_on_linear_and_constant(oracle):
    n = (20, 16, 1)
    shp = padded_shape(2, *n)
    I, J = np.meshgrid(np.arange(shp[0]), np.arange(shp[1]), indexing="ij")
    dx, dy = 0.1, 0.2
    f = (1.5 * I * dx - 0.7 * J * dy + 3.0).ravel()
    for sgn in (1.0, -1.0):
        u = np.full(f.size, sgn)
        fx, fy, _ = oracle.weno5(2, n, (dx, dy, 1.0), f, u, -u, u)
        interior = np.zeros(shp, bool)
        interior[3:-3, 3:-3] = True
        assert np.allclose(fx.reshape(shp)[interior], 1.5, rtol=1e-12)
        assert np.allclose(fy.reshape(shp)[interior], -0.7, rtol=1e-12)
    fx, _, _ = oracle.weno5(2, n, (dx, dy, 1.0), np.full(f.size, 2.0), u, u, u)
    assert np.all(fx == 0.0)


def test_weno_fifth_order(oracle):
    errs = []
    for nx in (32, 64, 128):
        shp = padded_shape(1, nx)
        dx = 1.0 / nx
        x = (np.arange(shp[0]) - 3 + 0.5) * dx
        f = np.sin(2 * math.pi * x)
        fx, _, _ = oracle.weno5(1, (nx, 1, 1), (dx, 1, 1), f, np.ones_like(f), f * 0, f * 0)
        errs.append(np.abs(fx[3:-3] - 2 * math.pi * np.cos(2 * math.pi * x[3:-3])).max())
    orders = np.log2(np.array(errs[:-1]) / np.array(errs[1:]))
    assert orders[-1] > 4.5, orders


def test_rk3_closed_forms(oracle):
    n = (24, 20, 1)
    h = (0.1, 0.1, 1.0)
    shp = padded_shape(2, *n)
    # constant field is invariant
    phi = np.full(int(np.prod(shp)), 0.7)
    u = np.full_like(phi, 0.3)
    oracle.tvd_rk3(2, n, h, 0.01, ["neumann"] * 4, [0.0] * 4, phi, u, -u, u, 4)
    assert np.all(phi == 0.7)
    # linear field, constant velocity, periodic-free interior: exact shift
    I, J = np.meshgrid(np.arange(shp[0]), np.arange(shp[1]), indexing="ij")
    phi = (0.5 * I * 0.1 + 0.25 * J * 0.1).ravel()
    phi0 = phi.copy()
    dt = 0.01
    oracle.tvd_rk3(2, n, h, dt, ["dirichlet"] * 4, [0.0] * 4, phi, u, -u, u, 1)
    # away from the walls (BCs enter after the first stage) s = -(0.3*0.5 - 0.3*0.25)
    core = np.zeros(shp, bool)
    core[9:-9, 9:-9] = True
    expect = phi0 + dt * -(0.3 * 0.5 - 0.3 * 0.25)
    assert np.allclose(phi.reshape(shp)[core], expect.reshape(shp)[core], rtol=1e-12)


def test_density_closed_form(oracle):
    phi = np.array([-1.0, -0.05, 0.0, 0.05, 1.0, 0.025])
    rho = oracle.density(phi, 0.05, 1000.0, True)
    assert rho[0] == 1000.0 and rho[4] == 1.0 and rho[1] == 1000.0 and rho[3] == 1.0
    assert rho[2] == pytest.approx(500.5)
    H = 0.5 * (1 + 0.5 + math.sin(math.pi * 0.5) / math.pi)
    assert rho[5] == pytest.approx(H + (1 - H) * 1000.0)
    assert np.allclose(oracle.density(phi, 0.05, 0.2, False)[[0, 4]], [1.0, 0.2])


def test_bc_closed_forms(oracle):
    n = (5, 4, 1)
    shp = padded_shape(2, *n)
    f = np.random.default_rng(4).standard_normal(int(np.prod(shp)))
    A = f.reshape(shp)
    g = f.copy()
    oracle.bc_apply(g, 2, n, (0.5, 0.5, 1.0), ["neumann", "periodic", "no_slip", "slip"][0:1] * 4, [0.2] * 4)
    G = g.reshape(shp)
    # Neumann df=0.2, h=0.5: each ghost adds df*h outward
    assert np.allclose(G[8, 3:7], A[7, 3:7] + 0.1) and np.allclose(G[9, 3:7], A[7, 3:7] + 0.2)
    assert np.allclose(G[2, 3:7], A[3, 3:7] - 0.1)
    g = f.copy()
    oracle.bc_apply(g, 2, n, (0.5, 0.5, 1.0), ["periodic"] * 4, [0.0] * 4)
    G = g.reshape(shp)
    assert np.array_equal(G[8:11, 3:7], A[3:6, 3:7]) and np.array_equal(G[0:3, 3:7], A[5:8, 3:7])
    g = f.copy()
    oracle.bc_apply(g, 2, n, (0.5, 0.5, 1.0), ["dirichlet"] * 4, [1.0] * 4)
    G = g.reshape(shp)
    # last interior cell overwritten (rows j=4,5 are untouched by the y pass)
    assert np.allclose(G[7, 4:6], A[6, 4:6] / 3 + 2 / 3)
    assert np.allclose(G[8, 4:6], 2 * G[7, 4:6] - A[6, 4:6])
    g = f.copy()
    oracle.bc_apply(g, 2, n, (0.5, 0.5, 1.0), ["no_slip"] * 4, [0.0] * 4, stag=0)
    G = g.reshape(shp)
    assert np.all(G[7, 3:7] == 0.0) and np.all(G[2, 3:7] == 0.0)  # wall faces
    assert np.array_equal(G[8, 3:7], -A[6, 3:7])


def _compact_rhs_line(f, h, kind):
    """the right-hand sides of oracle.c's orc_compact_rhs, same expressions
    in IEEE doubles (so the same values)"""
    n = len(f)
    s, ss = [0.0] * n, [0.0] * n
    s[0] = (-3.5 * f[0] + 4.0 * f[1] + -0.5 * f[2]) / h
    ss[0] = (9.0 * f[0] + -12.0 * f[1] + 3.0 * f[2]) / h / h
    s[n - 1] = -(-3.5 * f[n - 1] + 4.0 * f[n - 2] + -0.5 * f[n - 3]) / h
    ss[n - 1] = (9.0 * f[n - 1] + -12.0 * f[n - 2] + 3.0 * f[n - 3]) / h / h
    c3, c2, c1 = -0.06096119008109, 1.99692238016218, -1.93596119008109
    for i in range(1, n - 1):
        if kind == 0:
            s[i] = 15.0 / 16.0 * (f[i + 1] - f[i - 1]) / h
        elif kind == 1:
            s[i] = (c1 * f[i - 1] + c2 * f[i] + c3 * f[i + 1]) / h
        else:
            s[i] = -(c3 * f[i - 1] + c2 * f[i] + c1 * f[i + 1]) / h
        ss[i] = (3.0 * f[i - 1] - 6.0 * f[i] + 3.0 * f[i + 1]) / h / h
    return s, ss


def _twin_exact(c, s, ss):
    """the twin back substitution in exact rational arithmetic"""
    from fractions import Fraction as Q
    n = len(s)
    a0, a1, a2, b0, b1, b2, aa0, aa1, aa2, bb0, bb1, bb2 = [[Q(x) for x in c[k * n:(k + 1) * n]] for k in range(12)]
    S, SS = [Q(x) for x in s], [Q(x) for x in ss]
    for i in range(1, n):
        S[i], SS[i] = S[i] - (a0[i] * S[i - 1] + b0[i] * SS[i - 1]), SS[i] - (aa0[i] * S[i - 1] + bb0[i] * SS[i - 1])
    for i in range(n - 1, -1, -1):
        si, ssi = S[i], SS[i]
        if i < n - 1:
            si = si - (a2[i] * S[i + 1] + b2[i] * SS[i + 1])
            ssi = ssi - (aa2[i] * S[i + 1] + bb2[i] * SS[i + 1])
        den = a1[i] * bb1[i] - aa1[i] * b1[i]
        S[i], SS[i] = -(b1[i] * ssi - bb1[i] * si) / den, (a1[i] * ssi - aa1[i] * si) / den
    return np.array([float(x) for x in S]), np.array([float(x) for x in SS])


@pytest.mark.parametrize("kind", [0, 1, 2], ids=["ccd", "uccd-up", "uccd-down"])
def test_compact_extended_sweep_against_exact(oracle, kind):
    """orc_compact_derivatives_ld (the twin sweep in long double) against the
    same factorised system solved in exact rational arithmetic on a rough
    1-D line: within a few units of the last place of the exact values,
    and never farther from them than the double sweep."""
    import ctypes as C
    L = oracle.lib
    dp = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
    L.orc_ccd_coefficients.argtypes = [C.c_int, C.c_double, dp]
    L.orc_uccd_coefficients.argtypes = [C.c_int, C.c_double, dp, dp]
    nx = 40
    n, h = nx + 6, 1.0 / nx
    rng = np.random.default_rng(5)
    f = np.sin(9.0 * np.arange(n) * h) + 0.2 * rng.standard_normal(n)
    c = np.zeros(12 * n)
    if kind == 0:
        L.orc_ccd_coefficients(n, h, c)
    else:
        other = np.zeros(12 * n)
        L.orc_uccd_coefficients(n, h, *((c, other) if kind == 1 else (other, c)))
    s, ss = _compact_rhs_line(list(f), h, kind)
    if kind == 2:  # the downwind ends are the upwind ones
        su, ssu = _compact_rhs_line(list(f), h, 1)
        s[0], ss[0], s[-1], ss[-1] = su[0], ssu[0], su[-1], ssu[-1]
    es, ess = _twin_exact(c, s, ss)
    vel = np.full(n, 1.0 if kind == 1 else -1.0)
    outs = {}
    for name, fn in (("double", L.orc_compact_derivatives), ("long", L.orc_compact_derivatives_ld)):
        d1, d2 = np.zeros(n), np.zeros(n)
        z = np.zeros(n)
        fn(1, nx, 1, 1, h, 1.0, 1.0, f, vel.ctypes.data if kind else None, None, None, int(kind > 0), d1, z, z,
           d2, z, z)
        outs[name] = d1
        if kind == 0:
            outs[name + "2"] = d2
    for key, want in (("", es), ("2", ess)):
        if key and kind:
            continue  # UCCD's second derivative mixes both systems
        err_d = np.max(np.abs(outs["double" + key] - want))
        err_l = np.max(np.abs(outs["long" + key] - want))
        scale = np.max(np.abs(want))
        assert err_l <= 4 * np.finfo(float).eps * scale, (key, err_l / scale)
        assert err_l <= err_d, (key, err_l, err_d)
