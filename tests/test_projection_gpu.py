"""GPU projection step (scihpc_proj_* through the C ABI) against the
reference's fast_pressure_correction fixtures and the oracle.

Bars: p_hat bit-exact; FPC pressure and velocities within 1e-10 relative of
the reference (the reference-faithful multigrid differs from the
reference's Eigen arithmetic at rounding level, amplified by the cycles);
the variable-density projection within 1e-8 of the oracle composite solved
to the same tolerance, and its corrected velocity divergence-free to the
solve tolerance.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from golden_util import load, spacing
from scihpc_amd import PREC_FP64, PREC_MIXED, Multigrid, Projection, padded_shape
from scihpc_amd import workloads as W

FPC = load("fpc")


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.mark.parametrize("case", sorted(FPC), ids=str)
def test_fpc_step_matches_reference(case):
    c = FPC[case]
    ndim, n = int(c["ndim"]), [int(x) for x in c["n"]]
    dx, dy, dz = spacing(c)
    print(f"fpc step {case}", flush=True)
    mg = Multigrid(ndim, *n, dx, dy, dz, "neumann")
    pj = Projection(ndim, *n, dx, dy, dz, [int(t) for t in c["types"]])
    u, v, w, p, ph, cyc = pj.fpc_step(mg, c["rho"], c["p_old"], c["p_old2"], c["u"], c["v"], c["w"],
                                      float(c["dt"]), float(c["dt_old"]), float(c["dt_old2"]), float(c["rho12"]),
                                      float(c["tol"]))
    assert cyc >= 1
    np.testing.assert_array_equal(ph, c["phat_out"])
    for got, key in [(p, "p_out"), (u, "u_out"), (v, "v_out"), (w, "w_out")]:
        assert rel(got, c[key]) < 1e-10, (key, rel(got, c[key]))


def _field(ndim, n, seed):
    """Padded density (1000:1 blob) and wall-consistent random face velocities."""
    rng = np.random.default_rng(seed)
    shp = padded_shape(ndim, *n)
    idx = np.indices(shp).astype(float)
    x = (idx[0] - 2.5) / n[0]
    y = (idx[1] - 2.5) / n[1]
    z = (idx[2] - 2.5) / n[2] if ndim == 3 else 0.5 * np.ones(shp)
    phi = 0.25 - np.sqrt((x - 0.45) ** 2 + (y - 0.55) ** 2 + (z - 0.5) ** 2 * (ndim == 3))
    rho = W.density(phi, 1.5 / n[0], 1000.0).ravel()
    u, v, w = (rng.standard_normal(shp).ravel() for _ in range(3))
    return rho, u, v, w


@pytest.mark.parametrize("ndim,n", [(2, (64, 96, 1)), (3, (24, 24, 32))], ids=["2d", "3d"])
def test_fpc_step_matches_oracle(oracle, ndim, n):
    dx, dy, dz = (1.0 / n[0],) * 3
    rho, u, v, w = _field(ndim, n, 5)
    types = [3, 3, 4, 4, 3, 3]
    zero = np.zeros(6)
    for stag, a in enumerate((u, v, w)):  # walls consistent with the BCs: compatible Neumann problem
        oracle.lib.orc_bc_apply(ndim, *n, dx, dy, dz, np.array(types, np.int32), zero, stag, a)
    rng = np.random.default_rng(6)
    p_old, p_old2 = rng.standard_normal(u.size), rng.standard_normal(u.size)
    args = (0.01, 0.009, 0.011, 0.001, 1e-9)
    mg = Multigrid(ndim, *n, dx, dy, dz, "neumann")
    pj = Projection(ndim, *n, dx, dy, dz, types)
    print(f"fpc step {ndim}d {n}", flush=True)
    gu, gv, gw, gp, gph, cyc = pj.fpc_step(mg, rho, p_old, p_old2, u, v, w, *args)
    ou, ov, ow, op, oph = u.copy(), v.copy(), w.copy(), np.zeros_like(u), np.zeros_like(u)
    ocyc = oracle.lib.orc_fpc_step(ndim, *n, dx, dy, dz, *args, 100, np.array(types, np.int32), rho, p_old, p_old2,
                                   ou, ov, ow, op, oph)
    assert cyc == ocyc
    np.testing.assert_array_equal(gph, oph)
    for got, ref in [(gp, op), (gu, ou), (gv, ov), (gw, ow)]:
        assert rel(got, ref) < 1e-10


@pytest.mark.parametrize("ndim,n", [(2, (64, 96, 1)), (3, (24, 24, 32))], ids=["2d", "3d"])
def test_fpc_step_rb_cycles(oracle, ndim, n):
    """The FPC step with red-black V(2,2) cycles in its loop (cycle="rb"):
    the same absolute stopping rule reached (more cycles than the
    reference's full cycles, which smooth 20 + 10 times on the finest level
    and descend repeatedly, but each far cheaper), and pressure and velocity
    within what two solutions of residual <= ppe_tol may differ by from the
    oracle's reference-cycle step."""
    dx, dy, dz = (1.0 / n[0],) * 3
    rho, u, v, w = _field(ndim, n, 5)
    types = [3, 3, 4, 4, 3, 3]
    zero = np.zeros(6)
    for stag, a in enumerate((u, v, w)):
        oracle.lib.orc_bc_apply(ndim, *n, dx, dy, dz, np.array(types, np.int32), zero, stag, a)
    rng = np.random.default_rng(6)
    p_old, p_old2 = rng.standard_normal(u.size), rng.standard_normal(u.size)
    args = (0.01, 0.009, 0.011, 0.001, 1e-9)
    mg = Multigrid(ndim, *n, dx, dy, dz, "neumann")
    pj = Projection(ndim, *n, dx, dy, dz, types)
    print(f"fpc step rb cycles {ndim}d {n}", flush=True)
    gu, gv, gw, gp, gph, cyc = pj.fpc_step(mg, rho, p_old, p_old2, u, v, w, *args, cycle="rb")
    ou, ov, ow, op, oph = u.copy(), v.copy(), w.copy(), np.zeros_like(u), np.zeros_like(u)
    ocyc = oracle.lib.orc_fpc_step(ndim, *n, dx, dy, dz, *args, 100, np.array(types, np.int32), rho, p_old, p_old2,
                                   ou, ov, ow, op, oph)
    print(f"cycles rb {cyc} reference {ocyc}; p {rel(gp, op):.2e} u {rel(gu, ou):.2e}", flush=True)
    assert 1 <= cyc < 100 and 1 <= ocyc < 100
    np.testing.assert_array_equal(gph, oph)
    for got, ref in [(gp, op), (gu, ou), (gv, ov), (gw, ow)]:
        assert rel(got, ref) < 1e-6


def test_fpc_step_rejects_unknown_cycle():
    """scihpc_fpc_params.cycle other than 0 or 1 is a bad argument."""
    import ctypes as C

    from scihpc_amd import _capi

    n = (16, 16, 1)
    mg = Multigrid(2, *n, 1 / 16, 1 / 16, 1.0, "neumann")
    pj = Projection(2, *n, 1 / 16, 1 / 16, 1.0, [3, 3, 4, 4, 3, 3])
    T = int(np.prod(padded_shape(2, *n)))
    a = [np.ones(T) for _ in range(8)]
    prm = _capi.FpcParams(0.01, 0.01, 0.01, 0.001, 1e-9, 10, 2)
    cyc = C.c_int()
    rc = _capi.lib().scihpc_proj_fpc_step(pj.h, mg.h, C.byref(prm), *a, C.byref(cyc))
    assert rc == _capi.BAD_ARG
    assert b"cycle" in _capi.lib().scihpc_last_error()


def _interior_div(ndim, n, u, v, w, h):
    shp = padded_shape(ndim, *n)
    U, V, Wz = (a.reshape(shp) for a in (u, v, w))
    g = 3
    sl = tuple(slice(g, g + n[a]) if a < ndim else slice(0, 1) for a in range(3))
    d = (U[sl] - np.roll(U, 1, 0)[sl]) / h + (V[sl] - np.roll(V, 1, 1)[sl]) / h
    if ndim == 3:
        d = d + (Wz[sl] - np.roll(Wz, 1, 2)[sl]) / h
    return d


@pytest.mark.parametrize("precision", [PREC_FP64, PREC_MIXED], ids=["fp64", "mixed"])
@pytest.mark.parametrize("ndim,n", [(2, (48, 64, 1)), (3, (12, 12, 16))], ids=["2d", "3d"])
def test_variable_density_projection(oracle, ndim, n, precision):
    h = 1.0 / n[0]
    rho, u, v, w = _field(ndim, n, 7)
    types = [3, 3, 3, 3, 3, 3]
    zero = np.zeros(6)
    for stag, a in enumerate((u, v, w)):
        oracle.lib.orc_bc_apply(ndim, *n, h, h, h, np.array(types, np.int32), zero, stag, a)
    dt = 0.01
    mg = Multigrid(ndim, *n, h, h, h, "neumann")
    pj = Projection(ndim, *n, h, h, h, types)
    print(f"projection {ndim}d {n} precision={precision}", flush=True)
    gu, gv, gw, gp, st = pj.step(mg, rho, u, v, w, dt, rel_tol=1e-10, precision=precision)
    assert st["converged"], st
    # the corrected face velocities are discretely divergence-free
    d0 = _interior_div(ndim, n, u, v, w, h)
    d1 = _interior_div(ndim, n, gu, gv, gw, h)
    assert np.linalg.norm(d1) < 1e-8 * np.linalg.norm(d0)
    # oracle composite: rhs, dense/red-black solve, correction, BCs
    rhs = np.zeros_like(u)
    oracle.lib.orc_projection_rhs(ndim, *n, h, h, h, dt, u, v, w, rhs)
    b = np.zeros(int(np.prod(n)))
    oracle.lib.orc_padded_to_mg(ndim, *n, rhs, b)
    b -= b.mean()
    om = oracle.mg(ndim, *n, h, h, h, "neumann")
    rho_mg = np.zeros_like(b)
    oracle.lib.orc_padded_to_mg(ndim, *n, rho, rho_mg)
    om.set_rho(rho_mg)
    x = om.dense_solve(b)
    x -= x.mean()
    op = np.zeros_like(u)
    oracle.lib.orc_mg_to_padded(ndim, *n, x, op)
    oracle.lib.orc_bc_apply(ndim, *n, h, h, h, np.zeros(6, np.int32), zero, -1, op)
    ou, ov, ow = u.copy(), v.copy(), w.copy()
    oracle.lib.orc_projection_velocity(ndim, *n, h, h, h, dt, rho, op, ou, ov, ow)
    for stag, a in enumerate((ou, ov, ow)):
        oracle.lib.orc_bc_apply(ndim, *n, h, h, h, np.array(types, np.int32), zero, stag, a)
    for got, ref in [(gu, ou), (gv, ov), (gw, ow)]:
        assert rel(got, ref) < 1e-7, rel(got, ref)
