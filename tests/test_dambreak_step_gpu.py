"""One full dam-break step at the config-4 shape (1024x512 cells, 4x2 box,
water column level set, rho 1000:1, mu 100:1, gravity, no-slip walls) on
the GPU against the oracle, stage by stage, each stage fed the oracle's
previous output so that every comparison isolates one stage:

  1. level-set advection: tvd_rk3 with the WENO5 convection
     (runge_kutta.cpp:13-49, weno_solver.cpp:16-278, source.cpp:8-40);
  2. momentum predictor: find_source_sec + find_intermediate_velocity
     (projection_method.cpp:314-402);
  3. variable-density projection (projection_method.cpp:544-657) with the
     A(rho) flexible-CG multigrid solve in place of the SOR.

Bars: level set 1e-12 relative L2; momentum 1e-12 of each field's largest
value (device sin/cos/pow may differ from glibc's in the last bit); the
projection with FP64 solves to 1e-10 on both sides: pressure and velocity
within 1e-8 relative L2 (same algorithm and iteration count; what is left
is the FP64 floor of the 1000:1 operator).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from scihpc_amd import PREC_FP64, LevelSet, Multigrid, Projection

NX, NY = 1024, 512
H = 4.0 / NX
TYPES = np.array([3, 3, 3, 3, 3, 3], np.int32)  # no-slip velocity walls


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def maxrel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def test_dambreak_step_1024x512_vs_oracle(oracle):
    shp = (NX + 6, NY + 6, 1)
    T = int(np.prod(shp))
    n = (NX, NY, 1)
    h3 = (H, H, 1.0)
    I, J = np.meshgrid((np.arange(shp[0]) - 2.5) * H, (np.arange(shp[1]) - 2.5) * H, indexing="ij")
    phi = np.minimum(1.0 - I, 2.0 - J).ravel()  # signed distance to the column [0,1]x[0,2]
    oracle.bc_apply(phi, 2, n, h3, ["neumann"] * 4)
    # a collapsing-column-like smooth flow, wall-consistent
    u = (0.4 * np.sin(np.pi * I / 4) * np.cos(np.pi * J / 2)).ravel()
    v = (-0.4 * np.cos(np.pi * I / 4) * np.sin(np.pi * J / 2)).ravel()
    w = np.zeros(T)
    for stag, a in enumerate((u, v, w)):
        oracle.lib.orc_bc_apply(2, *n, *h3, TYPES, np.zeros(6), stag, a)
    nu, nv, nw = (np.zeros(T) for _ in range(3))
    oracle.lib.orc_node_velocity(2, *n, *h3, TYPES, u, v, w, nu, nv, nw)
    dt = 0.25 * H

    # 1. level set
    print("dam-break step: level set", flush=True)
    ls = LevelSet(2, NX, NY, 1, H, H, 1.0, bc_types=["neumann"] * 4)
    ls.set_phi(phi)
    ls.set_velocity(nu, nv)
    ls.tvd_rk3(dt, 1)
    phi_o = phi.copy()
    oracle.tvd_rk3(2, n, h3, dt, ["neumann"] * 4, [0.0] * 4, phi_o, nu, nv, nw, 1)
    e = rel(ls.phi().ravel(), phi_o)
    print(f"  level set rel {e:.2e}", flush=True)
    assert e < 1e-12

    # 2. momentum predictor (air/water: density 1e-3, viscosity 1e-2, Fr, Re)
    print("dam-break step: momentum", flush=True)
    prm8 = np.array([dt, 2000.0, 1.0, -1.0, 1.5 * H, 1e-3, 1e-2, 1.0])
    pj = Projection(2, NX, NY, 1, H, H, 1.0, [int(t) for t in TYPES])
    g = pj.momentum_step(Projection.flow_params(*prm8[:7], positive_ref=True), phi_o, u, v, w, nu, nv, nw,
                         want_rho=True)
    ou, ov, ow, onu, onv, onw = (a.copy() for a in (u, v, w, nu, nv, nw))
    osrc = [np.zeros(T) for _ in range(3)]
    orho = np.zeros(T)
    assert oracle.lib.orc_momentum_step(2, *n, *h3, prm8, TYPES, phi_o, ou, ov, ow, onu, onv, onw, *osrc,
                                        orho) == 0
    for k, got, want in zip(("u", "v", "nu", "nv", "rho"), (g[0], g[1], g[3], g[4], g[6]),
                            (ou, ov, onu, onv, orho)):
        d = maxrel(got, want)
        print(f"  momentum {k}: max diff / max {d:.2e}", flush=True)
        assert np.all(np.isfinite(got)) and d <= 1e-12, k

    # 3. projection on A(rho): FP64 flexible CG to 1e-10 on both sides
    print("dam-break step: projection", flush=True)
    mg = Multigrid(2, NX, NY, 1, H, H, 1.0, "neumann")
    gu, gv, gw, gp, st = pj.step(mg, orho, ou, ov, ow, dt, rel_tol=1e-10, precision=PREC_FP64, accel=1)
    assert st["converged"], st
    rhs = np.zeros(T)
    oracle.lib.orc_projection_rhs(2, *n, *h3, dt, ou, ov, ow, rhs)
    b = np.zeros(NX * NY)
    oracle.lib.orc_padded_to_mg(2, *n, rhs, b)
    b -= b.mean()
    om = oracle.mg(2, NX, NY, 1, H, H, 1.0, "neumann")
    rho_mg = np.zeros(NX * NY)
    oracle.lib.orc_padded_to_mg(2, *n, orho, rho_mg)
    om.set_rho(rho_mg)
    om.vec(0, "rhs")[:] = b
    it, _ = om.solve_fcg(1e-10, 200)
    x = om.vec(0, "sol").copy()
    op = np.zeros(T)
    oracle.lib.orc_mg_to_padded(2, *n, x, op)
    oracle.lib.orc_bc_apply(2, *n, *h3, np.zeros(6, np.int32), np.zeros(6), -1, op)
    pu, pv, pw = ou.copy(), ov.copy(), ow.copy()
    oracle.lib.orc_projection_velocity(2, *n, *h3, dt, orho, op, pu, pv, pw)
    for stag, a in enumerate((pu, pv, pw)):
        oracle.lib.orc_bc_apply(2, *n, *h3, TYPES, np.zeros(6), stag, a)
    print(f"  projection: GPU {st['cycles']} FCG iterations, oracle {it}", flush=True)
    assert st["cycles"] == it
    for k, got, want in (("p", gp, op), ("u", gu, pu), ("v", gv, pv)):
        e = rel(got, want)
        print(f"  projection {k}: rel {e:.2e}", flush=True)
        assert e < 1e-8, (k, e)
