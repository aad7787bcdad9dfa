"""GPU interface maintenance (scihpc_ls_* through the C ABI) against the
reference's fixtures (tests/golden/interface.npz) and the oracle.

Bars: CCD/UCCD derivatives and the UCCD RK3 are pure arithmetic with the
reference's rounding (no FMA): bit-identical with the sequential line
solver (the default).  The partitioned one (segment ends exchanged through
2x2 propagators, then every segment swept again from its true incoming
state) rounds the carried states differently; it must be as accurate as the
reference: its distance to the same computation with the twin sweeps in
long double (the oracle's _ld functions) at most 4 times the reference's
own distance, plus 4 ulp of the largest value; lsf_mass within
1e-14 relative (the GPU reduction order differs from the reference's serial
loop); mpls within 1e-10 of max|phi| (Heaviside/delta use the device's
sin/cos, and eta divides the mass gap by a sum); sign within 1 ulp-level
1e-15; redistancing within 1e-12 of max|phi| (the reference's pow(x, 2) is
glibc's, the device squares).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from golden_util import load, spacing
from scihpc_amd import LevelSet

IFACE = load("interface")


def handle(c):
    ndim, n = int(c["ndim"]), [int(x) for x in c["n"]]
    h = spacing(c)
    ls = LevelSet(ndim, *n, *h, [int(t) for t in c["types"]])
    ls.set_phi(c["phi"])
    ls.set_velocity(c["u"], c["v"], c["w"])
    return ls, ndim


def rel_max(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def check_lines(got, want, extended, solver, msg=""):
    """sequential: bit-identical to the reference's values `want`;
    partitioned: no farther from the long-double values `extended` than 4x
    the reference's own distance + 4 ulp of the largest value"""
    if solver == "sequential":
        np.testing.assert_array_equal(got, want, err_msg=msg)
        return
    ref_err = np.max(np.abs(want - extended))
    err = np.max(np.abs(got - extended))
    bar = 4.0 * ref_err + 4.0 * np.finfo(float).eps * np.max(np.abs(extended))
    assert err <= bar, (msg, err, ref_err)


LINE_MODES = pytest.mark.parametrize("solver", ["sequential", "partitioned"])


@LINE_MODES
@pytest.mark.parametrize("case", sorted(IFACE), ids=str)
def test_compact_derivatives_match_reference(oracle, case, solver):
    c = IFACE[case]
    print(f"compact derivatives {case} {solver}", flush=True)
    ls, ndim = handle(c)
    ls.set_line_solver(solver)
    n, h = [int(x) for x in c["n"]], spacing(c)
    for up, tag in ((False, "ccd"), (True, "uccd")):
        got = ls.compact_derivatives(c["phi"], upwind=up)
        ext = [np.zeros(c["phi"].size) for _ in range(6)]
        oracle.lib.orc_compact_derivatives_ld(ndim, *n, *h, c["phi"], *[c[k].ctypes.data for k in "uvw"], int(up),
                                              *ext)
        for q, k in enumerate(("fx", "fy", "fz", "fxx", "fyy", "fzz")):
            if q % 3 >= ndim:
                continue
            check_lines(got[q].ravel(), c[f"{tag}_{k}"], ext[q], solver, f"{tag} {k}")


@LINE_MODES
@pytest.mark.parametrize("case", sorted(IFACE), ids=str)
def test_uccd_rk3_matches_reference(oracle, case, solver):
    c = IFACE[case]
    print(f"uccd rk3 {case} {solver}", flush=True)
    ls, ndim = handle(c)
    ls.set_line_solver(solver)
    ls.set_scheme("uccd")
    ls.tvd_rk3(float(c["dt"]), 2)
    ext = c["phi"].copy()
    oracle.lib.orc_tvd_rk3_uccd_ld(ndim, *[int(x) for x in c["n"]], *spacing(c), float(c["dt"]),
                                   np.ascontiguousarray(c["types"], np.int32), np.zeros(6), ext, c["u"], c["v"],
                                   c["w"], 2)
    check_lines(ls.phi().ravel(), c["rk3_phi"], ext, solver)


@pytest.mark.parametrize("case", sorted(IFACE), ids=str)
def test_mass_and_mpls_match_reference(case):
    c = IFACE[case]
    print(f"mpls {case}", flush=True)
    ls, _ = handle(c)
    lw, ratio, dt = float(c["ls_width"]), float(c["ratio"]), float(c["dt"])
    m = ls.mass(lw, ratio)
    assert abs(m - float(c["mass"])) <= 1e-14 * abs(float(c["mass"]))
    ls.mpls(dt, lw, ratio, float(c["mpls_mass0"]), count=3)
    assert rel_max(ls.phi().ravel(), c["mpls_phi"]) <= 1e-10
    assert rel_max(ls.field("delta").ravel(), c["mpls_delta"]) <= 1e-12
    # the correction moved the mass toward its target
    assert abs(ls.mass(lw, ratio) - float(c["mpls_mass0"])) < abs(m - float(c["mpls_mass0"]))


@pytest.mark.parametrize("case", sorted(IFACE), ids=str)
def test_sign_and_redistance_match_reference(case):
    c = IFACE[case]
    print(f"redistance {case}", flush=True)
    ls, _ = handle(c)
    lw, dt = float(c["ls_width"]), float(c["dt"])
    ls.find_sign(lw)
    assert rel_max(ls.field("sign").ravel(), c["sign"]) <= 1e-15
    ls.set_phi(1.3 * c["phi"])
    ls.set_field("sign", c["sign"])
    ls.set_field("delta", c["mpls_delta"])
    ls.redistance(dt, float(c["redist_rdt"]), 3)
    assert rel_max(ls.phi().ravel(), c["redist_phi"]) <= 1e-12


@pytest.mark.parametrize("case", sorted(IFACE), ids=str)
def test_stabilize_gradient_matches_reference(case):
    """stabilized_upon_gradient: the Godunov gradient with the stored sign
    and phi divided by its largest value, within the redistancing bar (the
    reference squares with glibc's pow)."""
    c = IFACE[case]
    print(f"stabilize gradient {case}", flush=True)
    ls, _ = handle(c)
    ls.set_phi(1.7 * c["phi"])
    ls.set_field("sign", c["sign"])
    grad, mx = ls.stabilize_gradient()
    assert rel_max(grad.ravel(), c["stab_grad"]) <= 1e-12
    assert abs(mx - np.max(c["stab_grad"])) <= 1e-12 * np.max(c["stab_grad"])
    assert rel_max(ls.phi().ravel(), c["stab_phi"]) <= 1e-12


@LINE_MODES
@pytest.mark.parametrize("ndim,n", [(2, (160, 96, 1)), (3, (40, 32, 36))], ids=["2d", "3d"])
def test_interface_maintenance_matches_oracle_at_size(oracle, ndim, n, solver):
    """Larger grids than the fixtures (line solves over many lines, several
    blocks of reductions): UCCD RK3 against the oracle (bit-identical with
    the sequential line solver), mpls and redistancing within their bars."""
    print(f"interface oracle {ndim}d {n} {solver}", flush=True)
    rng = np.random.default_rng(8)
    lens = (1.0, 0.7, 0.8)
    h = [lens[a] / n[a] if a < ndim else 1.0 for a in range(3)]
    shp = tuple(n[a] + 6 if a < ndim else 1 for a in range(3))
    idx = np.indices(shp).astype(float)
    X = [(idx[a] - 3 + 0.5) * h[a] for a in range(3)]
    phi = (0.2 - np.sqrt(sum((X[a] - [0.5, 0.33, 0.4][a]) ** 2 for a in range(ndim)))).ravel()
    T = phi.size
    u, v, w = (0.5 * rng.standard_normal(T) for _ in range(3))
    types = np.zeros(6, np.int32)
    vals = np.zeros(6)
    ls = LevelSet(ndim, *n, *h, [0] * 6)
    ls.set_line_solver(solver)
    ls.set_phi(phi)
    ls.set_velocity(u, v, w)
    ls.set_scheme("uccd")
    dt = 0.2 * min(h[:ndim])
    ls.tvd_rk3(dt, 2)
    want = phi.copy()
    oracle.lib.orc_tvd_rk3_uccd(ndim, *n, *h, dt, types, vals, want, u, v, w, 2)
    ext = phi.copy()
    oracle.lib.orc_tvd_rk3_uccd_ld(ndim, *n, *h, dt, types, vals, ext, u, v, w, 2)
    check_lines(ls.phi().ravel(), want, ext, solver)
    if solver == "partitioned":  # the rest starts from the reference's phi
        ls.set_phi(want)
    lw, ratio = 1.5 * h[0], 1e-3
    dv = float(np.prod(h[:ndim]))
    m0 = oracle.lib.orc_lsf_mass(ndim, *n, dv, lw, ratio, 1, want)
    assert abs(ls.mass(lw, ratio) - m0) <= 1e-13 * abs(m0)
    ls.mpls(dt, lw, ratio, m0 * 1.001, count=2)
    delta = np.zeros(T)
    oracle.lib.orc_mpls_euler(ndim, *n, *h, dt, lw, ratio, 1, m0 * 1.001, types, vals, want, 2, delta.ctypes.data)
    assert rel_max(ls.phi().ravel(), want) <= 1e-10
    sign = np.zeros(T)
    oracle.lib.orc_find_sign(T, want, lw, sign)
    ls.find_sign(lw)
    ls.redistance(dt, 0.25 * h[0], 2)
    oracle.lib.orc_redistance_rk3(ndim, *n, *h, dt, 0.25 * h[0], sign, delta, types, vals, want, 2)
    assert rel_max(ls.phi().ravel(), want) <= 1e-9


@LINE_MODES
@pytest.mark.parametrize("n", [(1100, 40, 1), (1800, 8, 1), (70, 1030, 1), (2300, 6, 1), (500, 20, 1)],
                         ids=["1106-x", "1806-x", "1036-y", "2306-x", "506-x"])
def test_compact_derivatives_long_lines_match_oracle(oracle, n, solver):
    """Long lines: 506 points (16 lines a workgroup, 57 segments of 9 points),
    1036 and 1106 (8 lines, up to 123 segments of 9), 1806 (8 lines, 107
    segments of 17) and 2306 (more than the segments hold: the sequential
    kernel either way), along x and along the contiguous y axis: CCD and
    UCCD derivatives against the oracle."""
    print(f"compact derivatives long lines {n} {solver}", flush=True)
    rng = np.random.default_rng(21)
    h = [1.0 / n[0], 0.6 / n[1], 1.0]
    shp = (n[0] + 6, n[1] + 6, 1)
    idx = np.indices(shp).astype(float)
    X, Y = (idx[0] - 2.5) * h[0], (idx[1] - 2.5) * h[1]
    phi = (np.sin(7.0 * X) * np.cos(5.0 * Y) + 0.1 * rng.standard_normal(shp)).ravel()
    T = phi.size
    u, v = rng.standard_normal(T), rng.standard_normal(T)
    w = np.zeros(T)
    ls = LevelSet(2, *n, *h, [0] * 6)
    ls.set_line_solver(solver)
    ls.set_velocity(u, v, w)
    for up in (False, True):
        got = ls.compact_derivatives(phi, upwind=up)
        want = [np.zeros(T) for _ in range(6)]
        ext = [np.zeros(T) for _ in range(6)]
        vp = [a.ctypes.data for a in (u, v, w)] if up else [None] * 3
        oracle.lib.orc_compact_derivatives(2, *n, *h, phi, *vp, int(up), *want)
        oracle.lib.orc_compact_derivatives_ld(2, *n, *h, phi, *vp, int(up), *ext)
        mode = "sequential" if max(n) + 6 > 2176 else solver  # longer lines sweep sequentially either way
        for q in (0, 1, 3, 4):
            check_lines(got[q].ravel(), want[q], ext[q], mode, f"upwind={up} output {q}")
