"""GPU interface maintenance (scihpc_ls_* through the C ABI) against the
reference's fixtures (tests/golden/interface.npz) and the oracle.

Bars: CCD/UCCD derivatives and the UCCD RK3 are pure arithmetic with the
reference's rounding (no FMA) and must be bit-identical; lsf_mass within
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


@pytest.mark.parametrize("case", sorted(IFACE), ids=str)
def test_compact_derivatives_match_reference(case):
    c = IFACE[case]
    print(f"compact derivatives {case}", flush=True)
    ls, ndim = handle(c)
    for up, tag in ((False, "ccd"), (True, "uccd")):
        got = ls.compact_derivatives(c["phi"], upwind=up)
        for q, k in enumerate(("fx", "fy", "fz", "fxx", "fyy", "fzz")):
            if q % 3 >= ndim:
                continue
            np.testing.assert_array_equal(got[q].ravel(), c[f"{tag}_{k}"], err_msg=f"{tag} {k}")


@pytest.mark.parametrize("case", sorted(IFACE), ids=str)
def test_uccd_rk3_matches_reference(case):
    c = IFACE[case]
    print(f"uccd rk3 {case}", flush=True)
    ls, _ = handle(c)
    ls.set_scheme("uccd")
    ls.tvd_rk3(float(c["dt"]), 2)
    np.testing.assert_array_equal(ls.phi().ravel(), c["rk3_phi"])


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


@pytest.mark.parametrize("ndim,n", [(2, (160, 96, 1)), (3, (40, 32, 36))], ids=["2d", "3d"])
def test_interface_maintenance_matches_oracle_at_size(oracle, ndim, n):
    """Larger grids than the fixtures (line solves over many lines, several
    blocks of reductions): UCCD RK3 bit-identical to the oracle, mpls and
    redistancing within their bars."""
    print(f"interface oracle {ndim}d {n}", flush=True)
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
    ls.set_phi(phi)
    ls.set_velocity(u, v, w)
    ls.set_scheme("uccd")
    dt = 0.2 * min(h[:ndim])
    ls.tvd_rk3(dt, 2)
    want = phi.copy()
    oracle.lib.orc_tvd_rk3_uccd(ndim, *n, *h, dt, types, vals, want, u, v, w, 2)
    np.testing.assert_array_equal(ls.phi().ravel(), want)
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
