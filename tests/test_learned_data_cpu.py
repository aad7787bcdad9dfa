"""CPU checks of the learned guess's data plumbing (no GPU): the 3-D bubble
workload, the held-out dam-break fixture and the bubble configurations."""
import numpy as np

from scihpc_amd import learned as LRN
from scihpc_amd import workloads as W


def test_bubble_3d_workload():
    n = 32
    w = W.bubble_3d(n, center=(0.4, 0.5, 0.6), radius=0.2, seed=5)
    rho = w["rho"].reshape(n, n, n)
    b = w["rhs"].reshape(n, n, n)
    assert rho.min() == 1.0 and rho.max() == 1000.0
    # the sphere's centre cell is light, the corners heavy
    h = 1.0 / n
    k, j, i = int(0.6 / h), int(0.5 / h), int(0.4 / h)
    assert rho[k, j, i] == 1.0 and rho[0, 0, 0] == 1000.0
    # light volume ~ 4/3 pi r^3
    assert abs((rho < 500).mean() - 4 / 3 * np.pi * 0.2 ** 3) < 0.01
    # b is the divergence of a wall-tangent velocity: mean-free (all-Neumann compatibility)
    assert abs(b.mean()) < 1e-12 * np.abs(b).max()
    assert np.abs(b).max() > 0


def test_bubble3d_configs_clear_the_walls():
    for cfg in LRN.bubble3d_configs(50, 123):
        c, r = np.array(cfg["center"]), cfg["radius"]
        assert r == 0.2 and np.all(c - r >= 0.05 - 1e-12) and np.all(c + r <= 0.95 + 1e-12)


def test_dambreak_fixture_problems():
    probs = LRN.dambreak_fixture_problems()
    assert sorted(probs) == [655, 1310]
    for w in probs.values():
        assert w["rhs"].shape == (1024 * 512,) and w["rho"].shape == (1024 * 512,)
        assert w["rho"].min() == 1.0 and 800 < w["rho"].max() < 820  # air 1, water 1000 / 1.226
        assert abs(w["rhs"].mean()) < 1e-9 * np.abs(w["rhs"]).max()
    # the water column sits at the left wall: the bottom-left cell is water, the top-right air
    rho = probs[655]["rho"].reshape(512, 1024)
    assert rho[0, 0] > 800 and rho[-1, -1] == 1.0
