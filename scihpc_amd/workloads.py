"""Seeded synthetic inputs of the benchmark configurations.

All arrays are returned in the multigrid layout (x fastest: numpy shape
(nz, ny, nx) or (ny, nx) flattened C-order), which is the layout of
multigrid_base::of() (multigrid_base.cpp:8-14).

Workloads
  poisson_dirichlet_2d : 2-D constant-coefficient Poisson on the unit square,
      256x256 cells, f = 2 pi^2 sin(pi x) sin(pi y), homogeneous Dirichlet
      (operator of init_DirichletBC, multigrid_base.cpp:293-345).
  bubble_2d            : 2-D variable-density pressure Poisson, 512x1024
      cells in a 1x2 box, circle of radius 0.25 at (0.5, 0.5), rho 1 inside /
      1000 outside through the smoothed Heaviside of width 1.5h (lsm.cpp:8-18,
      wrapper_func.cpp:387-402), b = MAC divergence of a seeded smooth
      wall-tangent velocity (16 Fourier modes per component, amplitude
      N(0,1)/|k|^2), mean removed; all-Neumann walls.
  poisson_3d / plane_3d : 256^3 unit cube; Dirichlet RHS from seeded modes
      k <= 8 (N(0,1)/|k|^2); variable variant with a 1000:1 jump across
      z = 0.5 + 0.05 sin(2 pi x) sin(2 pi y), cosine modes, Neumann walls.
  bubble_3d            : 3-D bubble problem, 128^3 cells in the unit cube,
      sphere of radius 0.2 at a given centre, rho 1 inside / 1000 outside
      (smoothed Heaviside of width 1.5h), b = MAC divergence of a seeded
      smooth wall-tangent velocity (16 modes per component), mean removed;
      all-Neumann walls.
"""
from __future__ import annotations

import math

import numpy as np

SEED = 20261202


def heaviside(phi, w):
    """lsm.cpp:8-18."""
    h = 0.5 * (1.0 + phi / w + np.sin(math.pi * phi / w) / math.pi)
    return np.where(phi < -w, 0.0, np.where(phi > w, 1.0, h))


def density(phi, w, ratio, positive_ref=True):
    """wrapper_func.cpp:387-402."""
    h = heaviside(phi, w)
    return h + (1.0 - h) * ratio if positive_ref else (1.0 - h) + h * ratio


def poisson_dirichlet_2d(n=256):
    h = 1.0 / n
    xc = (np.arange(n) + 0.5) * h
    X, Y = np.meshgrid(xc, xc, indexing="xy")  # shape (ny, nx), x fastest
    # reference operator is A = +lap_h, so A p = -f with f = 2 pi^2 sin sin
    rhs = -2.0 * math.pi ** 2 * np.sin(math.pi * X) * np.sin(math.pi * Y)
    exact = np.sin(math.pi * X) * np.sin(math.pi * Y)
    return dict(ndim=2, shape=(n, n, 1), spacing=(h, h, 1.0), bc="dirichlet",
                rhs=rhs.ravel(), exact=exact.ravel())


def _modes(rng, count, kmax, ndim):
    ks = rng.integers(1, kmax + 1, size=(count, ndim))
    amp = rng.standard_normal(count) / np.sum(ks.astype(float) ** 2, axis=1)
    return ks, amp


def bubble_2d(nx=512, ny=1024, lx=1.0, ly=2.0, ratio=1000.0, seed=SEED + 2, center=(0.5, 0.5),
              radius=0.25):
    rng = np.random.default_rng(seed)
    dx, dy = lx / nx, ly / ny
    h = min(dx, dy)
    xc = (np.arange(nx) + 0.5) * dx
    yc = (np.arange(ny) + 0.5) * dy
    X, Y = np.meshgrid(xc, yc, indexing="xy")
    phi = radius - np.sqrt((X - center[0]) ** 2 + (Y - center[1]) ** 2)
    rho = density(phi, 1.5 * h, ratio)
    # face velocities: u at x-faces i+1/2 (i = -1..nx-1), v at y-faces
    xf = np.arange(nx + 1) * dx
    yf = np.arange(ny + 1) * dy
    ku, au = _modes(rng, 16, 16, 2)
    kv, av = _modes(rng, 16, 16, 2)
    u = np.zeros((ny, nx + 1))
    v = np.zeros((ny + 1, nx))
    for (k1, k2), a in zip(ku, au):
        u += a * np.outer(np.cos(math.pi * k2 * yc / ly), np.sin(math.pi * k1 * xf / lx))
    for (k1, k2), a in zip(kv, av):
        v += a * np.outer(np.sin(math.pi * k2 * yf / ly), np.cos(math.pi * k1 * xc / lx))
    u[:, 0] = u[:, -1] = 0.0
    v[0, :] = v[-1, :] = 0.0
    # projection_method.cpp:556-559 (dt = 1)
    b = (u[:, 1:] - u[:, :-1]) / dx + (v[1:, :] - v[:-1, :]) / dy
    b -= b.mean()
    return dict(ndim=2, shape=(nx, ny, 1), spacing=(dx, dy, 1.0), bc="neumann",
                rho=rho.ravel(), rhs=b.ravel(), phi=phi.ravel())


def _sep_field(coef, sx, sy, sz):
    """f[z,y,x] = sum_abc coef[a,b,c] sx[a,x] sy[b,y] sz[c,z] (separable)."""
    t = np.einsum("abc,cz->abz", coef, sz)
    t = np.einsum("abz,by->azy", t, sy)
    return np.einsum("azy,ax->zyx", t, sx, optimize=True)


def poisson_3d(n=256, kmax=8, seed=SEED + 3, variable=False, ratio=1000.0):
    rng = np.random.default_rng(seed)
    h = 1.0 / n
    xc = (np.arange(n) + 0.5) * h
    k = np.arange(1, kmax + 1)
    kk = k[:, None, None] ** 2 + k[None, :, None] ** 2 + k[None, None, :] ** 2
    coef = rng.standard_normal((kmax, kmax, kmax)) / kk
    trig = np.cos if variable else np.sin
    s = trig(math.pi * k[:, None] * xc[None, :])
    f = _sep_field(coef, s, s, s)
    out = dict(ndim=3, shape=(n, n, n), spacing=(h, h, h))
    if variable:
        f -= f.mean()
        X, Y = xc[None, :], xc[:, None]
        surf = 0.5 + 0.05 * np.sin(2 * math.pi * X) * np.sin(2 * math.pi * Y)  # (y, x)
        phi = xc[:, None, None] - surf[None, :, :]  # (z, y, x): z above surface > 0
        rho = density(phi, 1.5 * h, ratio)  # 1 above, ratio below
        out.update(bc="neumann", rho=rho.ravel(), rhs=f.ravel())
    else:
        out.update(bc="dirichlet", rhs=f.ravel())
    return out


def bubble_3d(n=128, center=(0.5, 0.5, 0.5), radius=0.2, ratio=1000.0, seed=SEED + 4, modes=16, kmax=8):
    rng = np.random.default_rng(seed)
    h = 1.0 / n
    xc = (np.arange(n) + 0.5) * h
    xf = np.arange(n + 1) * h
    Z, Y, X = np.meshgrid(xc, xc, xc, indexing="ij")
    phi = radius - np.sqrt((X - center[0]) ** 2 + (Y - center[1]) ** 2 + (Z - center[2]) ** 2)
    rho = density(phi, 1.5 * h, ratio)
    # face velocity of component a: sin(pi k x_a) along its own axis (zero on
    # the walls it crosses), cos(pi k x_b) along the others
    def comp(axis):
        ks, amp = _modes(rng, modes, kmax, 3)
        f = np.zeros([n + 1 if 2 - ax == axis else n for ax in range(3)])  # (z, y, x) shape
        for (k1, k2, k3), a in zip(ks, amp):
            fx = (np.sin if axis == 0 else np.cos)(math.pi * k1 * (xf if axis == 0 else xc))
            fy = (np.sin if axis == 1 else np.cos)(math.pi * k2 * (xf if axis == 1 else xc))
            fz = (np.sin if axis == 2 else np.cos)(math.pi * k3 * (xf if axis == 2 else xc))
            f += a * fz[:, None, None] * fy[None, :, None] * fx[None, None, :]
        return f
    u, v, w = comp(0), comp(1), comp(2)
    b = (u[:, :, 1:] - u[:, :, :-1] + v[:, 1:, :] - v[:, :-1, :] + w[1:, :, :] - w[:-1, :, :]) / h
    b -= b.mean()
    return dict(ndim=3, shape=(n, n, n), spacing=(h, h, h), bc="neumann", rho=rho.ravel(), rhs=b.ravel(),
                phi=phi.ravel())
