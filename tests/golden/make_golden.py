"""Write the golden fixtures of tests/golden/*.npz from the compiled SciHPC
reference (oracle/_ref/libscihpc_ref.so, built by `make -C oracle ref` from
the reference's own sources).  The reference is built without OpenMP, so
its reductions run in one fixed order.  Fixtures hold seeded inputs and the
reference's outputs only.

    python tests/golden/make_golden.py
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle.pyoracle import BC, Reference, padded_shape  # noqa: E402


def ext_for(ndim, n, lengths):
    e = []
    for a in range(3):
        e += [0.0, lengths[a] if a < ndim else 0.0]
    return np.array(e)


def mg_cases(ref):
    out = {}
    specs = [
        ("mg2d_dir", 2, (40, 40, 1), (1.0, 1.0, 0), "dirichlet", "v", 3),
        ("mg2d_neu", 2, (24, 48, 1), (1.0, 2.0, 0), "neumann", "v", 3),
        ("mg2d_neu3", 2, (27, 27, 1), (1.0, 1.0, 0), "neumann", "v", 2),
        ("mg3d_neu", 3, (12, 12, 12), (1.0, 1.0, 1.0), "neumann", "f", 2),
        ("mg3d_dir", 3, (8, 8, 16), (1.0, 1.0, 2.0), "dirichlet", "v", 2),
    ]
    for name, ndim, n, lens, bc, kind, cycles in specs:
        rng = np.random.default_rng(sum(map(ord, name)))
        ext = ext_for(ndim, n, lens)
        m = ref.mg(ndim, *n, ext, bc)
        N = m.shapes[0][4]
        b = rng.standard_normal(N)
        if bc == "neumann":
            b -= b.mean()
        m.set(0, "rhs", b)
        m.set(0, "sol", np.zeros(N))
        sols, res = [], []
        for _ in range(cycles):
            (ref.lib.ref_mg_vcycle if kind == "v" else ref.lib.ref_mg_fullcycle)(m.h)
            res.append(ref.lib.ref_mg_residual(m.h, 0))
            sols.append(m.get(0, "sol"))
        out[name] = dict(ndim=ndim, n=np.array(n), ext=ext, bc=BC[bc], kind=kind, rhs=b,
                         sols=np.array(sols), res=np.array(res),
                         shapes=np.array(m.shapes, dtype=np.int64))
        # single pieces on a fresh hierarchy: residual, relax(3), restriction,
        # prolongation, coarse solve
        m2 = ref.mg(ndim, *n, ext, bc)
        x0 = rng.standard_normal(N)
        m2.set(0, "rhs", b)
        m2.set(0, "sol", x0)
        r0 = ref.lib.ref_mg_residual(m2.h, 0)
        res0 = m2.get(0, "res")
        ref.lib.ref_mg_relax(m2.h, 0, 3)
        relax3 = m2.get(0, "sol")
        ref.lib.ref_mg_restriction(m2.h, 0)
        coarse_rhs = m2.get(1, "rhs")
        Nc = m2.shapes[1][4]
        xc = rng.standard_normal(Nc)
        m2.set(1, "sol", xc)
        m2.set(0, "sol", relax3)
        ref.lib.ref_mg_prolongation(m2.h, 1)
        prolonged = m2.get(0, "sol")
        out[name].update(x0=x0, r0=r0, res0=res0, relax3=relax3, coarse_rhs=coarse_rhs, xc=xc,
                         prolonged=prolonged)
    return out


def bc_cases(ref):
    out = {}
    rng = np.random.default_rng(7)
    for ndim, n in [(2, (6, 5, 1)), (3, (4, 4, 5)), (1, (7, 1, 1))]:
        shp = padded_shape(ndim, *n)
        ext = ext_for(ndim, n, (1.0, 2.0, 3.0))
        for t in range(5):
            for stag in [-1] + list(range(ndim)):
                f = rng.standard_normal(int(np.prod(shp)))
                types = np.full(6, t, np.int32)
                vals = rng.standard_normal(6)
                g = f.copy()
                ref.lib.ref_bc_apply(ndim, *n, ext, types, vals, stag, g)
                out[f"bc{ndim}d_t{t}_s{stag}"] = dict(ndim=ndim, n=np.array(n), ext=ext, types=types,
                                                     vals=vals, stag=stag, f=f, g=g)
        # mixed types (no periodic mixing)
        f = rng.standard_normal(int(np.prod(shp)))
        types = np.array([0, 1, 3, 4, 1, 0], np.int32)
        vals = rng.standard_normal(6)
        g = f.copy()
        ref.lib.ref_bc_apply(ndim, *n, ext, types, vals, -1, g)
        out[f"bc{ndim}d_mixed"] = dict(ndim=ndim, n=np.array(n), ext=ext, types=types, vals=vals,
                                       stag=-1, f=f, g=g)
    return out


def smooth_field(shp, rng, ndim):
    grids = np.meshgrid(*[np.linspace(0, 1, s) for s in shp], indexing="ij")
    f = np.zeros(shp)
    for _ in range(4):
        k = rng.uniform(1, 4, 3)
        ph = rng.uniform(0, 6.28, 3)
        f += rng.standard_normal() * np.sin(k[0] * 6.28 * grids[0] + ph[0]) * np.cos(
            k[1] * 6.28 * grids[1] + ph[1]) * np.cos(k[2] * 6.28 * grids[2] + ph[2])
    return f


def weno_cases(ref):
    out = {}
    rng = np.random.default_rng(11)
    for ndim, n, lens in [(1, (40, 1, 1), (1.0, 1, 1)), (2, (24, 20, 1), (1.0, 2.0, 1)),
                          (3, (10, 12, 8), (1.0, 1.0, 1.0))]:
        shp = padded_shape(ndim, *n)
        ext = ext_for(ndim, n, lens)
        for kind in ["smooth", "step"]:
            f = smooth_field(shp, rng, ndim)
            if kind == "step":
                f = np.where(f > 0.1, 1.0, -0.5) + 0.01 * f
            u, v, w = (smooth_field(shp, rng, ndim) for _ in range(3))
            fx, fy, fz = (np.zeros(int(np.prod(shp))) for _ in range(3))
            args = [np.ascontiguousarray(a.ravel()) for a in (f, u, v, w)]
            ref.lib.ref_weno_derivatives(ndim, *n, ext, *args, fx, fy, fz)
            out[f"weno{ndim}d_{kind}"] = dict(ndim=ndim, n=np.array(n), ext=ext, f=args[0], u=args[1],
                                              v=args[2], w=args[3], fx=fx, fy=fy, fz=fz)
    return out


def rk3_cases(ref):
    out = {}
    rng = np.random.default_rng(13)
    for ndim, n, lens, types in [(2, (32, 24, 1), (1.0, 1.0, 1), [0, 0, 0, 0, 0, 0]),
                                 (2, (20, 30, 1), (2.0, 3.0, 1), [2, 2, 2, 2, 0, 0]),
                                 (1, (50, 1, 1), (2.0, 1, 1), [2, 2, 0, 0, 0, 0]),
                                 (3, (12, 10, 8), (1.0, 1.0, 1.0), [0, 0, 3, 3, 1, 1])]:
        shp = padded_shape(ndim, *n)
        ext = ext_for(ndim, n, lens)
        grids = np.meshgrid(*[(np.arange(s) - 2.5) / max(n[0], 1) for s in shp], indexing="ij")
        phi = 0.3 - np.sqrt(sum((g - 0.5) ** 2 for g in grids[:ndim]))
        u, v, w = (smooth_field(shp, rng, ndim) for _ in range(3))
        h = min(lens[a] / n[a] for a in range(ndim))
        dt = 0.25 * h
        types = np.array(types, np.int32)
        vals = np.zeros(6)
        p = np.ascontiguousarray(phi.ravel())
        ref.lib.ref_bc_apply(ndim, *n, ext, types, vals, -1, p)
        p0 = p.copy()
        uu, vv, ww = (np.ascontiguousarray(a.ravel()) for a in (u, v, w))
        ref.lib.ref_tvd_rk3_weno(ndim, *n, ext, dt, types, vals, p, uu, vv, ww, 3)
        out[f"rk3_{ndim}d_{''.join(map(str, types[:2 * ndim]))}"] = dict(
            ndim=ndim, n=np.array(n), ext=ext, dt=dt, types=types, vals=vals, phi0=p0, u=uu, v=vv, w=ww,
            phi3=p.copy(), nsteps=3)
    return out


def field_cases(ref):
    out = {}
    rng = np.random.default_rng(17)
    for ndim, n in [(2, (16, 12, 1)), (3, (8, 6, 10))]:
        shp = padded_shape(ndim, *n)
        ext = ext_for(ndim, n, (1.0, 1.0, 1.0))
        tot = int(np.prod(shp))
        phi = rng.uniform(-0.2, 0.2, tot)
        rho = np.zeros(tot)
        ref.lib.ref_density(ndim, *n, ext, phi, 0.05, 1000.0, 1, rho)
        rho2 = np.zeros(tot)
        ref.lib.ref_density(ndim, *n, ext, phi, 0.05, 0.2, 0, rho2)
        p = rng.standard_normal(tot)
        u, v, w = (rng.standard_normal(tot) for _ in range(3))
        coef = np.zeros(8 * tot)
        pp, uu, vv, ww = p.copy(), u.copy(), v.copy(), w.copy()
        ref.lib.ref_projection(ndim, *n, ext, 0.01, 7, 1e-30, 1e-30, rho, pp, uu, vv, ww, coef)
        fcoef = np.zeros(8 * tot)
        phat = rng.standard_normal(tot)
        ref.lib.ref_fpc_coefficients(ndim, *n, ext, 0.001, rho, phat, fcoef)
        out[f"fields{ndim}d"] = dict(ndim=ndim, n=np.array(n), ext=ext, phi=phi, rho=rho, rho2=rho2, p=p, u=u,
                                     v=v, w=w, coef=coef, p_out=pp, u_out=uu, v_out=vv, w_out=ww, dt=0.01,
                                     sweeps=7, phat=phat, fcoef=fcoef, rho12=0.001)
    return out


def fpc_cases(ref):
    """One FPC projection step (pressure guess, FPC coefficients, rhs,
    full-cycle multigrid from p = 0, velocity correction, velocity BCs)."""
    out = {}
    rng = np.random.default_rng(23)
    for ndim, n, vt in [(2, (16, 16, 1), (3, 3, 3, 3, 3, 3)), (3, (8, 8, 8), (3, 3, 4, 4, 3, 3))]:
        shp = padded_shape(ndim, *n)
        ext = ext_for(ndim, n, (1.0, 1.0, 1.0))
        tot = int(np.prod(shp))
        phi = rng.uniform(-0.3, 0.3, tot)
        rho = np.zeros(tot)
        ref.lib.ref_density(ndim, *n, ext, phi, 0.1, 1000.0, 1, rho)
        p_old, p_old2 = rng.standard_normal(tot), rng.standard_normal(tot)
        u, v, w = (rng.standard_normal(tot) for _ in range(3))
        types = np.array(vt, np.int32)
        # walls as the velocity wrapper leaves them, so that div u sums to zero
        # and the Neumann pressure problem is compatible
        for stag, arr in enumerate((u, v, w)):
            ref.lib.ref_bc_apply(ndim, *n, ext, types, np.zeros(6), stag, arr)
        dt, dt_old, dt_old2, rho12, tol = 0.01, 0.009, 0.011, 0.001, 1e-8
        uu, vv, ww = u.copy(), v.copy(), w.copy()
        pp, ph = np.zeros(tot), np.zeros(tot)
        ref.lib.ref_fpc_step(ndim, *n, ext, dt, dt_old, dt_old2, rho12, tol, types, rho, p_old, p_old2,
                             uu, vv, ww, pp, ph)
        out[f"fpc{ndim}d"] = dict(ndim=ndim, n=np.array(n), ext=ext, types=types, dt=dt, dt_old=dt_old,
                                  dt_old2=dt_old2, rho12=rho12, tol=tol, rho=rho, p_old=p_old, p_old2=p_old2,
                                  u=u, v=v, w=w, u_out=uu, v_out=vv, w_out=ww, p_out=pp, phat_out=ph)
    return out


def vtk_cases(ref):
    """vtkWriter files (create, two add_scalar, one add_vector, close) of
    seeded padded fields; the fixture keeps the file's bytes."""
    import tempfile
    out = {}
    rng = np.random.default_rng(29)
    for ndim, n, ext in [(1, (9, 1, 1), [0.5, 2.0, 0, 0, 0, 0]),
                         (2, (12, 8, 1), [0.0, 1.0, -0.5, 2.0, 0, 0]),
                         (3, (6, 5, 4), [0.0, 1.0, -0.5, 2.0, 0.25, 3.0])]:
        shp = padded_shape(ndim, *n)
        phi, p, u, v, w = (rng.standard_normal(shp).ravel() for _ in range(5))
        with tempfile.TemporaryDirectory() as d:
            path = ref.vtk_write(os.path.join(d, "snap"), 3, ndim, n, ext, [("phi", phi), ("pressure", p)],
                                 [("nvel", (u, v, w))])
            data = np.frombuffer(open(path, "rb").read(), np.uint8)
        out[f"vtk{ndim}d"] = dict(ndim=ndim, n=np.array(n), ext=np.array(ext), phi=phi, p=p, u=u, v=v, w=w,
                                  file=data)
    return out


def momentum_cases(ref):
    """find_source_sec + find_intermediate_velocity (one momentum predictor
    step) from seeded velocities (walls as the velocity wrapper leaves
    them), node velocities and previous sources, phi = signed distance to a
    circle/sphere; plus node_from_face + apply_nvel_bc.  With surface
    tension the reference leaves NaN in some corner ghost cells (zero level-
    set gradient there gives 0/0 curvature); the fixture keeps them."""
    out = {}
    rng = np.random.default_rng(31)
    lens = (1.0, 2.0, 1.5)
    for tag, ndim, n, types, prm in [
            ("2d_st", 2, (16, 12, 1), (3, 3, 3, 3, 3, 3), (0.01, 100.0, 2.0, 5.0, 1.5 / 16, 1e-3, 0.01, 1)),
            ("2d_slip", 2, (16, 12, 1), (3, 4, 4, 3, 3, 3), (0.01, 100.0, -1.0, -1.0, 1.5 / 16, 1e-3, 0.02, 0)),
            ("3d_st", 3, (8, 7, 6), (3, 3, 3, 3, 3, 3), (0.01, 50.0, 3.0, 5.0, 0.2, 1e-3, 0.01, 1)),
            ("3d_slip", 3, (8, 7, 6), (4, 3, 3, 4, 3, 4), (0.01, 50.0, 3.0, -1.0, 0.2, 0.01, 0.05, 0))]:
        shp = padded_shape(ndim, *n)
        T = int(np.prod(shp))
        ext = ext_for(ndim, n, lens)
        h = [lens[a] / n[a] if a < ndim else 1.0 for a in range(3)]
        idx = np.indices(shp).astype(float)
        ctr = (0.43, 0.71, 0.52)
        d2 = sum(((idx[a] - 3 + 0.5) * h[a] - ctr[a]) ** 2 for a in range(ndim))
        phi = (0.3 - np.sqrt(d2)).ravel()
        u, v, w, nu, nv, nw, su, sv, sw = (rng.standard_normal(T) for _ in range(9))
        if ndim == 2:
            w[:] = 0.0
            sw[:] = 0.0
        types = np.array(types, np.int32)
        for stag, arr in enumerate((u, v, w)):
            ref.lib.ref_bc_apply(ndim, *n, ext, types, np.zeros(6), stag, arr)
        prm = np.array(prm, float)
        outs = [x.copy() for x in (u, v, w, nu, nv, nw, su, sv, sw)]
        rho = np.zeros(T)
        ref.lib.ref_momentum_step(ndim, *n, ext, prm, types, phi, *outs, rho)
        d = dict(ndim=ndim, n=np.array(n), ext=ext, types=types, prm=prm, phi=phi, u=u, v=v, w=w, nu=nu, nv=nv,
                 nw=nw, su=su, sv=sv, sw=sw, rho_out=rho)
        for k, x in zip(("u", "v", "w", "nu", "nv", "nw", "su", "sv", "sw"), outs):
            d[k + "_out"] = x
        # node velocities of the predicted face velocities
        nd = [np.zeros(T) for _ in range(3)]
        ref.lib.ref_node_velocity(ndim, *n, ext, types, outs[0], outs[1], outs[2], *nd)
        for k, x in zip(("node_u", "node_v", "node_w"), nd):
            d[k] = x
        out[f"mom{tag}"] = d
    return out


def interface_cases(ref):
    """Interface maintenance: CCD and UCCD line derivatives, tvd_rk3 with the
    UCCD convection term, lsf_mass, Euler steps with mpls, find_sign and
    tvd_rk3 with lsf_redistance_lambda, from a perturbed circle/sphere
    level set and seeded velocities."""
    out = {}
    rng = np.random.default_rng(37)
    lens = (1.0, 2.0, 1.5)
    for ndim, n, types in [(2, (20, 14, 1), (0, 0, 0, 0, 0, 0)), (3, (9, 8, 7), (0, 0, 0, 0, 0, 0))]:
        shp = padded_shape(ndim, *n)
        T = int(np.prod(shp))
        ext = ext_for(ndim, n, lens)
        h = [lens[a] / n[a] if a < ndim else 1.0 for a in range(3)]
        idx = np.indices(shp).astype(float)
        X = [(idx[a] - 3 + 0.5) * h[a] for a in range(3)]
        ctr = (0.45, 0.8, 0.6)
        r0 = np.sqrt(sum((X[a] - ctr[a]) ** 2 for a in range(ndim)))
        phi = (0.3 - r0 + 0.02 * np.sin(5 * X[0]) * np.cos(3 * X[1])).ravel()
        u, v, w = (rng.standard_normal(T) for _ in range(3))
        types = np.array(types, np.int32)
        vals = np.zeros(6)
        d = dict(ndim=ndim, n=np.array(n), ext=ext, types=types, phi=phi, u=u, v=v, w=w)
        for up in (0, 1):
            outs = [np.zeros(T) for _ in range(6)]
            vp = [x.ctypes.data for x in (u, v, w)]
            ref.lib.ref_compact_derivatives(ndim, *n, ext, phi, *vp, up, *outs)
            tag = "uccd" if up else "ccd"
            for k, x in zip(("fx", "fy", "fz", "fxx", "fyy", "fzz"), outs):
                d[f"{tag}_{k}"] = x
        dt, lw, ratio = 0.004, 1.5 * h[0], 1e-3
        a = phi.copy()
        ref.lib.ref_tvd_rk3_uccd(ndim, *n, ext, dt, types, vals, a, u, v, w, 2)
        d["rk3_phi"] = a
        mass = ref.lib.ref_lsf_mass(ndim, *n, ext, lw, ratio, 1, phi)
        d["mass"] = mass
        a = phi.copy()
        delta = np.zeros(T)
        ref.lib.ref_mpls_euler(ndim, *n, ext, dt, lw, ratio, 1, mass * 1.002, types, vals, a, 3, delta.ctypes.data)
        d.update(mpls_mass0=mass * 1.002, mpls_phi=a, mpls_delta=delta)
        sign = np.zeros(T)
        ref.lib.ref_find_sign(ndim, *n, ext, lw, phi, sign)
        d["sign"] = sign
        a = 1.3 * phi
        ref.lib.ref_redistance_rk3(ndim, *n, ext, dt, 0.25 * h[0], sign, delta, types, vals, a, 3)
        d.update(redist_rdt=0.25 * h[0], redist_phi=a, dt=dt, ls_width=lw, ratio=ratio)
        a = 1.7 * phi
        grad = np.zeros(T)
        ref.lib.ref_stabilize_gradient(ndim, *n, ext, sign, a, grad)
        d.update(stab_phi=a, stab_grad=grad)
        out[f"iface{ndim}d"] = d
    return out


def main():
    ref = Reference()
    groups = {"mg": mg_cases, "bc": bc_cases, "weno": weno_cases, "rk3": rk3_cases, "fields": field_cases,
              "fpc": fpc_cases, "vtk": vtk_cases, "momentum": momentum_cases,
              "interface": interface_cases}
    only = sys.argv[1:]
    if only:
        groups = {k: v for k, v in groups.items() if k in only}
    for gname, fn in groups.items():
        cases = fn(ref)
        flat = {}
        for cname, d in cases.items():
            for k, v in d.items():
                flat[f"{cname}/{k}"] = np.asarray(v)
        path = os.path.join(HERE, f"{gname}.npz")
        np.savez_compressed(path, **flat)
        print(path, os.path.getsize(path), "bytes", len(cases), "cases")


if __name__ == "__main__":
    main()
