"""ctypes bindings for the CPU checker (test infrastructure only).

``Oracle`` wraps ``oracle/liboracle.so`` (the C restatement in oracle.c);
``Reference`` wraps ``oracle/_ref/libscihpc_ref.so`` (the SciHPC reference
compiled from its own sources plus ref_harness.cpp) when that was built.
Only tests/, ``__graft_entry__.smoke()`` and bench.py's cpu_baseline leg
import this module; the HIP product path never does.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REF_LIB = os.path.join(HERE, "_ref", "libscihpc_ref.so")

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_vp = C.c_void_p
_i, _d, _l = C.c_int, C.c_double, C.c_long

GHC = 3
BC = {"neumann": 0, "dirichlet": 1, "periodic": 2, "no_slip": 3, "slip": 4}


def padded_shape(ndim, nx, ny=1, nz=1):
    n = [nx, ny, nz]
    return tuple(n[a] + 2 * GHC if a < ndim else 1 for a in range(3))


def _vtk_args(scalars, vectors):
    """(nscalar, names, ptrs, nvector, names, ptrs) for the VTK writers; the
    arrays are kept alive on the returned tuple's ctypes objects."""
    keep = []

    def ptrs(arrs):
        arrs = [np.ascontiguousarray(a, np.float64) for a in arrs]
        keep.extend(arrs)
        out = (C.c_void_p * max(1, len(arrs)))(*[a.ctypes.data for a in arrs])
        out._keep = arrs
        return out

    def names(ns):
        out = (C.c_char_p * max(1, len(ns)))(*[x.encode() for x in ns])
        return out

    sv = [a for _, a in scalars]
    vv = [c for _, t in vectors for c in t]
    return (len(scalars), names([nm for nm, _ in scalars]), ptrs(sv),
            len(vectors), names([nm for nm, _ in vectors]), ptrs(vv))


def _sig(lib, name, res, args):
    f = getattr(lib, name)
    f.restype = res
    f.argtypes = args
    return f


class Oracle:
    def __init__(self, path: str = LIB):
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} missing: run `make -C oracle` first")
        L = self.lib = C.CDLL(path)
        _sig(L, "orc_set_threads", _i, [_i])
        _sig(L, "orc_mg_levels", _i, [_i, _i, _i, _i, _i, _ip, _ip, _ip, _ip])
        _sig(L, "orc_mg_create", _vp, [_i, _i, _i, _i, _d, _d, _d, _i])
        _sig(L, "orc_mg_destroy", None, [_vp])
        _sig(L, "orc_mg_nlevels", _i, [_vp])
        _sig(L, "orc_mg_level_shape", None, [_vp, _i, _ip])
        _sig(L, "orc_mg_vec", C.POINTER(C.c_double), [_vp, _i, _i])
        _sig(L, "orc_mg_set_nocompat", None, [_vp, _i, _i])
        _sig(L, "orc_mg_set_rho", None, [_vp, _dp])
        _sig(L, "orc_mg_get_operator", None, [_vp, _i, _dp, _dp, _dp, _dp])
        _sig(L, "orc_mg_apply", None, [_vp, _i, _dp, _dp])
        _sig(L, "orc_mg_residual", _d, [_vp, _i])
        _sig(L, "orc_mg_compat", None, [_vp, _i, _i])
        _sig(L, "orc_mg_relax", None, [_vp, _i, _i])
        _sig(L, "orc_mg_restriction", None, [_vp, _i])
        _sig(L, "orc_mg_prolongation", None, [_vp, _i])
        _sig(L, "orc_mg_coarse_solve", None, [_vp])
        _sig(L, "orc_mg_vcycle_ref", None, [_vp])
        _sig(L, "orc_mg_fullcycle_ref", None, [_vp])
        _sig(L, "orc_mg_smooth_rb", None, [_vp, _i, _i, _i])
        _sig(L, "orc_mg_restrict_rb", None, [_vp, _i])
        _sig(L, "orc_mg_prolong_rb", None, [_vp, _i])
        _sig(L, "orc_mg_coarse_direct", None, [_vp])
        _sig(L, "orc_mg_vcycle_rb", None, [_vp, _i, _i])
        _sig(L, "orc_mg_solve_rb", _i, [_vp, _d, _i, _i, _i, _dp])
        _sig(L, "orc_mg_dense_solve_level0", _i, [_vp, _dp, _dp])
        _sig(L, "orc_mg_solve_fcg", _i, [_vp, _d, _i, _i, _i, _dp])
        _sig(L, "orc_mg_fcg_vec", C.POINTER(C.c_double), [_vp, _i])
        _sig(L, "orc_bc_apply", None, [_i, _i, _i, _i, _d, _d, _d, _ip, _dp, _i, _dp])
        _sig(L, "orc_weno5_derivatives", None,
             [_i, _i, _i, _i, _d, _d, _d, _dp, _dp, _dp, _dp, _dp, _dp, _dp])
        _sig(L, "orc_convection_weno", None,
             [_i, _i, _i, _i, _d, _d, _d, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp])
        _sig(L, "orc_tvd_rk3_weno", None,
             [_i, _i, _i, _i, _d, _d, _d, _d, _ip, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _i])
        _sig(L, "orc_density", None, [_l, _dp, _d, _d, _i, _dp])
        _sig(L, "orc_ppe_coefficients", None,
             [_i, _i, _i, _i, _d, _d, _d, _dp] + [_dp] * 7)
        _sig(L, "orc_projection_rhs", None, [_i, _i, _i, _i, _d, _d, _d, _d, _dp, _dp, _dp, _dp])
        _sig(L, "orc_projection_velocity", None,
             [_i, _i, _i, _i, _d, _d, _d, _d, _dp, _dp, _dp, _dp, _dp])
        _sig(L, "orc_projection_sor", _i,
             [_i, _i, _i, _i, _d, _i, _d, _d] + [_dp] * 10)
        _sig(L, "orc_fpc_coefficients", None,
             [_i, _i, _i, _i, _d, _d, _d, _d, _dp, _dp] + [_dp] * 8)
        _sig(L, "orc_fpc_rhs", None, [_i, _i, _i, _i, _d, _d, _d, _d, _d, _dp, _dp, _dp, _dp, _dp])
        _sig(L, "orc_mg_to_padded", None, [_i, _i, _i, _i, _dp, _dp])
        _sig(L, "orc_padded_to_mg", None, [_i, _i, _i, _i, _dp, _dp])
        _sig(L, "orc_fpc_velocity", None,
             [_i, _i, _i, _i, _d, _d, _d, _d, _d, _dp, _dp, _dp, _dp, _dp, _dp])
        _sig(L, "orc_pressure_guess", None, [_l, _d, _d, _dp, _dp, _dp])
        _sig(L, "orc_fpc_step", _i, [_i, _i, _i, _i, _d, _d, _d, _d, _d, _d, _d, _d, _i, _ip] + [_dp] * 8)
        _sig(L, "orc_momentum_step", _i, [_i, _i, _i, _i, _d, _d, _d, _dp, _ip, _dp] + [_dp] * 10)
        _sig(L, "orc_node_velocity", None, [_i, _i, _i, _i, _d, _d, _d, _ip] + [_dp] * 6)
        _sig(L, "orc_compact_derivatives", None, [_i, _i, _i, _i, _d, _d, _d, _dp, _vp, _vp, _vp, _i] + [_dp] * 6)
        _sig(L, "orc_compact_derivatives_ld", None, [_i, _i, _i, _i, _d, _d, _d, _dp, _vp, _vp, _vp, _i] + [_dp] * 6)
        _sig(L, "orc_tvd_rk3_uccd", None, [_i, _i, _i, _i, _d, _d, _d, _d, _ip, _dp, _dp, _dp, _dp, _dp, _i])
        _sig(L, "orc_tvd_rk3_uccd_ld", None, [_i, _i, _i, _i, _d, _d, _d, _d, _ip, _dp, _dp, _dp, _dp, _dp, _i])
        _sig(L, "orc_lsf_mass", _d, [_i, _i, _i, _i, _d, _d, _d, _i, _dp])
        _sig(L, "orc_mpls_euler", None, [_i, _i, _i, _i, _d, _d, _d, _d, _d, _d, _i, _d, _ip, _dp, _dp, _i, _vp])
        _sig(L, "orc_find_sign", None, [_l, _dp, _d, _dp])
        _sig(L, "orc_redistance_rk3", None, [_i, _i, _i, _i, _d, _d, _d, _d, _d, _dp, _dp, _ip, _dp, _dp, _i])
        _sig(L, "orc_stabilize_gradient", _d, [_i, _i, _i, _i, _d, _d, _d, _dp, _dp, _dp])
        _sig(L, "orc_vtk_write", _i, [C.c_char_p, _i, _i, _i, _i, _dp, _dp, _i, C.c_void_p, C.c_void_p,
                                      _i, C.c_void_p, C.c_void_p])

    # -------------------------------------------------------------- MG
    def levels(self, ndim, nx, ny, nz=1):
        a = [np.zeros(32, np.int32) for _ in range(4)]
        L = self.lib.orc_mg_levels(ndim, nx, ny, nz, 32, *a)
        return [(int(a[0][i]), int(a[1][i]), int(a[2][i]), int(a[3][i])) for i in range(L)]

    def mg(self, ndim, nx, ny, nz=1, dx=1.0, dy=1.0, dz=1.0, bc="neumann"):
        return OracleMG(self, ndim, nx, ny, nz, dx, dy, dz, BC[bc])

    # ----------------------------------------------------- padded fields
    def bc_apply(self, f, ndim, n, h, types, values=None, stag=-1):
        t = np.ascontiguousarray([BC[x] if isinstance(x, str) else x for x in types], np.int32)
        v = np.zeros(6) if values is None else np.ascontiguousarray(values, np.float64)
        t = np.pad(t, (0, 6 - t.size))
        v = np.pad(v, (0, 6 - v.size))
        self.lib.orc_bc_apply(ndim, *n, *h, t, v, stag, f)

    def weno5(self, ndim, n, h, flux, u, v, w):
        fx, fy, fz = (np.zeros_like(flux) for _ in range(3))
        self.lib.orc_weno5_derivatives(ndim, *n, *h, flux, u, v, w, fx, fy, fz)
        return fx, fy, fz

    def tvd_rk3(self, ndim, n, h, dt, types, values, phi, u, v, w, nsteps=1):
        t = np.pad(np.ascontiguousarray([BC[x] if isinstance(x, str) else x for x in types], np.int32), (0, 0))
        t = np.pad(t, (0, 6 - t.size)).astype(np.int32)
        vals = np.pad(np.ascontiguousarray(values, np.float64), (0, 6 - len(values)))
        fx, fy, fz = (np.zeros_like(phi) for _ in range(3))
        self.lib.orc_tvd_rk3_weno(ndim, *n, *h, dt, t, vals, phi, u, v, w, fx, fy, fz, nsteps)
        return phi

    def vtk_write(self, path, ndim, n, origin, spacing, scalars=(), vectors=()):
        """Legacy VTK file of padded fields, as vtkWriter writes it."""
        args = _vtk_args(scalars, vectors)
        n = list(n) + [1] * (3 - len(n))
        rc = self.lib.orc_vtk_write(path.encode(), ndim, *n, np.ascontiguousarray(origin, np.float64),
                                    np.ascontiguousarray(spacing, np.float64), *args)
        if rc != 0:
            raise OSError(f"orc_vtk_write failed for {path}")

    def density(self, phi, ls_width, ratio, positive_ref=True):
        rho = np.zeros_like(phi)
        self.lib.orc_density(phi.size, phi, ls_width, ratio, int(positive_ref), rho)
        return rho


class OracleMG:
    def __init__(self, o: Oracle, ndim, nx, ny, nz, dx, dy, dz, bc):
        self.o, self.lib = o, o.lib
        self.ndim, self.bc = ndim, bc
        self.h = self.lib.orc_mg_create(ndim, nx, ny, nz, dx, dy, dz, bc)
        self.L = self.lib.orc_mg_nlevels(self.h)
        self.shapes = []
        for l in range(self.L):
            s = np.zeros(5, np.int32)
            self.lib.orc_mg_level_shape(self.h, l, s)
            self.shapes.append(tuple(int(x) for x in s))

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.orc_mg_destroy(self.h)
            self.h = None

    def vec(self, level, which):
        n = self.shapes[level][4]
        p = self.lib.orc_mg_vec(self.h, level, {"rhs": 0, "sol": 1, "res": 2}[which])
        return np.ctypeslib.as_array(p, shape=(n,))

    def set_rho(self, rho):
        self.lib.orc_mg_set_rho(self.h, np.ascontiguousarray(rho, np.float64).ravel())

    def operator(self, level):
        n = self.shapes[level][4]
        a = [np.zeros(n) for _ in range(4)]
        self.lib.orc_mg_get_operator(self.h, level, *a)
        return a

    def apply(self, level, x):
        y = np.zeros(self.shapes[level][4])
        self.lib.orc_mg_apply(self.h, level, np.ascontiguousarray(x, np.float64), y)
        return y

    def residual(self, level=0):
        return self.lib.orc_mg_residual(self.h, level)

    def solve_rb(self, rel_tol, max_cycles=100, nu1=2, nu2=2):
        hist = np.zeros(max_cycles + 1)
        c = self.lib.orc_mg_solve_rb(self.h, rel_tol, max_cycles, nu1, nu2, hist)
        return c, hist[:c]

    def solve_fcg(self, rel_tol, max_iter=100, nu1=2, nu2=2):
        """FCG(1) around the red-black V-cycle from x = 0 (orc_mg_solve_fcg);
        returns (iterations, relative residual after each)."""
        hist = np.zeros(max_iter + 1)
        c = self.lib.orc_mg_solve_fcg(self.h, rel_tol, max_iter, nu1, nu2, hist)
        return c, hist[:c]

    def fcg_vec(self, which):
        """x, r, p, q, p_old, q_old or z of the last orc_mg_solve_fcg (a copy)."""
        k = ["x", "r", "p", "q", "p_old", "q_old", "z"].index(which)
        n = self.shapes[0][4]
        return np.ctypeslib.as_array(self.lib.orc_mg_fcg_vec(self.h, k), shape=(n,)).copy()

    def dense_solve(self, b):
        x = np.zeros(self.shapes[0][4])
        rc = self.lib.orc_mg_dense_solve_level0(self.h, np.ascontiguousarray(b, np.float64), x)
        if rc != 0:
            raise RuntimeError("dense solve failed")
        return x


class Reference:
    """The compiled SciHPC reference (only where oracle/_ref was built)."""

    def __init__(self, path: str = REF_LIB):
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        L = self.lib = C.CDLL(path)
        _sig(L, "ref_mg_create", _vp, [_i, _i, _i, _i, _dp, _i])
        _sig(L, "ref_mg_nlevels", _i, [_vp])
        _sig(L, "ref_mg_level_shape", None, [_vp, _i, _ip])
        _sig(L, "ref_mg_set", None, [_vp, _i, _i, _dp])
        _sig(L, "ref_mg_get", None, [_vp, _i, _i, _dp])
        _sig(L, "ref_mg_set_nocompat", None, [_vp, _i, _i])
        _sig(L, "ref_mg_apply", None, [_vp, _i, _dp, _dp])
        _sig(L, "ref_mg_residual", _d, [_vp, _i])
        _sig(L, "ref_mg_relax", None, [_vp, _i, _i])
        _sig(L, "ref_mg_restriction", None, [_vp, _i])
        _sig(L, "ref_mg_prolongation", None, [_vp, _i])
        _sig(L, "ref_mg_solve", None, [_vp])
        _sig(L, "ref_mg_vcycle", None, [_vp])
        _sig(L, "ref_mg_fullcycle", None, [_vp])
        _sig(L, "ref_bc_apply", None, [_i, _i, _i, _i, _dp, _ip, _dp, _i, _dp])
        _sig(L, "ref_weno_derivatives", None, [_i, _i, _i, _i, _dp] + [_dp] * 7)
        _sig(L, "ref_tvd_rk3_weno", None,
             [_i, _i, _i, _i, _dp, _d, _ip, _dp, _dp, _dp, _dp, _dp, _i])
        _sig(L, "ref_density", None, [_i, _i, _i, _i, _dp, _dp, _d, _d, _i, _dp])
        _sig(L, "ref_projection", None,
             [_i, _i, _i, _i, _dp, _d, _i, _d, _d, _dp, _dp, _dp, _dp, _dp, _dp])
        _sig(L, "ref_fpc_coefficients", None, [_i, _i, _i, _i, _dp, _d, _dp, _dp, _dp])
        _sig(L, "ref_fpc_step", _i, [_i, _i, _i, _i, _dp, _d, _d, _d, _d, _d, _ip] + [_dp] * 8)
        _sig(L, "ref_momentum_step", _i, [_i, _i, _i, _i, _dp, _dp, _ip, _dp] + [_dp] * 10)
        _sig(L, "ref_node_velocity", _i, [_i, _i, _i, _i, _dp, _ip] + [_dp] * 6)
        _sig(L, "ref_compact_derivatives", None, [_i, _i, _i, _i, _dp, _dp, _vp, _vp, _vp, _i] + [_dp] * 6)
        _sig(L, "ref_tvd_rk3_uccd", None, [_i, _i, _i, _i, _dp, _d, _ip, _dp, _dp, _dp, _dp, _dp, _i])
        _sig(L, "ref_lsf_mass", _d, [_i, _i, _i, _i, _dp, _d, _d, _i, _dp])
        _sig(L, "ref_mpls_euler", None, [_i, _i, _i, _i, _dp, _d, _d, _d, _i, _d, _ip, _dp, _dp, _i, _vp])
        _sig(L, "ref_find_sign", None, [_i, _i, _i, _i, _dp, _d, _dp, _dp])
        _sig(L, "ref_redistance_rk3", None, [_i, _i, _i, _i, _dp, _d, _d, _dp, _dp, _ip, _dp, _dp, _i])
        _sig(L, "ref_stabilize_gradient", None, [_i, _i, _i, _i, _dp, _dp, _dp, _dp])
        _sig(L, "ref_vtk_write", _i, [C.c_char_p, C.c_uint, _i, _i, _i, _i, _dp, _i, C.c_void_p, C.c_void_p,
                                      _i, C.c_void_p, C.c_void_p])

    def vtk_write(self, prefix, id, ndim, n, ext, scalars=(), vectors=()):
        """vtkWriter(grid, prefix): create(id), add_scalar per (name, padded),
        add_vector per (name, (u, v, w)), close().  Returns the file path."""
        args = _vtk_args(scalars, vectors)
        n = list(n) + [1] * (3 - len(n))
        self.lib.ref_vtk_write(prefix.encode(), id, ndim, *n, np.ascontiguousarray(ext, np.float64), *args)
        return f"{prefix}{id}.vtk"

    def mg(self, ndim, nx, ny, nz, ext, bc="neumann"):
        return RefMG(self, ndim, nx, ny, nz, ext, BC[bc])


class RefMG:
    def __init__(self, r: Reference, ndim, nx, ny, nz, ext, bc):
        self.lib = r.lib
        self.h = self.lib.ref_mg_create(ndim, nx, ny, nz, np.ascontiguousarray(ext, np.float64), bc)
        self.L = self.lib.ref_mg_nlevels(self.h)
        self.shapes = []
        for l in range(self.L):
            s = np.zeros(5, np.int32)
            self.lib.ref_mg_level_shape(self.h, l, s)
            self.shapes.append(tuple(int(x) for x in s))

    def get(self, level, which):
        out = np.zeros(self.shapes[level][4])
        self.lib.ref_mg_get(self.h, level, {"rhs": 0, "sol": 1, "res": 2, "buffer": 3}[which], out)
        return out

    def set(self, level, which, x):
        self.lib.ref_mg_set(self.h, level, {"rhs": 0, "sol": 1, "res": 2}[which],
                            np.ascontiguousarray(x, np.float64))
