"""MI355X (gfx950) projection-step hot path of SciHPC.

Host-side mirror of the reference's solver classes over the C ABI
(include/scihpc_hip.h): ``Multigrid`` stands for multigrid/multigrid_base
(multigrid.h:13-25, multigrid_base.h:17-55), ``LevelSet`` for weno_solver +
runge_kutta + bc_factory on one padded grid (weno_solver.h:17-48,
runge_kutta.h:12-23, bc_factory.h:37-59).  Every call runs HIP kernels;
nothing here computes on the CPU.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._capi import (BC_CODES, MG_DIRICHLET, MG_NEUMANN, NOT_CONVERGED, OK, PREC_FP64,
                    PREC_MIXED, FlowParams, FpcParams, ScihpcError, SolveOpts, SolveStats, check, lib)

__all__ = ["Multigrid", "LevelSet", "Projection", "UNet", "VtkWriter", "ScihpcError", "PREC_FP64",
           "PREC_MIXED", "padded_shape", "read_vtk"]

GHC = 3


def padded_shape(ndim, nx, ny=1, nz=1):
    n = (nx, ny, nz)
    return tuple(n[a] + 2 * GHC if a < ndim else 1 for a in range(3))


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64).ravel()


class Multigrid:
    """multigrid(scalar_data*, structured_grid*) on the GPU.

    ``bc`` is "neumann" (init_NeumannBC) or "dirichlet" (init_DirichletBC).
    Vectors are numpy arrays in the reference's MG layout (x fastest).
    """

    def __init__(self, ndim, nx, ny, nz=1, dx=1.0, dy=1.0, dz=1.0, bc="neumann"):
        self._lib = lib()
        h = C.c_void_p()
        code = MG_NEUMANN if bc == "neumann" else MG_DIRICHLET
        check(self._lib.scihpc_mg_create(ndim, nx, ny, nz, dx, dy, dz, code, C.byref(h)))
        self.h = h
        self.ndim = ndim
        L = C.c_int()
        check(self._lib.scihpc_mg_num_levels(h, C.byref(L)))
        self.level_num = L.value
        self.shapes = []
        for l in range(self.level_num):
            a, b, c, d = C.c_int(), C.c_int(), C.c_int(), C.c_int()
            check(self._lib.scihpc_mg_level_shape(h, l, C.byref(a), C.byref(b), C.byref(c), C.byref(d)))
            self.shapes.append((a.value, b.value, c.value, d.value))
        self.n = [s[0] * s[1] * s[2] for s in self.shapes]

    def close(self):
        if getattr(self, "h", None):
            self._lib.scihpc_mg_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ------------------------------------------------------------ data
    def set_density(self, rho):
        check(self._lib.scihpc_mg_set_density(self.h, _f64(rho)))

    def set_density_padded(self, rho_padded):
        check(self._lib.scihpc_mg_set_density_padded(self.h, _f64(rho_padded)))

    def set_rhs(self, x, level=0):
        check(self._lib.scihpc_mg_set_rhs(self.h, level, _f64(x)))

    def set_sol(self, x, level=0):
        check(self._lib.scihpc_mg_set_sol(self.h, level, _f64(x)))

    def set_rhs_device(self, dptr):
        """Level-0 b from a device array (int address, n FP64 values, MG
        layout, complete before the call): a device-to-device copy."""
        check(self._lib.scihpc_mg_set_rhs_device(self.h, C.c_void_p(int(dptr))))

    def get_sol_device(self, dptr):
        """Level-0 solution into a device array (int address)."""
        check(self._lib.scihpc_mg_get_sol_device(self.h, C.c_void_p(int(dptr))))

    def _get(self, fn, level):
        out = np.empty(self.n[level])
        check(fn(self.h, level, out))
        return out

    def rhs(self, level=0):
        return self._get(self._lib.scihpc_mg_get_rhs, level)

    def sol(self, level=0):
        return self._get(self._lib.scihpc_mg_get_sol, level)

    def res(self, level=0):
        return self._get(self._lib.scihpc_mg_get_res, level)

    def set_nocompat(self, level, flag):
        check(self._lib.scihpc_mg_set_nocompat(self.h, level, int(flag)))

    # ------------------------------------------------------- operators
    def residual(self, level=0):
        r = C.c_double()
        check(self._lib.scihpc_mg_residual(self.h, level, C.byref(r)))
        return r.value

    def apply(self, x, level=0):
        y = np.empty(self.n[level])
        check(self._lib.scihpc_mg_apply(self.h, level, _f64(x), y))
        return y

    def compatibility_condition(self, level, which="sol"):
        check(self._lib.scihpc_mg_compat(self.h, level, 0 if which == "rhs" else 1))

    # ------------------------------------ reference-faithful algorithm
    def relax(self, level, iters):
        check(self._lib.scihpc_mg_relax(self.h, level, iters))

    def restriction(self, level):
        check(self._lib.scihpc_mg_restriction(self.h, level))

    def prolongation(self, level):
        check(self._lib.scihpc_mg_prolongation(self.h, level))

    def coarse_solve(self):
        check(self._lib.scihpc_mg_coarse_solve(self.h))

    def v_cycle_ref(self):
        check(self._lib.scihpc_mg_vcycle_ref(self.h))

    def full_cycle_ref(self):
        check(self._lib.scihpc_mg_fullcycle_ref(self.h))

    # ------------------------------------------------ MI355X algorithm
    def v_cycle(self, nu1=2, nu2=2):
        check(self._lib.scihpc_mg_vcycle(self.h, nu1, nu2))

    def solve(self, rel_tol=1e-8, max_cycles=200, nu1=2, nu2=2, precision=PREC_FP64,
              use_learned=False, zero_guess=True, accel=0, learned_correction=0, lagged_check=0):
        """accel 0: stationary defect correction; 1: flexible CG around the V-cycle.
        use_learned: 1 judged guess with back-off after rejections, 3 judged
        on every solve, 2 forced (ablation).  learned_correction (mixed
        flexible CG): 1 the network's coarse-level correction of the defect
        after the first iteration, judged by the FP64 residual; 2 forced.
        lagged_check 1: the residual is read only from one iteration before
        the count the last converged solve of the same path needed (fewer
        host round trips; a solve may run a few iterations past rel_tol)."""
        o = SolveOpts(rel_tol, max_cycles, nu1, nu2, precision, int(use_learned), int(zero_guess),
                      int(accel), int(learned_correction), int(lagged_check))
        s = SolveStats()
        rc = check(self._lib.scihpc_mg_solve(self.h, C.byref(o), C.byref(s)), allow=(OK, NOT_CONVERGED))
        return {"converged": rc == OK, "cycles": s.cycles, "rel_residual": s.rel_residual,
                "ms": s.ms, "learned_accepted": s.learned_accepted,
                "learned_rejected": s.learned_rejected, "learned_residual": s.learned_residual,
                "learned_skipped": s.learned_skipped, "correction_accepted": s.correction_accepted,
                "correction_rejected": s.correction_rejected,
                "correction_residual": s.correction_residual}

    # ---------------------------------------------- learned guess
    def attach_learned(self, net, level=1):
        """Use UNet `net` (None detaches) for solve(use_learned=True)."""
        self._net = net
        check(self._lib.scihpc_mg_attach_learned(self.h, net.h if net is not None else None, level))

    def learned_features(self, level=1, with_target=False, with_faces=False):
        """(features [cells, 2], target [cells] or None, faces [cells, ndim]
        or None) of the current b, rho and sol on `level` (cells x fastest)."""
        nx, ny, nz = self.shapes[level][:3]
        cells = nx * ny * (nz if self.ndim == 3 else 1)
        feat = np.empty(2 * cells, np.float32)
        tgt = np.empty(cells, np.float32) if with_target else None
        faces = np.empty(self.ndim * cells, np.float32) if with_faces else None
        check(self._lib.scihpc_mg_learned_features(self.h, level, feat, tgt, faces))
        return feat.reshape(cells, 2), tgt, None if faces is None else faces.reshape(cells, self.ndim)

    def probe_kernel(self, which=6, capacity=4096):
        """Arm HIP events around every launch of kernel `which` inside the
        following solves (6: the fused FCG update + FP32 level-0
        pre-smoothing); -1 disarms."""
        check(self._lib.scihpc_mg_probe_kernel(self.h, which, capacity))

    def probe_result(self):
        """(launches, summed ms between the HIP events around them, summed
        ms of the kernels' own wall-clock spans, algorithmic bytes per
        launch) since the probe was armed or last read."""
        n, ms, kms, b = C.c_int(), C.c_double(), C.c_double(), C.c_double()
        check(self._lib.scihpc_mg_probe_result(self.h, C.byref(n), C.byref(ms), C.byref(kms), C.byref(b)))
        return n.value, ms.value, kms.value, b.value

    def time_kernel(self, which, iters=20):
        ms, b = C.c_double(), C.c_double()
        check(self._lib.scihpc_mg_time_kernel(self.h, which, iters, C.byref(ms), C.byref(b)))
        return ms.value, b.value


class UNet:
    """The learned pressure guess: 4-level U-Net (16/32/64/128 channels) with
    BF16 MFMA convolutions; no SciHPC counterpart.  height x width (x depth
    for the 3-D network: 3x3x3 convolutions, 2x2x2 pooling) is the grid of
    the multigrid level it serves (multiples of 8)."""

    N_LAYERS = 15  # 14 3x3 (3x3x3) convolutions + the 1x1 head

    def __init__(self, height, width, max_batch=1, seed=0, depth=None):
        self._lib = lib()
        h = C.c_void_p()
        if depth is None:
            check(self._lib.scihpc_unet_create(height, width, max_batch, seed, C.byref(h)))
        else:
            check(self._lib.scihpc_unet_create3(depth, height, width, max_batch, seed, C.byref(h)))
        self.h = h
        self.ndim = 2 if depth is None else 3
        self.D = 1 if depth is None else depth
        self.H, self.W, self.max_batch = height, width, max_batch
        self.pixels = self.D * height * width
        n = C.c_longlong()
        check(self._lib.scihpc_unet_num_params(h, C.byref(n)))
        self.num_params = n.value
        self.layers = []
        for i in range(self.N_LAYERS):
            a, b, c, o = C.c_int(), C.c_int(), C.c_int(), C.c_longlong()
            check(self._lib.scihpc_unet_layer(h, i, C.byref(a), C.byref(b), C.byref(c), C.byref(o)))
            self.layers.append(dict(cin=a.value, cout=b.value, level=c.value, offset=o.value))

    def close(self):
        if getattr(self, "h", None):
            self._lib.scihpc_unet_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def params(self):
        out = np.empty(self.num_params, np.float32)
        check(self._lib.scihpc_unet_get_params(self.h, out))
        return out

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == self.num_params
        check(self._lib.scihpc_unet_set_params(self.h, p))

    def grads(self):
        out = np.empty(self.num_params, np.float32)
        check(self._lib.scihpc_unet_get_grads(self.h, out))
        return out

    @property
    def out_scale(self):
        s = C.c_float()
        check(self._lib.scihpc_unet_get_out_scale(self.h, C.byref(s)))
        return s.value

    @out_scale.setter
    def out_scale(self, v):
        check(self._lib.scihpc_unet_set_out_scale(self.h, float(v)))

    def set_samples(self, feat, target, faces=None):
        """feat [n, pixels, 2], target [n, pixels], faces [n, pixels, ndim] or
        None (FP32, host; pixels x fastest); faces select the energy-norm
        loss.  Set out_scale first."""
        feat = np.ascontiguousarray(feat, np.float32)
        target = np.ascontiguousarray(target, np.float32)
        n = feat.shape[0]
        assert feat.size == n * self.pixels * 2 and target.size == n * self.pixels
        if faces is not None:
            faces = np.ascontiguousarray(faces, np.float32).ravel()
            assert faces.size == n * self.pixels * self.ndim
        check(self._lib.scihpc_unet_set_samples(self.h, n, feat.ravel(), target.ravel(), faces))

    def set_loss_smoothing(self, sweeps=0, omega=0.8):
        """Energy loss of the error after `sweeps` damped-Jacobi sweeps of the
        network grid's A(rho) (what the multigrid smoother leaves of it)."""
        check(self._lib.scihpc_unet_set_loss_smoothing(self.h, int(sweeps), float(omega)))

    @staticmethod
    def _idx(samples, flips=None):
        samples = np.asarray(samples, np.int32)
        flips = np.zeros_like(samples) if flips is None else np.asarray(flips, np.int32)
        return np.ascontiguousarray(np.stack([samples, flips], 1).ravel(), np.int32)

    def forward(self, samples, flips=None):
        idx = self._idx(samples, flips)
        out = np.empty(len(samples) * self.pixels, np.float32)
        check(self._lib.scihpc_unet_forward(self.h, len(samples), idx, out))
        return out.reshape(len(samples), self.pixels)

    def backward(self):
        loss = C.c_double()
        check(self._lib.scihpc_unet_backward(self.h, C.byref(loss)))
        return loss.value

    def train_step(self, samples, lr, flips=None, want_loss=True):
        idx = self._idx(samples, flips)
        loss = C.c_double()
        check(self._lib.scihpc_unet_train_step(self.h, len(samples), idx, float(lr),
                                               C.byref(loss) if want_loss else None))
        return loss.value if want_loss else None

    def sync(self):
        check(self._lib.scihpc_unet_sync(self.h))

    def time(self, n=1, train=False, iters=10):
        """(device ms, MFMA FLOPs) of one forward pass or training step."""
        ms, fl = C.c_double(), C.c_double()
        check(self._lib.scihpc_unet_time(self.h, n, 1 if train else 0, iters, C.byref(ms), C.byref(fl)))
        return ms.value, fl.value


class LevelSet:
    """weno_solver + runge_kutta + bc_factory for one padded grid, on the GPU."""

    def __init__(self, ndim, nx, ny, nz=1, dx=1.0, dy=1.0, dz=1.0, bc_types=None, bc_values=None):
        self._lib = lib()
        types = [BC_CODES[t] if isinstance(t, str) else int(t) for t in (bc_types or ["neumann"] * 6)]
        types = np.ascontiguousarray((types + [0] * 6)[:6], np.int32)
        vals = np.ascontiguousarray((list(bc_values or []) + [0.0] * 6)[:6], np.float64)
        h = C.c_void_p()
        check(self._lib.scihpc_ls_create(ndim, nx, ny, nz, dx, dy, dz, types, vals, C.byref(h)))
        self.h = h
        self.ndim = ndim
        self.shape = padded_shape(ndim, nx, ny, nz)
        self.size = int(np.prod(self.shape))

    def close(self):
        if getattr(self, "h", None):
            self._lib.scihpc_ls_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_phi(self, phi):
        check(self._lib.scihpc_ls_set_phi(self.h, _f64(phi)))

    def phi(self):
        out = np.empty(self.size)
        check(self._lib.scihpc_ls_get_phi(self.h, out))
        return out.reshape(self.shape)

    def set_velocity(self, u, v, w=None):
        w = np.zeros(self.size) if w is None else w
        check(self._lib.scihpc_ls_set_velocity(self.h, _f64(u), _f64(v), _f64(w)))

    def weno5_derivatives(self, flux):
        fx, fy, fz = (np.zeros(self.size) for _ in range(3))
        check(self._lib.scihpc_ls_weno5_derivatives(self.h, _f64(flux), fx, fy, fz))
        return fx.reshape(self.shape), fy.reshape(self.shape), fz.reshape(self.shape)

    def tvd_rk3(self, dt, nsteps=1):
        check(self._lib.scihpc_ls_tvd_rk3(self.h, dt, nsteps))

    def apply_bc(self, f, stag=-1):
        a = _f64(f).copy()
        check(self._lib.scihpc_ls_apply_bc(self.h, a, stag))
        return a.reshape(self.shape)

    def density(self, ls_width, ratio, positive_ref=True):
        out = np.empty(self.size)
        check(self._lib.scihpc_ls_density(self.h, ls_width, ratio, int(positive_ref), out))
        return out.reshape(self.shape)

    def time_step(self, dt, iters=10):
        ms, b = C.c_double(), C.c_double()
        check(self._lib.scihpc_ls_time_step(self.h, dt, iters, C.byref(ms), C.byref(b)))
        return ms.value, b.value

    # ------------------------------------------- interface maintenance
    def set_scheme(self, scheme):
        """'weno5' (default) or 'uccd' (the reference's convection term)."""
        code = {"weno5": 0, "uccd": 1}[scheme] if isinstance(scheme, str) else int(scheme)
        check(self._lib.scihpc_ls_set_scheme(self.h, code))

    def set_line_solver(self, solver):
        """CCD/UCCD line systems: "sequential" (the default, one sweep per
        line in the reference's order, bit-identical to matrix_solver.cpp's
        twin_bks) or "partitioned" (segments swept in parallel, carried
        states exchanged through 2x2 propagators; within the sequential
        sweep's own rounding error)."""
        code = {"sequential": 0, "partitioned": 1}[solver]
        check(self._lib.scihpc_ls_set_line_solver(self.h, code))

    def compact_derivatives(self, f, upwind=True):
        """UCCD (by the handle's velocity) or CCD derivatives of a padded
        field: (fx, fy, fz, fxx, fyy, fzz)."""
        outs = [np.zeros(self.size) for _ in range(6)]
        check(self._lib.scihpc_ls_compact_derivatives(self.h, _f64(f), int(upwind), *outs))
        return tuple(o.reshape(self.shape) for o in outs)

    def mass(self, ls_width, ratio, positive_ref=True):
        m = C.c_double()
        check(self._lib.scihpc_ls_mass(self.h, ls_width, ratio, int(positive_ref), C.byref(m)))
        return m.value

    def mpls(self, dt, ls_width, ratio, mass0, count=1, positive_ref=True):
        check(self._lib.scihpc_ls_mpls(self.h, dt, ls_width, ratio, int(positive_ref), mass0, count))

    def find_sign(self, ls_width):
        check(self._lib.scihpc_ls_find_sign(self.h, ls_width))

    def find_delta(self, ls_width):
        check(self._lib.scihpc_ls_find_delta(self.h, ls_width))

    def set_field(self, which, a):
        check(self._lib.scihpc_ls_set_field(self.h, {"sign": 0, "delta": 1}[which], _f64(a)))

    def field(self, which):
        out = np.zeros(self.size)
        check(self._lib.scihpc_ls_get_field(self.h, {"sign": 0, "delta": 1}[which], out))
        return out.reshape(self.shape)

    def redistance(self, dt, rdt, nsteps=1):
        check(self._lib.scihpc_ls_redistance(self.h, dt, rdt, nsteps))

    def stabilize_gradient(self):
        """stabilized_upon_gradient (godunov_gradient.cpp:159-181): phi /=
        max Godunov gradient (stored sign); returns (gradient, maximum)."""
        grad = np.zeros(self.size)
        mx = C.c_double()
        check(self._lib.scihpc_ls_stabilize_gradient(self.h, grad, C.byref(mx)))
        return grad.reshape(self.shape), mx.value


class Projection:
    """The projection step on padded fields around a Multigrid of the same
    grid: the FPC step of projection_method (fast_pressure_correction,
    projection_method.cpp:659-753, with pressure_guess_solution :803-815 and
    init_ppe_fpc_coefficients :467-541) and the variable-density projection
    (projection_method.cpp:544-657) with the A(rho) solve."""

    def __init__(self, ndim, nx, ny, nz=1, dx=1.0, dy=1.0, dz=1.0, vel_bc_types=None):
        self._lib = lib()
        types = [BC_CODES[t] if isinstance(t, str) else int(t) for t in (vel_bc_types or ["no_slip"] * 6)]
        types = np.ascontiguousarray((types + [0] * 6)[:6], np.int32)
        h = C.c_void_p()
        check(self._lib.scihpc_proj_create(ndim, nx, ny, nz, dx, dy, dz, types, C.byref(h)))
        self.h = h
        self.shape = padded_shape(ndim, nx, ny, nz)

    def close(self):
        if getattr(self, "h", None):
            self._lib.scihpc_proj_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def fpc_step(self, mg, rho, p_old, p_old2, u, v, w, dt, dt_old, dt_old2, rho12, ppe_tol=1e-4,
                 max_cycles=100, cycle="reference"):
        """Returns (u, v, w, p, p_hat, cycles) as padded arrays.  cycle:
        "reference" (the reference's full cycles) or "rb" (red-black
        V(2,2) cycles, same loop and stopping rule)."""
        u, v, w = (_f64(a).copy() for a in (u, v, w))
        p = np.zeros_like(u)
        ph = np.zeros_like(u)
        prm = FpcParams(dt, dt_old, dt_old2, rho12, ppe_tol, max_cycles, {"reference": 0, "rb": 1}[cycle])
        cyc = C.c_int()
        check(self._lib.scihpc_proj_fpc_step(self.h, mg.h, C.byref(prm), _f64(rho), _f64(p_old), _f64(p_old2),
                                             u, v, w, p, ph, C.byref(cyc)))
        return u, v, w, p, ph, cyc.value

    @staticmethod
    def flow_params(dt, reynolds, froude=-1.0, weber=-1.0, ls_width=None, density_ratio=1e-3,
                    viscosity_ratio=1e-2, positive_ref=True, h=None):
        """problem_parameters of the momentum step (global.h:25-54);
        ls_width defaults to 1.5 h as the drivers set it."""
        if ls_width is None:
            ls_width = 1.5 * h
        return FlowParams(dt, reynolds, froude, weber, ls_width, density_ratio, viscosity_ratio, int(positive_ref))

    def momentum_step(self, prm, phi, u, v, w, nu, nv, nw, want_rho=False):
        """find_source_sec + find_intermediate_velocity (projection_method.cpp:
        314-402).  Returns (u, v, w, nu, nv, nw[, rho]) as new padded arrays;
        the sources stay in the handle for the next step's AB2."""
        out = [_f64(a).copy() for a in (u, v, w, nu, nv, nw)]
        rho = np.zeros_like(out[0]) if want_rho else None
        check(self._lib.scihpc_proj_momentum_step(self.h, C.byref(prm), _f64(phi), *out,
                                                  rho.ctypes.data if rho is not None else None))
        return tuple(out) + ((rho,) if want_rho else ())

    def set_sources(self, su, sv, sw):
        check(self._lib.scihpc_proj_set_sources(self.h, _f64(su), _f64(sv), _f64(sw)))

    def sources(self):
        n = int(np.prod(self.shape))
        s = [np.zeros(n) for _ in range(3)]
        check(self._lib.scihpc_proj_get_sources(self.h, *s))
        return tuple(s)

    def find_sources(self, prm, phi, u, v, w, nu, nv, nw, want_rho=False):
        """find_source_sec alone (projection_method.cpp:314-382): the handle's
        sources rotate, the velocity is not changed.  Returns (nu, nv, nw[,
        rho])."""
        out = [_f64(a).copy() for a in (nu, nv, nw)]
        rho = np.zeros_like(out[0]) if want_rho else None
        check(self._lib.scihpc_proj_find_sources(self.h, C.byref(prm), _f64(phi), _f64(u), _f64(v), _f64(w),
                                                 *out, rho.ctypes.data if rho is not None else None))
        return tuple(out) + ((rho,) if want_rho else ())

    def intermediate_velocity(self, prm, u, v, w):
        """find_intermediate_velocity (projection_method.cpp:384-402): AB2 with
        the handle's sources and the velocity BCs.  Returns (u, v, w)."""
        out = [_f64(a).copy() for a in (u, v, w)]
        check(self._lib.scihpc_proj_intermediate_velocity(self.h, C.byref(prm), *out))
        return tuple(out)

    def node_velocity(self, u, v, w, nu, nv, nw):
        """node_from_face + apply_nvel_bc; returns (nu, nv, nw)."""
        out = [_f64(a).copy() for a in (nu, nv, nw)]
        check(self._lib.scihpc_proj_node_velocity(self.h, _f64(u), _f64(v), _f64(w), *out))
        return tuple(out)

    def time_momentum(self, prm, iters=20):
        ms = C.c_double()
        check(self._lib.scihpc_proj_time_momentum(self.h, C.byref(prm), iters, C.byref(ms)))
        return ms.value

    def step(self, mg, rho, u, v, w, dt, rel_tol=1e-8, precision=PREC_FP64, accel=1, max_cycles=200):
        """Variable-density projection; returns (u, v, w, p, stats)."""
        u, v, w = (_f64(a).copy() for a in (u, v, w))
        p = np.zeros_like(u)
        o = SolveOpts(rel_tol, max_cycles, 2, 2, precision, 0, 1, accel)
        s = SolveStats()
        rc = check(self._lib.scihpc_proj_step(self.h, mg.h, dt, C.byref(o), _f64(rho), u, v, w, p, C.byref(s)),
                   allow=(OK, NOT_CONVERGED))
        return u, v, w, p, {"converged": rc == OK, "cycles": s.cycles, "rel_residual": s.rel_residual}


class VtkWriter:
    """vtkWriter (vtkWriter.h:15-35): legacy VTK snapshots of padded fields,
    byte-identical to the reference's files.  The device reorders and
    byte-swaps each field; a background thread writes the file.

        w = VtkWriter(2, (nx, ny), origin=(0, 0), spacing=(dx, dy), prefix="out/dambreak")
        w.begin(3); w.add_scalar("phi", phi); w.add_vector("nvel", u, v, w_); w.close()
    """

    def __init__(self, ndim, n, origin, spacing, prefix):
        self._lib = lib()
        n = list(n) + [1] * (3 - len(n))
        org = np.zeros(3)
        sp = np.zeros(3)
        org[:ndim] = origin[:ndim]
        sp[:ndim] = spacing[:ndim]
        h = C.c_void_p()
        check(self._lib.scihpc_vtk_create(ndim, n[0], n[1], n[2], org, sp, prefix.encode(), C.byref(h)))
        self.h = h
        self.prefix = prefix
        self.shape = padded_shape(ndim, *n)

    def close_handle(self):
        if getattr(self, "h", None):
            self._lib.scihpc_vtk_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close_handle()
        except Exception:
            pass

    def begin(self, id):
        check(self._lib.scihpc_vtk_begin(self.h, int(id)))
        return f"{self.prefix}{int(id)}.vtk"

    def _host(self, a):
        a = _f64(a)
        if a.size != int(np.prod(self.shape)):
            raise ValueError(f"field has {a.size} values, padded grid has {int(np.prod(self.shape))}")
        return a

    def add_scalar(self, name, field):
        f = self._host(field)
        check(self._lib.scihpc_vtk_add_scalar(self.h, name.encode(), f.ctypes.data, 0))

    def add_vector(self, name, u, v, w):
        u, v, w = (self._host(a) for a in (u, v, w))
        check(self._lib.scihpc_vtk_add_vector(self.h, name.encode(), u.ctypes.data, v.ctypes.data,
                                              w.ctypes.data, 0))

    def add_scalar_device(self, name, dptr):
        """Device pointer (int) to a padded field, complete before the call."""
        check(self._lib.scihpc_vtk_add_scalar(self.h, name.encode(), int(dptr), 1))

    def add_vector_device(self, name, du, dv, dw):
        check(self._lib.scihpc_vtk_add_vector(self.h, name.encode(), int(du), int(dv), int(dw), 1))

    def close(self):
        check(self._lib.scihpc_vtk_close(self.h))

    def time_pack(self, ncomp=1, iters=50):
        ms = C.c_double()
        check(self._lib.scihpc_vtk_time_pack(self.h, ncomp, iters, C.byref(ms)))
        return ms.value


def read_vtk(path):
    """Reads a legacy BINARY STRUCTURED_POINTS file written by vtkWriter (or
    VtkWriter) into {"dims", "origin", "spacing", "cells", <name>: array}.
    Scalars come back with shape (nz, ny, nx), vectors (nz, ny, nx, 3), in
    the file's big-endian FP64 converted to native floats."""
    raw = open(path, "rb").read()
    pos = 0
    out = {}

    def line():
        nonlocal pos
        e = raw.index(b"\n", pos)
        t = raw[pos:e].decode()
        pos = e + 1
        return t

    for _ in range(4):
        line()
    out["dims"] = tuple(int(x) for x in line().split()[1:])
    out["origin"] = tuple(float(x) for x in line().split()[1:])
    out["spacing"] = tuple(float(x) for x in line().split()[1:])
    cells = int(line().split()[1])
    out["cells"] = cells
    shape = tuple(max(d - 1, 1) for d in out["dims"])[::-1]
    while pos < len(raw):
        head = line().split()
        if head[0] == "SCALARS":
            line()  # LOOKUP_TABLE
            a = np.frombuffer(raw, ">f8", cells, pos).astype(np.float64).reshape(shape)
            pos += 8 * cells
        elif head[0] == "VECTORS":
            a = np.frombuffer(raw, ">f8", 3 * cells, pos).astype(np.float64).reshape(shape + (3,))
            pos += 24 * cells
        else:
            raise ValueError(f"unexpected VTK section {head[0]!r} in {path}")
        out[head[1]] = a
    return out
