"""Batched TDGL: many GLSolver instances advanced together.

The reference runs one GLSolver per process, one time step at a time
(svirl/solvers/td.py:342-366). Here a list of solvers of the same precision
(each with its own grid, kappa, field, epsilon map and mask) advances by one
persistent launch per time step (glx_batch_td_run): every instance is swept
by its own team of workgroups with team-local barriers, so each instance
keeps its own Jacobi sweep counts and stop rule, exactly as if it had been
run alone. The step launches are replayed from a captured HIP graph.

The tile shape — sweeps per psi pass (the tile depth, which sets the
strip width: 64 - 2 x depth columns per wave), fewest rows per wave when
sizing the teams (the tile height) and workgroups per CU — is picked by an
autotuner: the first steps of a run time candidate shapes on the real state
(the results do not depend on the shape), one axis at a time from a default
(coordinate descent: depth, then height, then occupancy, each axis keeping
its fastest value), and the rest of the run uses the fastest; the choice is
cached per batch signature. Teams are sized by each instance's wave-rows
times its sweeps per step in the previous run, so instances that iterate
more get more workgroups.

    from svirl_amd import GLSolver
    from svirl_amd.batch import BatchTD
    sims = [GLSolver(...) for _ in range(256)]
    BatchTD(sims).td(dt=0.1, Nt=5000)

BatchCG minimises many instances at once by the reference's Polak-Ribiere
CG (cg.py:452-551): one workgroup per instance runs whole iterations with
FP64 reductions and an FP64 device line search, and each instance stops on
its own gradient-norm test.
"""
import ctypes

import numpy as np

from svirl_amd import _native as nat
from svirl_amd._native import BatchCGInstance, BatchInstance

# (signature) -> shape (wg_per_cu, min_rows, psi_depth)
_TUNED = {}

DEFAULT_SHAPE = (0, 8, 3)
# the tuner's axes, in the order it visits them: (index in the shape, values)
AXES = ((2, (3, 4, 2)), (1, (8, 4, 12, 16, 24)), (0, (0, 2, 1, 3)))
# shapes the tuner times at most: the default, then the other values of each axis
TUNE_RUNS = 1 + sum(len(v) - 1 for _, v in AXES)


class BatchTD:
    """Advance a list of GLSolver objects together (FP32 or FP64, all the
    same). use_graph: replay the step launches from a HIP graph.
    autotune: pick the tile shape on the first steps (AXES; needs
    tune_need() steps of the first td() call).
    noalloc: run through glx_batch_td_run_noalloc (no device allocation per
    call, its workspace is the library's fixed scratch array; no graph)."""

    def __init__(self, solvers, use_graph=True, autotune=True, tune_steps=4, noalloc=False):
        if not solvers:
            raise ValueError("BatchTD needs at least one solver")
        dt0 = solvers[0].cfg.dtype
        if any(s.cfg.dtype is not dt0 for s in solvers):
            raise ValueError("all solvers of a batch must share one dtype")
        self.solvers = list(solvers)
        self.prec = nat.prec_of(dt0)
        self.noalloc = bool(noalloc)
        self.use_graph = bool(use_graph) and not self.noalloc
        self.autotune = bool(autotune)
        self.tune_steps = int(tune_steps)
        self.shape = None
        self.timings = {}
        self.last_kernel_ms = 0.0
        self.weights = None   # per-instance sweeps per step of the last run (team sizing)

    def tune_need(self):
        """Steps the autotuner takes from the first td() call."""
        return self.tune_steps * (TUNE_RUNS + 1) + 1

    def _signature(self):
        return tuple((int(s.cfg.Nx), int(s.cfg.Ny), bool(s.params.solveA)) for s in self.solvers) + (self.prec,)

    def _tds(self, dt, Nt):
        tds = []
        for s in self.solvers:
            s.solve._init_td()
            td = s.solve._td
            td._set_iterator_options('order_parameter', dt=dt, Nt=Nt)
            td._set_iterator_options('vector_potential', dt=dt, Nt=Nt)
            td._stability_warnings()
            tds.append(td)
        return tds

    def _run(self, tds, nsteps, shape):
        n = len(tds)
        arr = (BatchInstance * n)()
        bufs = []
        for k, td in enumerate(tds):
            if td._has_fixed_vortices():
                raise NotImplementedError("fixed vortices / phase lock are not supported in batched runs")
            g, prm, P, pc, Q, qc = td._prepare(0, td.dt)
            if prm.langevin_psi > 1e-32 or prm.langevin_A > 1e-32:
                raise NotImplementedError("Langevin noise is not supported in batched runs")
            s = self.solvers[k]
            I = arr[k]
            I.g = g
            I.p = prm
            I.eps_map = nat.ptr(s.params.linear_coefficient_h())
            I.mt = nat.ptr(s.mesh.material_tiling_h())
            for j in range(3):
                I.psi[j] = P[j].ptr
                I.ab[j] = Q[j].ptr
            I.psi_cur = pc
            I.ab_cur = qc
            bufs.append((P, Q))
        sw = (ctypes.c_int32 * (2 * n * nsteps))()
        wt = None
        if self.weights is not None:
            wt = np.ascontiguousarray(self.weights, dtype=np.float64)
        wp = wt.ctypes.data if wt is not None else None
        if self.noalloc:
            nat.call("glx_batch_td_run_noalloc", self.prec, arr, n, nsteps, sw, shape[0], shape[1], wp, shape[2])
        else:
            nat.call("glx_batch_td_run", self.prec, arr, n, nsteps, sw, 1 if self.use_graph else 0, shape[0],
                     shape[1], wp, shape[2])
        kms = ctypes.c_float()
        nat.call("glx_td_last_kernel_ms", ctypes.byref(kms))
        self.last_kernel_ms = kms.value
        sweeps = np.frombuffer(sw, dtype=np.int32).reshape(nsteps, n, 2)
        # the next run sizes the teams by each instance's sweeps per step
        self.weights = np.maximum(sweeps.sum(axis=2).mean(axis=0), 1.0)
        for k, td in enumerate(tds):
            P, Q = bufs[k]
            td._finish(P, arr[k].psi_cur, Q, arr[k].ab_cur, sweeps[:, k, :], 0)
        return kms.value

    def td(self, dt=0.1, Nt=1000):
        """Nt time steps of dt for every instance (td.py:342-366 per
        instance)."""
        tds = self._tds(dt, Nt)
        done = 0
        sig = self._signature()
        if self.shape is None:
            self.shape = _TUNED.get(sig)
        if self.shape is None and self.autotune and Nt >= self.tune_need():
            if self.weights is None:  # learn the per-instance sweep counts first
                self._run(tds, self.tune_steps, DEFAULT_SHAPE)
                done += self.tune_steps

            def timed(shape):
                nonlocal done
                if shape not in self.timings:
                    self.timings[shape] = self._run(tds, self.tune_steps, shape) / self.tune_steps
                    done += self.tune_steps
                return self.timings[shape]

            best = DEFAULT_SHAPE
            timed(best)
            for axis, values in AXES:
                for v in values:
                    cand = best[:axis] + (v,) + best[axis + 1:]
                    if timed(cand) < self.timings[best]:
                        best = cand
            self.shape = best
            _TUNED[sig] = self.shape
        if self.shape is None:
            self.shape = DEFAULT_SHAPE
        if Nt - done > 0:
            self._run(tds, Nt - done, self.shape)
        for s in self.solvers:
            s.solve.vortices_detected = False


class BatchCG:
    """Free-energy minimisation of a list of GLSolver objects (one dtype;
    all kappa < inf or all kappa = inf). Each instance stops when
    ||grad F|| <= rtol * ||grad F||_0 or after max_iter iterations."""

    def __init__(self, solvers):
        if not solvers:
            raise ValueError("BatchCG needs at least one solver")
        dt0 = solvers[0].cfg.dtype
        if any(s.cfg.dtype is not dt0 for s in solvers):
            raise ValueError("all solvers of a batch must share one dtype")
        if len({bool(s.params.solveA) for s in solvers}) != 1:
            raise ValueError("all solvers of a batch must have kappa < inf, or all kappa = inf")
        self.solvers = list(solvers)
        self.prec = nat.prec_of(dt0)
        self.last_kernel_ms = 0.0

    def cg(self, max_iter=3000, rtol=1e-8, iters_per_launch=50, stall_iters=0, fp64_finish=False):
        """Returns (iterations, converged flags, final relative gradient
        norms) per instance.

        stall_iters > 0 stops an instance whose relative gradient norm has
        not halved within that many iterations. fp64_finish (FP32 batches,
        with a stall rule) continues the stalled instances in FP64 from
        where they stopped, against the FP32 run's initial gradient norm,
        and rounds their result back into the FP32 state. After the call,
        self.stalled and self.fp64_iters hold the per-instance stall flags
        and FP64-finish iteration counts."""
        if fp64_finish and self.prec != nat.GLX_F32:
            raise ValueError("the FP64 finish is for FP32 batches")
        if fp64_finish and stall_iters < 1:
            raise ValueError("the FP64 finish needs a stall rule (stall_iters >= 1)")
        n = len(self.solvers)
        arr = (BatchCGInstance * n)()
        for k, gl in enumerate(self.solvers):
            c, p, v = gl.cfg, gl.params, gl.vars
            if p.fixed_vortices._vpi is not None or p.fixed_vortices._phase_lock_ns is not None:
                raise NotImplementedError("fixed vortices are not supported by the batched CG")
            v._psi.sync_d()
            v._vp.sync_d()
            I = arr[k]
            I.g = nat.Grid(int(c.Nx), int(c.Ny), float(c.dx), float(c.dy))
            I.kappa2 = float(p.gl_parameter_squared_h())
            I.eps = float(p.linear_coefficient_scalar_h())
            I.H = float(p.homogeneous_external_field)
            I.eps_map = nat.ptr(p.linear_coefficient_h())
            I.mt = nat.ptr(gl.mesh.material_tiling_h())
            I.psi = v.order_parameter_h().ptr
            I.ab = nat.ptr(v.vector_potential_h()) if p.solveA else None
        nat.call("glx_batch_cg_run", self.prec, arr, n, int(max_iter), float(rtol), int(iters_per_launch),
                 int(stall_iters), int(bool(fp64_finish)))
        kms = ctypes.c_float()
        nat.call("glx_td_last_kernel_ms", ctypes.byref(kms))
        self.last_kernel_ms = kms.value
        for gl in self.solvers:
            gl.vars._psi.need_dtoh_sync()
            gl.vars._vp.need_dtoh_sync()
            gl.solve.vortices_detected = False
        iters = np.array([arr[k].iters for k in range(n)])
        conv = np.array([bool(arr[k].converged) for k in range(n)])
        grel = np.array([arr[k].grad_rel for k in range(n)])
        self.stalled = np.array([bool(arr[k].stalled) for k in range(n)])
        self.fp64_iters = np.array([arr[k].fp64_iters for k in range(n)])
        return iters, conv, grel
