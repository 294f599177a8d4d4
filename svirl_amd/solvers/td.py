"""Time-dependent GL: backward-Euler psi and A updates by Jacobi
iteration (svirl/solvers/td.py), run as persistent GPU launches.

Each launch advances several time steps with every Jacobi sweep and its
stop test on the device (glx_td_run). The state buffers rotate among three
device arrays per field instead of the reference's right-hand-side copy and
pointer swap; vars._psi / vars._vp always point at the one holding the
current state. With fixed vortices or a phase lock the reference's per-step
hooks run between one-step launches (TD._run_fixed).
"""
import ctypes

import numpy as np

from svirl_amd import _native as nat
from svirl_amd._native import DeviceArray


class TD:
    MAX_SWEEPS = 1024  # Jacobi cap per solve (td.py:164, 274)

    def __init__(self, par, mesh, _vars, params, observables):
        self.par = par
        self.mesh = mesh
        self.vars = _vars
        self.params = params
        self.fixed_vortices = params.fixed_vortices
        self.observables = observables
        self.cfg = cfg = par.cfg
        self.solveA = self.params.solveA
        self._random_t = np.uint32(1)
        if cfg.random_seed is not None:
            self._random_t = np.uint32(cfg.random_seed)
        # stop criteria clamps (td.py:51-66)
        if cfg.dtype is np.float32:
            cfg.stop_criterion_order_parameter = max(cfg.stop_criterion_order_parameter, 1e-6)
            cfg.stop_criterion_vector_potential = max(cfg.stop_criterion_vector_potential, 1e-6)
        else:
            cfg.stop_criterion_order_parameter = max(cfg.stop_criterion_order_parameter, 1e-12)
            cfg.stop_criterion_vector_potential = max(cfg.stop_criterion_vector_potential, 1e-12)
        cfg.stop_criterion_order_parameter = cfg.dtype(cfg.stop_criterion_order_parameter)
        cfg.stop_criterion_vector_potential = cfg.dtype(cfg.stop_criterion_vector_potential)
        self.Nt = self.dt = self.T = None
        self.NtA = self.dtA = self.TA = None
        self._rhs_psi = None
        self._rhs_ab = None
        self.sweeps = []          # (S_psi, S_A) per time step
        self.wg = 0               # persistent grid (0: chosen per grid size)
        self.steps_per_launch = None
        self.td_energies = []

    # -- options (td.py:75-107) ---------------------------------------------------
    def _set_iterator_options(self, iterator_type, Nt=None, dt=None, T=None, mandatory_definition=True):
        assert iterator_type in ['order_parameter', 'vector_potential']
        if Nt is not None and T is not None and dt is None:
            dt = float(T) / Nt
        elif Nt is not None and T is None and dt is not None:
            T = float(dt) * Nt
        elif Nt is None and T is not None and dt is not None:
            Nt = int(np.round(T / dt))
        elif Nt is not None and T is not None and dt is not None:
            assert np.isclose(T, dt * Nt)
        if mandatory_definition:
            assert isinstance(dt, (np.floating, float, np.integer, int)) and dt >= 0.0
            assert isinstance(Nt, (np.integer, int)) and Nt >= 0
        cd = self.cfg.dtype
        if iterator_type == 'order_parameter':
            self.Nt = np.int32(Nt) if Nt is not None else None
            self.dt = cd(dt) if dt is not None else None
            self.T = cd(T) if T is not None else None
        else:
            self.NtA = np.int32(Nt) if Nt is not None else None
            self.dtA = cd(dt) if dt is not None else None
            self.TA = cd(T) if T is not None else None

    def _stability_warnings(self, psi=True, A=True):
        cfg = self.cfg
        if psi and self.dt is not None:
            d = self.dt / min(cfg.dx, cfg.dy) ** 2
            if d > 1.0:
                print('Warning (order parameter):  dt/min(dx,dy)^2 = %g is too large' % d)
        if A and self.solveA and self.dtA is not None:
            d = self.dtA * self.params.gl_parameter ** 2 / (self.params.normal_conductivity * min(cfg.dx, cfg.dy) ** 2)
            if d > 1.0:
                print('Warning (vector potential):  dt*kappa^2/(sigma*min(dx,dy)^2) = %g is too large' % d)
            d = self.dtA / (self.params.normal_conductivity * min(cfg.dx, cfg.dy))
            if d > 1.0:
                print('Warning (vector potential):  dt/(sigma*min(dx,dy)) = %g is too large' % d)

    # -- buffers ---------------------------------------------------------------------
    def _rotation(self, garray):
        """The field's three rotation buffers and the index of the current
        one (GArray.device_pool, shared with the CG)."""
        return garray.device_pool()

    def _default_wg(self):
        """Persistent grid: fewer workgroups for small grids (about 32 rows
        per wave); 0 (the library's default: two workgroups per CU for the
        psi solve and three for the A solve, the measured optima at 4097^2)
        once a grid fills two workgroups per CU."""
        cfg = self.cfg
        strips = -(-int(cfg.Nx) // 62)
        rows = strips * int(cfg.Ny)
        waves = max(1, rows // 32)
        maxg = ctypes.c_int32()
        nat.call("glx_td_max_grid", self.par.prec, 256, ctypes.byref(maxg))
        cus = nat.device_info()["num_cus"]
        wg = int(max(1, min(maxg.value, -(-waves // 4))))
        return 0 if wg >= min(maxg.value, 2 * cus) else wg

    def _launch_steps(self):
        if self.steps_per_launch:
            return int(self.steps_per_launch)
        return int(max(1, min(1000, 2.0e8 // max(1, int(self.cfg.N)))))

    # -- the time loop ------------------------------------------------------------------
    def _has_fixed_vortices(self):
        return self.fixed_vortices._vpi is not None or self.fixed_vortices._phase_lock_ns is not None

    def _prepare(self, eqn_code, dt):
        """Buffers and glx_td_params of a run (shared with batched runs)."""
        cfg = self.cfg
        v = self.vars
        v._psi.sync_d()
        v._vp.sync_d()
        P, pc = self._rotation(v._psi)
        Q, qc = self._rotation(v._vp)
        p = self.params
        lp = float(p.order_parameter_Langevin_coefficient)
        lA = float(p.vector_potential_Langevin_coefficient)
        if lp > 1e-32 and self._rhs_psi is None:
            self._rhs_psi = DeviceArray(P[0].size, P[0].dtype)
        if lA > 1e-32 and self._rhs_ab is None:
            self._rhs_ab = DeviceArray(Q[0].size, Q[0].dtype)
        prm = nat.TDParams()
        prm.dt = float(dt)
        prm.eps = float(p.linear_coefficient_scalar_h())
        prm.kappa2 = float(p.gl_parameter_squared_h())
        prm.rho = float(p._rho)
        prm.H = float(p.homogeneous_external_field)
        prm.stop_psi = float(cfg.stop_criterion_order_parameter)
        prm.stop_A = float(cfg.stop_criterion_vector_potential)
        prm.langevin_psi = lp
        prm.langevin_A = lA
        prm.solveA = 1 if self.solveA else 0
        prm.max_sweeps = self.MAX_SWEEPS
        prm.eqn = eqn_code
        prm.rand_t = int(self._random_t)
        g = nat.Grid(int(cfg.Nx), int(cfg.Ny), float(cfg.dx), float(cfg.dy))
        return g, prm, P, pc, Q, qc

    def _finish(self, P, pc, Q, qc, sweeps, eqn_code):
        """Record a run's sweep counts and point the variables at the final
        buffers."""
        arr = np.asarray(sweeps, dtype=np.int32).reshape(-1, 2)
        self.sweeps.extend(map(tuple, arr.tolist()))
        per_step = (1 if eqn_code != 2 else 0) + (1 if (self.solveA and eqn_code != 1) else 0)
        self._random_t = np.uint32((int(self._random_t) + per_step * len(arr)) & 0xffffffff)
        v = self.vars
        v._psi._gdata = P[pc]
        v._vp._gdata = Q[qc]
        v._psi.need_dtoh_sync()
        v._vp.need_dtoh_sync()

    # -- the time loop ------------------------------------------------------------------
    def _run(self, nsteps, eqn_code, dt, a_src=None):
        if nsteps <= 0:
            return
        g, prm, P, pc, Q, qc = self._prepare(eqn_code, dt)
        prm.a_src = nat.ptr(a_src)
        p = self.params
        Pp = (ctypes.c_void_p * 3)(*[b.ptr for b in P])
        Qp = (ctypes.c_void_p * 3)(*[b.ptr for b in Q])
        cpc, cqc = ctypes.c_int32(pc), ctypes.c_int32(qc)
        wg = self.wg or self._default_wg()
        chunk = self._launch_steps()
        done = 0
        while done < nsteps:
            n = min(chunk, nsteps - done)
            prm.rand_t = int(self._random_t)
            sw = (ctypes.c_int32 * (2 * n))()
            nat.call("glx_td_run", self.par.prec, g, ctypes.byref(prm), nat.ptr(p.linear_coefficient_h()),
                     nat.ptr(self.mesh.material_tiling_h()), Pp, ctypes.byref(cpc), Qp, ctypes.byref(cqc),
                     nat.ptr(self._rhs_psi), nat.ptr(self._rhs_ab), n, sw, wg, 256)
            self._finish(P, cpc.value, Q, cqc.value, np.frombuffer(sw, dtype=np.int32), eqn_code)
            done += n

    def _lock_ns_device(self):
        ns = self.fixed_vortices._phase_lock_ns
        if ns is None:
            return None
        cached = getattr(self, '_lock_cache', None)
        if cached is None or cached[0].shape != ns.shape or not np.array_equal(cached[0], ns):
            cached = (np.array(ns, copy=True), DeviceArray.from_host(ns, np.int32))
            self._lock_cache = cached
        return cached[1]

    def _run_fixed(self, nsteps, eqn_code, dt):
        """Fixed vortices and phase lock, with the reference's hooks around
        one-step launches: vp += vpi before and vp -= vpi after each psi
        solve (td.py:118-154; with eqn='order_parameter' the addition is
        made once before the loop, td.py:221-231), |psi| at the locked nodes
        after it (td.py:205-216), vpi += vp before and vpi -= vp after each
        A solve (td.py:256-266, 318-325), all over the first N entries as
        in the reference. The A solve reads its source from vpi
        (glx_td_params.a_src) when vpi is set."""
        fv = self.fixed_vortices
        vpi = fv.irregular_vector_potential_h()
        ns = self._lock_ns_device()
        N = int(self.cfg.N)
        prec = self.par.prec
        v = self.vars

        def xpy(x, y):
            nat.call("glx_axpy", prec, x.ptr, y.ptr, x.ptr, 1.0, N)

        def xmy(x, y):
            nat.call("glx_axmy", prec, x.ptr, y.ptr, x.ptr, 1.0, N)

        def psi_step():
            self._run(1, 1, dt)
            if ns is not None:
                nat.call("glx_td_phase_lock", prec, v._psi.get_d_obj().ptr, int(ns.size), ns.ptr)
            return self.sweeps.pop()[0]

        def A_step():
            if vpi is not None:
                xpy(vpi, v._vp.get_d_obj())
            self._run(1, 2, dt, a_src=vpi)
            if vpi is not None:
                xmy(vpi, v._vp.get_d_obj())
            return self.sweeps.pop()[1]

        v._psi.sync_d()
        v._vp.sync_d()
        if vpi is not None:
            fv._vpi.sync_d()
        if eqn_code == 1:
            if vpi is not None:
                xpy(v._vp.get_d_obj(), vpi)
            for _ in range(nsteps):
                self.sweeps.append((psi_step(), 0))
                if vpi is not None:
                    xmy(v._vp.get_d_obj(), vpi)
        elif eqn_code == 2:
            for _ in range(nsteps):
                self.sweeps.append((0, A_step()))
        else:
            for _ in range(nsteps):
                if vpi is not None:
                    xpy(v._vp.get_d_obj(), vpi)
                sp = psi_step()
                if vpi is not None:
                    xmy(v._vp.get_d_obj(), vpi)
                self.sweeps.append((sp, A_step() if self.solveA else 0))
        if vpi is not None:
            fv._vpi.need_dtoh_sync()
        v._vp.need_dtoh_sync()

    def _steps(self, nsteps, eqn_code, dt):
        if self._has_fixed_vortices():
            self._run_fixed(nsteps, eqn_code, dt)
        else:
            self._run(nsteps, eqn_code, dt)

    def _solve(self, dt=None, Nt=None, T=None, eqn=None):
        """td.py:342-377: eqn None (psi then A each step), 'order_parameter'
        or 'vector_potential'. The A update uses the psi time step dt
        (td.py:278), as in the reference."""
        if eqn == "order_parameter":
            self._set_iterator_options('order_parameter', dt=dt, Nt=Nt, T=T)
            self._stability_warnings(psi=True, A=False)
            self._steps(int(self.Nt), 1, self.dt)
        elif eqn == "vector_potential":
            if not self.solveA:
                return
            self._set_iterator_options('vector_potential', dt=dt, Nt=Nt, T=T)
            self._stability_warnings(psi=False, A=True)
            if self.dt is None:
                raise TypeError("the vector-potential update uses the order-parameter time step, which is not set yet")
            self._steps(int(self.NtA), 2, self.dt)
        else:
            self._set_iterator_options('order_parameter', dt=dt, Nt=Nt, T=T)
            self._set_iterator_options('vector_potential', dt=dt, Nt=Nt, T=T)
            self._stability_warnings()
            self.td_energies = []
            self._steps(int(self.Nt), 0, self.dt)
