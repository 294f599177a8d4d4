"""Free-energy minimisation by non-linear conjugate gradients with
Polak-Ribiere beta and an exact polynomial line search
(svirl/solvers/cg.py), on libglx kernels.

Per iteration: gradients (GPU), beta from two FP64 device reductions, the
search direction update (GPU), the line-search polynomial coefficients
(GPU, FP64 accumulation), the minimum of the polynomial, the update (GPU)
and the energy (GPU). Directions, previous gradients and betas persist
across cg() calls as in the reference (cg.py:54-71).

The line search follows `line_search`: 'host' (the FP64 default) reads the
coefficients back and solves them on the host exactly as the reference does
(numpy polyroots / scipy BFGS), which keeps the FP64 form's trajectory the
reference's; 'device' (the FP32-mixed default) minimises the same FP64
coefficients on the device and applies the update there (glx_cg_line_step),
so an iteration needs no host round trip but the energy read for the stop
test.
"""
import ctypes

import numpy as np
import scipy.optimize

from svirl_amd import _native as nat
from svirl_amd._native import DeviceArray


class CG:
    def __init__(self, par, mesh, _vars, params, observables):
        self.par = par
        self.mesh = mesh
        self.vars = _vars
        self.params = params
        self.fixed_vortices = params.fixed_vortices
        self.observables = observables
        self.cfg = cfg = par.cfg
        self.__convergence_rtol = cfg.convergence_rtol
        self.__reduction_vector_length = 17 if self.params.solveA else 5
        cd = cfg.dtype_complex
        self.__gdir_psi = DeviceArray.zeros(cfg.N, cd)
        self.__gdir_psi_next = DeviceArray(cfg.N, cd)  # the fused direction update writes here
        self.__gjac_psi = DeviceArray.zeros(cfg.N, cd)
        self.__gjac_psi_prev = DeviceArray.zeros(cfg.N, cd)
        # beta stays in device memory, as in the reference: [beta_psi, beta_A,
        # staged beta_psi, staged beta_A] (FP64 slots)
        self.__gbeta = DeviceArray.zeros(4, np.float64)
        if self.params.solveA:
            self.__gdir_A = DeviceArray.zeros(cfg.Nab, cfg.dtype)
            self.__gdir_A_next = DeviceArray(cfg.Nab, cfg.dtype)
            self.__gjac_A = DeviceArray.zeros(cfg.Nab, cfg.dtype)
            self.__gjac_A_prev = DeviceArray.zeros(cfg.Nab, cfg.dtype)
        self.__c = np.zeros((5, 5), dtype=cfg.dtype) if self.params.solveA else np.zeros(5, dtype=cfg.dtype)
        self.__galpha = DeviceArray.zeros(2, np.float64)
        self.line_search = 'device' if np.dtype(cfg.dtype) == np.float32 else 'host'
        self.cg_energies = []

    def _grid(self):
        c = self.cfg
        return nat.Grid(int(c.Nx), int(c.Ny), float(c.dx), float(c.dy))

    def _args(self):
        p = self.params
        return dict(k2=float(p.gl_parameter_squared_h()), eps=float(p.linear_coefficient_scalar_h()),
                    epsmap=nat.ptr(p.linear_coefficient_h()), H=float(p.homogeneous_external_field),
                    mt=nat.ptr(self.mesh.material_tiling_h()), abei=nat.ptr(p.external_irregular_vector_potential_h()),
                    ab=nat.ptr(self.vars.vector_potential_h()), psi=self.vars.order_parameter_h().ptr)

    # -- kernels exposed like the reference (used by its tests) -----------------------
    @property
    def _free_energy_jacobian_psi(self):
        """dF/d(Re psi, Im psi) * dxdy as a flat complex device array
        (cg.py:112-146)."""
        self.vars._psi.sync_d()
        self.vars._vp.sync_d()
        a = self._args()
        self.__gjac_psi.fill(0)
        nat.call("glx_cg_jacobian_psi", self.par.prec, self._grid(), a['k2'], a['eps'], a['epsmap'], a['H'], a['mt'],
                 a['psi'], a['abei'], a['ab'], self.__gjac_psi.ptr)
        return self.__gjac_psi

    @property
    def _free_energy_jacobian_A(self):
        """dF/dA (boundary links carry the reference's factor 2 on the field
        part) as a flat real device array [a; b] (cg.py:149-181)."""
        if not self.params.solveA:
            return None
        self.vars._psi.sync_d()
        self.vars._vp.sync_d()
        a = self._args()
        self.__gjac_A.fill(0)
        nat.call("glx_cg_jacobian_A", self.par.prec, self._grid(), a['k2'], a['H'], a['mt'], a['psi'], a['abei'],
                 a['ab'], self.__gjac_A.ptr)
        return self.__gjac_A

    def _free_energy_conjgrad_coef_psi(self, gdir_psi):
        """c0..c4 of F(psi + alpha dpsi) (cg.py:184-224)."""
        self.vars._psi.sync_d()
        self.vars._vp.sync_d()
        a = self._args()
        out = (ctypes.c_double * 5)()
        nat.call("glx_cg_coef_psi", self.par.prec, self._grid(), a['k2'], a['eps'], a['H'], a['mt'], a['psi'],
                 gdir_psi.ptr, a['abei'], a['ab'], out)
        self.__c = np.zeros(5, dtype=self.cfg.dtype) if self.__c.ndim != 1 else self.__c
        self.__c[:] = out[:]
        return self.__c

    def _cg_alpha_psi_min(self):
        """Smallest real root >= 0 of the cubic derivative (cg.py:227-235)."""
        c = self.__c
        am = np.polynomial.polynomial.polyroots([c[1], 2.0 * c[2], 3.0 * c[3], 4.0 * c[4]])
        am = am[np.isclose(am.imag, 0)].real
        am = am[am >= 0]
        return np.min(am)

    def _free_energy_conjgrad_coef(self, gdir_psi, gdir_A):
        """The 17 coefficients packed into a 5x5 polyval2d matrix
        (cg.py:325-375)."""
        self.vars._psi.sync_d()
        self.vars._vp.sync_d()
        a = self._args()
        out = (ctypes.c_double * 17)()
        nat.call("glx_cg_coef", self.par.prec, self._grid(), a['k2'], a['eps'], a['H'], a['mt'], a['psi'],
                 gdir_psi.ptr, a['abei'], a['ab'], gdir_A.ptr, out)
        g = np.array(out[:])
        c = np.zeros((5, 5), dtype=self.cfg.dtype)
        c[0, :] = g[0:5]
        c[1, :] = g[5:10]
        c[2, :] = g[10:15]
        c[3, 0] = g[15]
        c[4, 0] = g[16]
        self.__c = c
        return self.__c

    def _cg_alpha_min(self, alpha0=[0.0, 0.0], tol=1e-8):
        """scipy BFGS on the (alpha_psi, alpha_A) polynomial (cg.py:378-419)."""
        c = self.__c
        cj0 = np.polynomial.polynomial.polyder(c, axis=0)
        cj1 = np.polynomial.polynomial.polyder(c, axis=1)

        def f(alpha):
            return np.polynomial.polynomial.polyval2d(alpha[0], alpha[1], c)

        def j(alpha):
            return np.array([np.polynomial.polynomial.polyval2d(alpha[0], alpha[1], cj0),
                             np.polynomial.polynomial.polyval2d(alpha[0], alpha[1], cj1)])

        r = scipy.optimize.minimize(f, x0=np.array(alpha0), jac=j, method='BFGS', tol=tol)
        return r.x

    # -- hidden state (checkpoint / resume) ------------------------------------------------
    def _state(self):
        """The state that carries over between cg() calls in the reference
        (cg.py:54-71): the search directions and the betas."""
        d = {"cg_dir_psi": self.__gdir_psi.get(), "cg_beta": self.__gbeta.get()}
        if self.params.solveA:
            d["cg_dir_A"] = self.__gdir_A.get()
        return d

    def _set_state(self, d):
        self.__gdir_psi.set(d["cg_dir_psi"])
        self.__gbeta.set(d["cg_beta"])
        if self.params.solveA and "cg_dir_A" in d:
            self.__gdir_A.set(d["cg_dir_A"])

    # -- helpers ----------------------------------------------------------------------------
    @property
    def _beta_psi(self):
        """Polak-Ribiere beta of the psi direction (device scalar, read back)."""
        return self.cfg.dtype(self.__gbeta.get()[0])

    @property
    def _beta_A(self):
        return self.cfg.dtype(self.__gbeta.get()[1])

    def _beta_ptr(self, k):
        return ctypes.c_void_p(self.__gbeta.ptr + 8 * k)

    def _beta(self, jac, jac_prev, k):
        """beta_k = max(sum g(g - gp) / sum gp^2, 0), FP64 sums, left on the
        device (cg.py:422-449)."""
        nat.call("glx_cg_pr_beta", self.par.prec, jac.ptr, jac_prev.ptr, self._reals(jac), self._beta_ptr(k))

    def _axmy_beta(self, x, y, z, k):
        """z = beta_k x - y with the device beta (cg.py:283-290)."""
        nat.call("glx_axmy_dev", self.par.prec, x.ptr, y.ptr, z.ptr, self._beta_ptr(k), self._reals(x))

    def _reals(self, arr):
        return arr.size * (2 if arr.dtype.kind == 'c' else 1)

    def _axpy(self, x, y, z, alpha):
        nat.call("glx_axpy", self.par.prec, x.ptr, y.ptr, z.ptr, float(alpha), self._reals(x))

    # -- minimisation loops ----------------------------------------------------------------------
    #
    # Both loops follow cg.py:238-322 / 452-551 step for step in two passes
    # per iteration. The direction update rides in the line-search
    # coefficient pass (glx_cg_direction_step). The update of the state, its
    # gradient, energy and Polak-Ribiere sums share the next pass
    # (glx_cg_update_gradient_pass): it gives E_i for the stop test and the
    # gradient and the (staged) beta of iteration i+1. The staged beta only
    # becomes the solver's beta when iteration i+1 runs, so the state left
    # behind when the loop stops is the reference's.
    def __gradient_pass(self, withA):
        a = self._args()
        F = ctypes.c_double()
        nat.call("glx_cg_gradient_pass", self.par.prec, self._grid(), a['k2'], a['eps'], a['epsmap'], a['H'], a['mt'],
                 a['psi'], a['abei'], a['ab'], self.__gjac_psi.ptr, self.__gjac_psi_prev.ptr,
                 self.__gjac_A.ptr if withA else None, self.__gjac_A_prev.ptr if withA else None,
                 self._beta_ptr(2), ctypes.byref(F))
        return F.value

    def __direction_step(self, withA, c=None, beta_slot=0, swap=True):
        """dir = beta dir - jac for psi (and A) fused into the line-search
        coefficients (glx_cg_direction_step): with c (a ctypes array of 5 or
        17 doubles) the coefficients come back to the host; without, the
        device line search leaves the alphas in device memory. beta comes
        from the device slots beta_slot, beta_slot + 1 (0: the solver's
        betas, 2: the staged ones). The new directions are written to the
        second buffers, which become the current ones when swap is set."""
        a = self._args()
        nat.call("glx_cg_direction_step", self.par.prec, self._grid(), a['k2'], a['eps'], a['H'], a['mt'], a['psi'],
                 a['abei'], a['ab'], self.__gdir_psi.ptr, self.__gjac_psi.ptr, self.__gdir_psi_next.ptr,
                 self.__gdir_A.ptr if withA else None, self.__gjac_A.ptr if withA else None,
                 self.__gdir_A_next.ptr if withA else None, self._beta_ptr(beta_slot), c,
                 None if c is not None else self.__galpha.ptr)
        if swap:
            self.__swap_directions(withA)

    def __swap_directions(self, withA):
        self.__gdir_psi, self.__gdir_psi_next = self.__gdir_psi_next, self.__gdir_psi
        if withA:
            self.__gdir_A, self.__gdir_A_next = self.__gdir_A_next, self.__gdir_A

    def __update_gradient_pass(self, withA, wait=True):
        """The update state += alpha dir (alphas in device memory) fused into
        the gradient pass of the updated state (glx_cg_update_gradient_pass);
        the updated state goes to another buffer of the field's pool, which
        becomes the current one."""
        a = self._args()
        v = self.vars
        P, pc = v._psi.device_pool()
        psi_out = P[(pc + 1) % 3]
        ab_out = None
        if withA:
            Q, qc = v._vp.device_pool()
            ab_out = Q[(qc + 1) % 3]
        F = ctypes.c_double()
        nat.call("glx_cg_update_gradient_pass", self.par.prec, self._grid(), a['k2'], a['eps'], a['epsmap'], a['H'],
                 a['mt'], a['psi'], a['abei'], a['ab'], self.__gdir_psi.ptr, self.__gdir_A.ptr if withA else None,
                 self.__galpha.ptr, psi_out.ptr, nat.ptr(ab_out), self.__gjac_psi.ptr, self.__gjac_psi_prev.ptr,
                 self.__gjac_A.ptr if withA else None, self.__gjac_A_prev.ptr if withA else None,
                 self._beta_ptr(2), ctypes.byref(F) if wait else None)
        v._psi._gdata = psi_out
        if withA:
            v._vp._gdata = ab_out
        return F.value if wait else None

    def __commit_beta(self):
        nat.call("glx_memcpy_d2d", self._beta_ptr(0), self._beta_ptr(2), 16)

    def __loop(self, n_iter, withA):
        self.vars._psi.sync_d()
        self.vars._vp.sync_d()
        self.cg_energies = []
        self.__gradient_pass(withA)  # gradient of the starting state
        if self.line_search == 'device':
            self.__device_loop(n_iter, withA)
        else:
            self.__host_loop(n_iter, withA)
        self.vars._psi.need_dtoh_sync()
        if withA:
            self.vars._vp.need_dtoh_sync()

    def __converged(self, i):
        return i > 0 and np.abs(self.cg_energies[i] / self.cg_energies[i - 1] - 1.0) < self.__convergence_rtol

    def __device_loop(self, n_iter, withA):
        """The loop with the device line search. The next iteration's
        direction step (from the staged betas) is enqueued before this
        iteration's energy is read, so the GPU does not wait for the stop
        test; it writes only the second direction buffers and the alphas,
        so when the loop stops there the directions and betas are left as
        if it had not run."""
        cd = self.cfg.dtype
        if n_iter <= 0:
            return
        self.__direction_step(withA)
        for i in range(n_iter):
            if withA:
                self.__gjac_A, self.__gjac_A_prev = self.__gjac_A_prev, self.__gjac_A
            self.__gjac_psi, self.__gjac_psi_prev = self.__gjac_psi_prev, self.__gjac_psi
            self.__update_gradient_pass(withA, wait=False)
            more = i + 1 < n_iter
            if more:
                self.__direction_step(withA, beta_slot=2, swap=False)
            F = ctypes.c_double()
            nat.call("glx_cg_energy_wait", ctypes.byref(F))
            self.cg_energies.append(cd(F.value))
            if self.__converged(i):
                break
            if more:
                self.__commit_beta()
                self.__swap_directions(withA)

    def __host_loop(self, n_iter, withA):
        """The loop with the reference's host line search."""
        cd = self.cfg.dtype
        for i in range(n_iter):
            if i > 0:
                self.__commit_beta()
            if withA:
                out = (ctypes.c_double * 17)()
                self.__direction_step(True, out)
                g = np.array(out[:])
                c = np.zeros((5, 5), dtype=cd)
                c[0, :], c[1, :], c[2, :] = g[0:5], g[5:10], g[10:15]
                c[3, 0], c[4, 0] = g[15], g[16]
                self.__c = c
                alpha_psi, alpha_A = self._cg_alpha_min()
                self.__galpha.set(np.array([float(cd(alpha_psi)), float(cd(alpha_A))]))
            else:
                out = (ctypes.c_double * 5)()
                self.__direction_step(False, out)
                self.__c = np.zeros(5, dtype=cd) if self.__c.ndim != 1 else self.__c
                self.__c[:] = out[:]
                self.__galpha.set(np.array([float(cd(self._cg_alpha_psi_min())), 0.0]))
            if withA:
                self.__gjac_A, self.__gjac_A_prev = self.__gjac_A_prev, self.__gjac_A
            self.__gjac_psi, self.__gjac_psi_prev = self.__gjac_psi_prev, self.__gjac_psi
            # the update with this iteration's alphas, then the gradient,
            # energy and Polak-Ribiere sums of the updated state
            self.cg_energies.append(cd(self.__update_gradient_pass(withA)))
            if self.__converged(i):
                break

    def __minimization_psi(self, n_iter=1000):
        """cg.py:238-322 (kappa = inf): the direction restarts from zero."""
        assert not self.params.solveA
        self.__gdir_psi.fill(0)
        self.__loop(n_iter, False)

    def __minimization(self, n_iter=1000):
        """cg.py:452-551 (kappa < inf): directions and betas carry over."""
        self.__loop(n_iter, True)

    def _solve(self, n_iter=1000):
        """Fixed vortices enter the CG only through the external irregular
        vector potential (params._vpei = external A + vpi, params.py:193-221)
        and the phase lock not at all, as in the reference's cg.py."""
        if self.params.solveA:
            self.__minimization(n_iter)
        else:
            self.__minimization_psi(n_iter)
