"""Free energy, currents, field and superfluid density
(svirl/observables/observables.py) on libglx kernels."""
import ctypes

import numpy as np

from svirl_amd import _native as nat


class Observables:
    def __init__(self, par, mesh, vars, params):
        self.par = par
        self.mesh = mesh
        self.vars = vars
        self.params = params
        self.cfg = par.cfg
        self._edge2 = None  # second link-sized buffer for the fused pass

    def _grid(self):
        c = self.cfg
        return nat.Grid(int(c.Nx), int(c.Ny), float(c.dx), float(c.dy))

    def _sync_device(self):
        self.vars._psi.sync_d()
        self.vars._vp.sync_d()
        if self.params._vpei is not None:
            self.params._vpei.sync_d()

    @property
    def superfluid_density(self):
        """|psi|^2 on nodes (observables.py:29-33)."""
        self.vars._psi.sync_d()
        psi = self.vars._psi.get_h()
        return np.square(psi.real) + np.square(psi.imag)

    @property
    def magnetic_field(self):
        """B per cell from abei + ab (observables.py:36-58)."""
        self._sync_device()
        out = self.vars._tmp_cell_var_h()
        nat.call("glx_magnetic_field", self.par.prec, self._grid(),
                 nat.ptr(self.params.external_irregular_vector_potential_h()),
                 nat.ptr(self.vars.vector_potential_h()), out.ptr)
        c = self.cfg
        return np.reshape(out.get(), (c.Nyc, c.Nxc)).T.copy()

    def _edge_pair(self, dev):
        c = self.cfg
        v = dev.get()
        a = np.reshape(v[:c.Na], (c.Nya, c.Nxa)).T.copy()
        b = np.reshape(v[c.Na:], (c.Nyb, c.Nxb)).T.copy()
        return a, b

    @property
    def supercurrent_density(self):
        """J_s on links (observables.py:61-83)."""
        self._sync_device()
        out = self.vars._tmp_edge_var_h()
        nat.call("glx_supercurrent_density", self.par.prec, self._grid(), nat.ptr(self.mesh.material_tiling_h()),
                 self.vars.order_parameter_h().ptr, nat.ptr(self.params.external_irregular_vector_potential_h()),
                 nat.ptr(self.vars.vector_potential_h()), out.ptr)
        return self._edge_pair(out)

    @property
    def current_density(self):
        """kappa^2 (curl B - H) on links; J_s when kappa = inf
        (observables.py:86-110)."""
        if not self.params.solveA:
            return self.supercurrent_density
        self._sync_device()
        out = self.vars._tmp_edge_var_h()
        nat.call("glx_current_density", self.par.prec, self._grid(), float(self.params.gl_parameter_squared_h()),
                 float(self.params.homogeneous_external_field),
                 nat.ptr(self.params.external_irregular_vector_potential_h()),
                 nat.ptr(self.vars.vector_potential_h()), out.ptr)
        return self._edge_pair(out)

    @property
    def normalcurrent_density(self):
        """J - J_s on links (observables.py:113-121); with A solved, J and
        J_s come from one fused pass (glx_bulk_observables)."""
        if not self.params.solveA:
            jx, jy = self.current_density
            jsx, jsy = self.supercurrent_density
            return jx - jsx, jy - jsy
        _, (jx, jy), (jsx, jsy) = self._bulk(False)
        return jx - jsx, jy - jsy

    def bulk_observables(self):
        """(B, (Jx, Jy), (Jsx, Jsy)) as magnetic_field, current_density and
        supercurrent_density return them, from one pass over the nodes."""
        return self._bulk(True)

    def _bulk(self, want_B):
        self._sync_device()
        c = self.cfg
        p = self.params
        if self._edge2 is None:
            self._edge2 = nat.DeviceArray(int(c.Na + c.Nb), c.dtype)
        B = self.vars._tmp_cell_var_h() if want_B else None
        J = self.vars._tmp_edge_var_h() if p.solveA else None
        Js = self._edge2
        nat.call("glx_bulk_observables", self.par.prec, self._grid(),
                 float(p.gl_parameter_squared_h()) if p.solveA else 0.0, float(p.homogeneous_external_field),
                 nat.ptr(self.mesh.material_tiling_h()), self.vars.order_parameter_h().ptr,
                 nat.ptr(p.external_irregular_vector_potential_h()), nat.ptr(self.vars.vector_potential_h()),
                 nat.ptr(B), nat.ptr(J), Js.ptr)
        Bh = np.reshape(B.get(), (c.Nyc, c.Nxc)).T.copy() if want_B else None
        js = self._edge_pair(Js)
        # J = J_s when kappa = inf (observables.py:86-110)
        jj = self._edge_pair(J) if J is not None else (js[0].copy(), js[1].copy())
        return Bh, jj, js

    @property
    def free_energy(self):
        """Total GL free energy (observables.py:124-149), FP64 accumulation."""
        self._sync_device()
        F = ctypes.c_double()
        p = self.params
        nat.call("glx_free_energy", self.par.prec, self._grid(), float(p.gl_parameter_squared_h()),
                 float(p.linear_coefficient_scalar_h()), nat.ptr(p.linear_coefficient_h()),
                 float(p.homogeneous_external_field), nat.ptr(self.mesh.material_tiling_h()),
                 self.vars.order_parameter_h().ptr, nat.ptr(p.external_irregular_vector_potential_h()),
                 nat.ptr(self.vars.vector_potential_h()), ctypes.byref(F))
        return self.cfg.dtype(F.value)

    def free_energy_fp64(self):
        """The same total as a Python float, without the cast to cfg.dtype."""
        self._sync_device()
        F = ctypes.c_double()
        p = self.params
        nat.call("glx_free_energy", self.par.prec, self._grid(), float(p.gl_parameter_squared_h()),
                 float(p.linear_coefficient_scalar_h()), nat.ptr(p.linear_coefficient_h()),
                 float(p.homogeneous_external_field), nat.ptr(self.mesh.material_tiling_h()),
                 self.vars.order_parameter_h().ptr, nat.ptr(p.external_irregular_vector_potential_h()),
                 nat.ptr(self.vars.vector_potential_h()), ctypes.byref(F))
        return F.value
