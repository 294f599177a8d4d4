"""GL parameters: epsilon, kappa, sigma, H, external potential, Langevin
coefficients (svirl/vars/params.py)."""
import numpy as np

from svirl_amd._native import DeviceArray
from svirl_amd.storage import GArray

from .fixed_vortices import FixedVortices


class Params:
    """Setters and getters of svirl/vars/params.py:10-327."""

    def __init__(self, mesh, vars):
        self.mesh = mesh
        self.vars = vars
        self.cfg = vars.cfg
        cfg = self.cfg
        self.fixed_vortices = FixedVortices(self.mesh, self.vars)
        self.solveA = False
        self.linear_coefficient = cfg.linear_coefficient
        self.gl_parameter = cfg.gl_parameter
        self.normal_conductivity = cfg.normal_conductivity
        self._H = cfg.dtype(0.0)
        self.homogeneous_external_field_reset = cfg.homogeneous_external_field
        self.ae, self.be = None, None
        self._vpei = None
        self.external_field = cfg.external_field
        self.order_parameter_Langevin_coefficient = cfg.order_parameter_Langevin_coefficient
        self.vector_potential_Langevin_coefficient = cfg.vector_potential_Langevin_coefficient

    # -- epsilon ---------------------------------------------------------------
    @property
    def linear_coefficient(self):
        cfg = self.cfg
        if self._eps_map is None:
            return np.full((cfg.Nx, cfg.Ny), self._eps_scalar, dtype=cfg.dtype)
        return self._eps_host.copy()

    @linear_coefficient.setter
    def linear_coefficient(self, linear_coefficient):
        cfg = self.cfg
        if callable(linear_coefficient):
            # params.py:58-60 refers to an undefined `mesh` here; the mesh of
            # this solver is what it means
            xg, yg = self.mesh.xy_grid
            lc = linear_coefficient(xg, yg)
        else:
            lc = linear_coefficient
        if np.isscalar(lc):
            self._eps_scalar = cfg.dtype(lc)
            self._eps_map = None
            self._eps_host = None
        else:
            lc = np.asarray(lc)
            assert lc.shape == (cfg.Nx, cfg.Ny)
            self._eps_host = lc.astype(cfg.dtype)
            self._eps_map = DeviceArray.from_host(np.ascontiguousarray(self._eps_host.T).ravel(), cfg.dtype)
            self._eps_scalar = None

    def linear_coefficient_h(self):
        """Device epsilon map (None for a scalar)."""
        return self._eps_map

    def linear_coefficient_scalar_h(self):
        """The scalar epsilon handed to the kernels: 0 when a map is set
        (params.py:79-83)."""
        if self._eps_map is None:
            return self._eps_scalar
        return self.cfg.dtype(0.0)

    # -- kappa -------------------------------------------------------------------
    @property
    def gl_parameter(self):
        return self._kappa

    @gl_parameter.setter
    def gl_parameter(self, gl_parameter):
        if gl_parameter is None or np.isnan(gl_parameter) or np.isinf(gl_parameter):
            gl_parameter = np.inf
        assert isinstance(gl_parameter, (np.floating, float, np.integer, int)) and (
            np.isposinf(gl_parameter) or gl_parameter > 0.0)
        self._kappa = self.cfg.dtype(gl_parameter)
        self.solveA = bool(not np.isposinf(self._kappa))

    def gl_parameter_squared_h(self):
        if self.solveA:
            return self.cfg.dtype(self.gl_parameter ** 2)
        return self.cfg.dtype(-1.0)

    # -- sigma -------------------------------------------------------------------
    @property
    def normal_conductivity(self):
        return self._sigma

    @normal_conductivity.setter
    def normal_conductivity(self, normal_conductivity):
        assert isinstance(normal_conductivity, (np.floating, float, np.integer, int)) and normal_conductivity > 0.0
        self._sigma = self.cfg.dtype(normal_conductivity)
        self._rho = self.cfg.dtype(1.0 / normal_conductivity)

    # -- homogeneous field -------------------------------------------------------
    @property
    def homogeneous_external_field(self):
        return self._H

    @homogeneous_external_field.setter
    def homogeneous_external_field(self, homogeneous_external_field):
        self._H = self.cfg.dtype(homogeneous_external_field)

    def _update_vector_potential(self, homogeneous_external_field, reset):
        """Symmetric gauge A = (-g (y - Ly/2) dH, (1-g)(x - Lx/2) dH), g = 1/2
        (params.py:135-170)."""
        assert isinstance(homogeneous_external_field, (np.floating, float, np.integer, int))
        cfg = self.cfg
        vp = self.vars._vp
        vp.sync()
        if reset:
            self._H = cfg.dtype(homogeneous_external_field)
            a, b = vp.get_vec_h()
            a.fill(0.0)
            b.fill(0.0)
            vp.need_htod_sync()
            vp.sync()
            delta_H = self._H
        else:
            delta_H = -self._H
            self._H = cfg.dtype(homogeneous_external_field)
            delta_H += self._H
        g = 0.5
        _, yg = self.mesh.xy_a_grid
        xg, _ = self.mesh.xy_b_grid
        a, b = vp.get_vec_h()
        a -= g * (yg - 0.5 * cfg.Ly) * delta_H
        b += (1.0 - g) * (xg - 0.5 * cfg.Lx) * delta_H
        vp.need_htod_sync()
        vp.sync()

    def _homogeneous_external_field_delta(self, homogeneous_external_field):
        self._update_vector_potential(homogeneous_external_field, reset=False)

    homogeneous_external_field_delta = property(fset=_homogeneous_external_field_delta)

    def _homogeneous_external_field_reset(self, homogeneous_external_field):
        self._update_vector_potential(homogeneous_external_field, reset=True)

    homogeneous_external_field_reset = property(fset=_homogeneous_external_field_reset)

    # -- external / irregular vector potential ----------------------------------
    def _update_gvpei(self):
        """params.py:193-221: vpei = (ae, be) + (ai, bi)."""
        assert (self.ae is None) == (self.be is None)
        ai, bi = None, None
        if self.fixed_vortices is not None and self.fixed_vortices._vpi is not None:
            ai, bi = self.fixed_vortices._vpi.get_vec_h()
        if self.ae is not None:
            vpei = (self.ae + ai, self.be + bi) if ai is not None else (self.ae, self.be)
        else:
            vpei = (ai, bi)
        if vpei[0] is None:
            self._vpei = None
            return
        cfg = self.cfg
        self._vpei = GArray(shape=[vpei[0].shape, vpei[1].shape], dtype=cfg.dtype)
        self._vpei.set_vec_h(np.asarray(vpei[0], dtype=cfg.dtype), np.asarray(vpei[1], dtype=cfg.dtype))
        self._vpei.sync()

    @property
    def external_vector_potential(self):
        assert (self.ae is None) == (self.be is None)
        if self.ae is not None:
            return self.ae, self.be
        return None

    @external_vector_potential.setter
    def external_vector_potential(self, external_vector_potential):
        cfg = self.cfg
        if external_vector_potential is not None:
            Ax, Ay = external_vector_potential
            assert (Ax is None) == (Ay is None)
        else:
            Ax = Ay = None
        if Ax is not None:
            assert Ax.shape == (cfg.Nxa, cfg.Nya)
            assert Ay.shape == (cfg.Nxb, cfg.Nyb)
            self.ae, self.be = Ax, Ay
        else:
            self.ae, self.be = None, None
        self._update_gvpei()

    @property
    def external_irregular_vector_potential(self):
        if self._vpei is not None:
            return self._vpei.get_vec_h()
        return None

    def external_irregular_vector_potential_h(self):
        if self._vpei is not None:
            return self._vpei.get_d_obj()
        return None

    @property
    def external_field(self):
        A = self.external_vector_potential
        if A is not None:
            Ax, Ay = A
            return -np.diff(Ax, axis=1) * self.cfg.idy + np.diff(Ay, axis=0) * self.cfg.idx
        return None

    @external_field.setter
    def external_field(self, external_field):
        cfg = self.cfg
        if external_field is not None:
            g = 0.5
            _, yg = self.mesh.xy_a_grid
            xg, _ = self.mesh.xy_b_grid
            Ax = -g * (yg - 0.5 * cfg.Ly) * external_field
            Ay = (1.0 - g) * (xg - 0.5 * cfg.Lx) * external_field
            self.external_vector_potential = (Ax, Ay)
        else:
            self.external_vector_potential = None

    # -- Langevin ------------------------------------------------------------------
    @property
    def order_parameter_Langevin_coefficient(self):
        return self._psi_langevin_c

    @order_parameter_Langevin_coefficient.setter
    def order_parameter_Langevin_coefficient(self, c):
        assert isinstance(c, (np.floating, float, np.integer, int))
        self._psi_langevin_c = self.cfg.dtype(c)

    @property
    def vector_potential_Langevin_coefficient(self):
        return self._ab_langevin_c

    @vector_potential_Langevin_coefficient.setter
    def vector_potential_Langevin_coefficient(self, c):
        assert isinstance(c, (np.floating, float, np.integer, int))
        self._ab_langevin_c = self.cfg.dtype(c)
