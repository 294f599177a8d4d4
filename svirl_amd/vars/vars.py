"""Order parameter and vector potential storage (svirl/vars/vars.py)."""
import numpy as np

from svirl_amd.storage import GArray


class Vars:
    """psi (complex, on nodes) and A = (a, b) (real, on links), with the
    getters/setters of svirl/vars/vars.py:53-134."""

    def __init__(self, par, mesh):
        self.par = par
        self.mesh = mesh
        self.cfg = par.cfg
        self._psi = None
        self._vp = None
        self._tmp_node_var = None
        self._tmp_edge_var = None
        self._tmp_cell_var = None
        self.solveA = bool(not np.isposinf(self.cfg.gl_parameter))
        cfg = self.cfg
        if isinstance(cfg.order_parameter, str) and cfg.order_parameter == 'random':
            self.order_parameter = 1.0
            self.randomize_order_parameter(level=cfg.random_level, seed=cfg.random_seed)
        else:
            self.order_parameter = cfg.order_parameter
        shapes = [(cfg.Nxa, cfg.Nya), (cfg.Nxb, cfg.Nyb)]
        self._vp = GArray(shape=shapes, dtype=cfg.dtype)

    @property
    def order_parameter(self):
        self._psi.sync()
        return self._psi.get_h().copy()

    @order_parameter.setter
    def order_parameter(self, order_parameter):
        cfg = self.cfg
        if isinstance(order_parameter, (np.complexfloating, complex, np.floating, float, np.integer, int)):
            order_parameter = cfg.dtype_complex(order_parameter) * np.ones((cfg.Nx, cfg.Ny), cfg.dtype_complex)
        order_parameter = np.asarray(order_parameter)
        assert order_parameter.shape == (cfg.Nx, cfg.Ny)
        if self._psi is None:
            self._psi = GArray(like=order_parameter.astype(cfg.dtype_complex))
        else:
            self._psi.set_h(order_parameter)
        self.set_order_parameter_to_zero_outside_material()
        self._psi.sync()

    def order_parameter_h(self):
        return self._psi.get_d_obj()

    def set_order_parameter_to_zero_outside_material(self):
        if self._psi is None or not self.mesh.have_material_tiling():
            return
        mt_at_nodes = self.mesh._get_material_tiling_at_nodes()
        self._psi.sync()
        psi = self._psi.get_h()
        psi[~mt_at_nodes] = 0.0
        self._psi.need_htod_sync()
        self._psi.sync()

    def randomize_order_parameter(self, level=1.0, seed=None):
        """|psi| = 1 - level*rand, phase = level*pi*(2*rand - 1), drawn from
        numpy's legacy MT19937 stream (vars.py:93-109)."""
        assert 0.0 <= level <= 1.0
        self._psi.sync()
        if seed is not None:
            np.random.seed(seed)
        N = self.cfg.N
        data = (1.0 - level * np.random.rand(N)) * np.exp(level * 1.0j * np.pi * (2.0 * np.random.rand(N) - 1.0))
        self._psi.set_h(data)
        self._psi.sync()

    @property
    def vector_potential(self):
        if self._vp is None:
            cfg = self.cfg
            return (np.zeros((cfg.Nxa, cfg.Nya), dtype=cfg.dtype), np.zeros((cfg.Nxb, cfg.Nyb), dtype=cfg.dtype))
        self._vp.sync()
        a, b = self._vp.get_vec_h()
        return a.copy(), b.copy()

    @vector_potential.setter
    def vector_potential(self, vector_potential):
        a, b = vector_potential
        self._vp.sync()
        self._vp.set_vec_h(np.asarray(a), np.asarray(b))
        self._vp.sync()

    def vector_potential_h(self):
        if self._vp is not None:
            return self._vp.get_d_obj()
        return None

    def _tmp_edge_var_h(self):
        if self._tmp_edge_var is None:
            cfg = self.cfg
            self._tmp_edge_var = GArray(shape=[(cfg.Nxa, cfg.Nya), (cfg.Nxb, cfg.Nyb)], dtype=cfg.dtype)
        return self._tmp_edge_var.get_d_obj()

    def _tmp_cell_var_h(self):
        if self._tmp_cell_var is None:
            cfg = self.cfg
            self._tmp_cell_var = GArray(shape=(cfg.Nxc, cfg.Nyc), dtype=cfg.dtype)
        return self._tmp_cell_var.get_d_obj()
