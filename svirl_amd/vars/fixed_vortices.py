"""Vortex seeding and fixed (pinned) vortices (svirl/vars/fixed_vortices.py).

Seeding phase windings into psi is host work. Fixed vortices (an irregular
vector potential vpi and phase-locked nodes) are built on the host like the
reference; the TD solver applies them with the reference's per-step hooks,
including their partial-length (first N entries) vpi updates, and the CG
sees them through the external irregular vector potential. Batched runs
reject them."""
import numpy as np

from svirl_amd.storage import GArray


class FixedVortices:
    def __init__(self, mesh, vars):
        self.mesh = mesh
        self.vars = vars
        self.cfg = vars.cfg
        self._vpi = None
        self._phase_lock_ns = None
        self._phase_lock_radius = self.cfg.phase_lock_radius
        if self.cfg.fixed_vortices_correction is None:
            self.cfg.fixed_vortices_correction = 'none'
        assert self.cfg.fixed_vortices_correction in ('none', 'cell centers', 'vertices')
        self.fixed_vortices_correction = self.cfg.fixed_vortices_correction
        self.fixed_vortices = self.cfg.fixed_vortices

    @staticmethod
    def _vortex_psi_abs(x, y):
        r2 = x ** 2 + y ** 2
        return (1.0 - np.exp(-r2)) / (1.0 + np.exp(-r2))

    def order_parameter_add_vortices(self, vortices, phase=True, deep=False):
        """Multiply psi by exp(i v atan2(y - vy, x - vx)) (and a core
        profile when deep) for each vortex (fixed_vortices.py:44-62)."""
        vx, vy, vv = self._vortices_format(vortices)
        xg, yg = self.mesh.xy_grid
        self.vars._psi.sync()
        psi = self.vars._psi.get_h()
        for k in range(vx.size):
            if phase:
                psi *= np.exp(1.0j * vv[k] * np.arctan2(yg - vy[k], xg - vx[k]))
            if deep:
                psi *= np.power(self._vortex_psi_abs(xg - vx[k], yg - vy[k]), np.abs(vv[k]))
        self.vars._psi.need_htod_sync()
        self.vars._psi.sync()

    def fixed_vortices_release(self):
        """Frees the pinned vortices (fixed_vortices.py:63-77): psi takes the
        phase of the irregular vector potential, exp(-i phase) with phase its
        line integral from node (0, 0) (_ab_phase), so that the windings it
        imposed become windings of psi itself; then the fixed vortices and
        the phase lock are cleared."""
        self.vars._psi.sync()
        psi = self.vars._psi.get_h()
        psi *= np.exp(-1.0j * self.fixed_vortices_phase)
        self.vars._psi.need_htod_sync()
        self.vars._psi.sync()
        self.fixed_vortices = None
        self.phase_lock_radius = None

    def _vortices_format(self, vortices):
        """Three equal-length arrays (x, y, vorticity) in the solver dtype
        from (x, y), (x, y, vorticity) or a dict with 'x', 'y' and an
        optional 'vorticity' (fixed_vortices.py:81-112). None stands for no
        entries; a single value is repeated to the others' length; the
        vorticity is 1 when only positions are given; entries with a NaN in
        any of the three are dropped. (For a dict the reference reads the
        vorticity as key 2, which raises; here it is key 'vorticity'.)"""
        if vortices is None:
            vortices = ([], [])
        assert isinstance(vortices, (list, tuple, dict)) and len(vortices) in (2, 3)
        if isinstance(vortices, dict):
            parts = (vortices['x'], vortices['y'], vortices.get('vorticity'))
        else:
            parts = tuple(vortices) + (None,) * (3 - len(vortices))
        x, y, v = (np.asarray([] if q is None else q).ravel() for q in parts)
        n = max(x.size, y.size, v.size)
        if x.size and not v.size:
            v = np.ones(1, dtype=int)
        assert all(q.size in (1, n) for q in (x, y, v))
        dt = self.cfg.dtype
        if n == 0:
            return tuple(np.zeros(0, dtype=dt) for _ in range(3))
        x, y, v = (np.broadcast_to(q, (n,)) for q in (x, y, v))
        keep = ~(np.isnan(x) | np.isnan(y) | np.isnan(v))
        return tuple(q[keep].astype(dt) for q in (x, y, v))

    def _fixed_vortices_correct(self):
        """Rounds the vorticities to integers and, by the chosen correction,
        moves each position to the nearest cell centre (where B lives) or
        grid vertex (where psi lives) (fixed_vortices.py:115-128)."""
        cfg = self.cfg
        self.fixed_vortices_vorticity = np.round(self.fixed_vortices_vorticity)
        offset = {'cell centers': 0.5, 'vertices': 0.0}.get(self.fixed_vortices_correction)
        if offset is not None:
            def snap(u, h):
                return h * (np.round(u / h + offset) - offset)
            self.fixed_vortices_x = snap(self.fixed_vortices_x, cfg.dx)
            self.fixed_vortices_y = snap(self.fixed_vortices_y, cfg.dy)

    def _ab_phase(self, a, b):
        cfg = self.cfg
        a = cfg.dx * np.r_[np.full((1, cfg.Nya), 0.0, dtype=cfg.dtype), a].cumsum(axis=0)
        b = cfg.dy * np.c_[np.full((cfg.Nxb, 1), 0.0, dtype=cfg.dtype), b].cumsum(axis=1)
        return np.repeat(b[0:1, :], cfg.Nx, axis=0) + a

    @property
    def irregular_vector_potential(self):
        cfg = self.cfg
        if self._vpi is None:
            return (np.zeros((cfg.Nxa, cfg.Nya), dtype=cfg.dtype), np.zeros((cfg.Nxb, cfg.Nyb), dtype=cfg.dtype))
        self._vpi.sync()
        return self._vpi.get_vec_h()

    def irregular_vector_potential_h(self):
        return self._vpi.get_d_obj() if self._vpi is not None else None

    @property
    def fixed_vortices_phase(self):
        ai, bi = self.irregular_vector_potential
        return self._ab_phase(ai, bi)

    @property
    def fixed_vortices(self):
        return (self.fixed_vortices_x.copy(), self.fixed_vortices_y.copy(), self.fixed_vortices_vorticity.copy())

    @fixed_vortices.setter
    def fixed_vortices(self, fixed_vortices):
        """Irregular vector potential of pinned vortices
        (fixed_vortices.py:168-199)."""
        cfg = self.cfg
        self.fixed_vortices_x, self.fixed_vortices_y, self.fixed_vortices_vorticity = self._vortices_format(fixed_vortices)
        self._fixed_vortices_correct()
        if self.fixed_vortices_x.size > 0:
            if self._vpi is None:
                self._vpi = GArray(shape=[(cfg.Nxa, cfg.Nya), (cfg.Nxb, cfg.Nyb)], dtype=cfg.dtype)
            ai, bi = self._vpi.get_vec_h()
            ai.fill(0.0)
            bi.fill(0.0)
            xg, yg = self.mesh.xy_grid
            for k in range(self.fixed_vortices_x.size):
                vangle = np.arctan2(yg - self.fixed_vortices_y[k], xg - self.fixed_vortices_x[k])
                vangle -= vangle[0, 0]
                ai += self.fixed_vortices_vorticity[k] * cfg.idx * np.diff(vangle, axis=0)
                bi += self.fixed_vortices_vorticity[k] * cfg.idy * np.diff(vangle, axis=1)
            self._vpi.need_htod_sync()
            self._vpi.sync()
        else:
            self._vpi = None
        self._update_phase_lock_ns()

    @property
    def phase_lock_radius(self):
        return self._phase_lock_radius

    @phase_lock_radius.setter
    def phase_lock_radius(self, radius):
        assert radius is None or (isinstance(radius, (np.floating, float, np.integer, int)) and radius > 0.0)
        self._phase_lock_radius = radius
        self._update_phase_lock_ns()

    def _update_phase_lock_ns(self):
        self._phase_lock_ns = None
        if self._phase_lock_radius is not None:
            cfg = self.cfg
            xg, yg = self.mesh.xy_grid
            lock = np.full((cfg.Nx, cfg.Ny), False, dtype=bool)
            for k in range(self.fixed_vortices_x.size):
                lock |= (np.square(xg - self.fixed_vortices_x[k]) + np.square(yg - self.fixed_vortices_y[k])
                         <= np.square(self._phase_lock_radius))
            ns = np.where(lock)[0].astype(np.int32)
            if ns.size > 0:
                self._phase_lock_ns = ns
