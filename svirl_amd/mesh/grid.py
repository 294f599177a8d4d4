"""Coordinates of nodes, links and cells and the material tiling
(svirl/mesh/grid.py:11-183)."""
import numpy as np

from svirl_amd._native import DeviceArray


class Grid:
    """Grid coordinates and the per-cell material tiling (bool (Nxc, Nyc);
    stored on the device as one byte per cell, i + Nxc*j)."""

    def __init__(self, cfg):
        self.cfg = cfg
        self._mt = None      # host bool (Nxc, Nyc)
        self._gmt = None     # device uint8 flat
        self.material_tiling = cfg.material_tiling

    def have_material_tiling(self):
        return self._mt is not None

    @property
    def xy(self):
        c = self.cfg
        return (np.linspace(0.0, c.Lx, num=c.Nx, endpoint=True, dtype=c.dtype),
                np.linspace(0.0, c.Ly, num=c.Ny, endpoint=True, dtype=c.dtype))

    @property
    def xy_grid(self):
        x, y = self.xy
        return np.meshgrid(x, y, indexing='ij')

    @property
    def xy_a(self):
        c = self.cfg
        return (np.linspace(0.5 * c.dx, c.Lx - 0.5 * c.dx, num=c.Nxa, endpoint=True, dtype=c.dtype),
                np.linspace(0.0, c.Ly, num=c.Nya, endpoint=True, dtype=c.dtype))

    @property
    def xy_a_grid(self):
        x, y = self.xy_a
        return np.meshgrid(x, y, indexing='ij')

    @property
    def xy_b(self):
        c = self.cfg
        return (np.linspace(0.0, c.Lx, num=c.Nxb, endpoint=True, dtype=c.dtype),
                np.linspace(0.5 * c.dy, c.Ly - 0.5 * c.dy, num=c.Nyb, endpoint=True, dtype=c.dtype))

    @property
    def xy_b_grid(self):
        x, y = self.xy_b
        return np.meshgrid(x, y, indexing='ij')

    @property
    def xy_c(self):
        c = self.cfg
        return (np.linspace(0.5 * c.dx, c.Lx - 0.5 * c.dx, num=c.Nxc, endpoint=True, dtype=c.dtype),
                np.linspace(0.5 * c.dy, c.Ly - 0.5 * c.dy, num=c.Nyc, endpoint=True, dtype=c.dtype))

    @property
    def xy_c_grid(self):
        x, y = self.xy_c
        return np.meshgrid(x, y, indexing='ij')

    @property
    def material_tiling(self):
        if self._mt is not None:
            return self._mt.copy()
        return np.full((self.cfg.Nxc, self.cfg.Nyc), True, dtype=bool)

    @material_tiling.setter
    def material_tiling(self, material_tiling):
        if callable(material_tiling):
            xg, yg = self.xy_c_grid
            mt = material_tiling(xg, yg)
        else:
            mt = material_tiling
        if mt is not None:
            mt = np.asarray(mt)
            assert mt.shape == (self.cfg.Nxc, self.cfg.Nyc)
            self._mt = mt.astype(bool)
            flat = np.ascontiguousarray(self._mt.T).reshape(-1).astype(np.uint8)
            self._gmt = DeviceArray.from_host(flat, np.uint8)
        else:
            self._mt = None
            self._gmt = None

    def material_tiling_h(self):
        """Device buffer of the tiling (None when absent)."""
        return self._gmt

    def _get_material_tiling_at_nodes(self):
        """grid.py:160-171: a node is in material if any adjacent cell is."""
        mt = self._mt
        c = self.cfg
        mt_c = np.full((1, c.Nyc), False, dtype=bool)
        mt_r = np.full((c.Nx, 1), False, dtype=bool)
        return np.logical_or.reduce((
            np.c_[mt_r, np.r_[mt_c, mt]],
            np.c_[mt_r, np.r_[mt, mt_c]],
            np.c_[np.r_[mt_c, mt], mt_r],
            np.c_[np.r_[mt, mt_c], mt_r],
        ))

    def interpolate_ab_array_to_c_array(self, a, b):
        return 0.5 * (a[:, :-1] + a[:, 1:]), 0.5 * (b[:-1, :] + b[1:, :])

    def interpolate_ab_array_to_c_array_abs(self, a, b):
        cx, cy = self.interpolate_ab_array_to_c_array(a, b)
        return np.sqrt(np.square(cx) + np.square(cy))
