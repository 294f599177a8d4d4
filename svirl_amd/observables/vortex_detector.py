"""Vortex detector (svirl/observables/vortex_detector.py): the per-cell
winding number (glx_vortex_cells) and the sub-cell triangulation of the
flagged cells (glx_vortex_triangulate) run on the GPU, which returns the
cells in ascending cell id, the order of the reference's loop over cells."""
import ctypes

import numpy as np

from svirl_amd import _native as nat

_VREC = 14


class VortexDetector:
    def __init__(self, _vars, params, solver):
        self.vars = _vars
        self.params = params
        self.fixed_vortices = params.fixed_vortices
        self.solver = solver
        self.cfg = _vars.cfg
        self._capacity = 1024
        self._rec = None
        self._xyv = None
        self._n_cells = 0
        self._cells_cache = None

    def unflatten_c(self, n):
        return n % self.cfg.Nxc, n // self.cfg.Nxc

    def flatten_c(self, i, j):
        return i + self.cfg.Nxc * j

    def _grid(self):
        c = self.cfg
        return nat.Grid(int(c.Nx), int(c.Ny), float(c.dx), float(c.dy))

    def _cells(self):
        """Runs the GPU winding kernel; the flagged cells' records stay on
        the device (self._rec), in ascending cell id. Returns their number."""
        c = self.cfg
        g = self._grid()
        self.vars._psi.sync_d()
        self.vars._vp.sync_d()
        abi = self.fixed_vortices.irregular_vector_potential_h()
        while True:
            if self._rec is None or self._rec.size < self._capacity * _VREC:
                self._rec = nat.DeviceArray(self._capacity * _VREC, np.float64)
            count = ctypes.c_int32()
            try:
                nat.call("glx_vortex_cells", nat.prec_of(c.dtype), g, self.vars.order_parameter_h().ptr,
                         self.vars.vector_potential_h().ptr, nat.ptr(abi), float(self.params.homogeneous_external_field),
                         self._rec.ptr, self._capacity, ctypes.byref(count))
                break
            except nat.GLXError as e:
                if e.code != nat.GLX_E_CAPACITY:
                    raise
                self._capacity = int(count.value) + 64
        return int(count.value)

    @property
    def vortex_cells(self):
        """(cell ids ascending, vorticity) of the flagged cells of the last
        detection, copied from the device on first use (the positions need
        only the triangulation's output)."""
        if self.solver.vortices_detected is not True:
            _ = self.vortices  # detection
        if self._cells_cache is None:
            n = self._n_cells
            rec = self._rec.get(n * _VREC).reshape(n, _VREC) if n else np.zeros((0, _VREC))
            self._cells_cache = (rec[:, 0].astype(np.int64), rec[:, 1].copy())
        return self._cells_cache

    @property
    def vortices(self):
        """Vortex positions and vorticities (x, y, v) with sub-cell precision
        (vortex_detector.py:31-156), in ascending cell id."""
        if self.solver.vortices_detected is not True:
            c = self.cfg
            n = self._cells()
            self._n_cells, self._cells_cache = n, None
            if n:
                if self._xyv is None or self._xyv.size < 4 * n:
                    self._xyv = nat.DeviceArray(4 * max(n, self._capacity), np.float64)
                nat.call("glx_vortex_triangulate", self._grid(), self._rec.ptr, n, self._xyv.ptr)
                xyv = self._xyv.get(4 * n).reshape(n, 4)
                det = xyv[xyv[:, 3] != 0.0, :3]
            else:
                det = np.zeros((0, 3))
            self._detected_vortices = det.astype(c.dtype)
            self.solver.vortices_detected = True
        d = self._detected_vortices
        return d[:, 0], d[:, 1], d[:, 2]
