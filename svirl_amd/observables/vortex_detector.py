"""Vortex detector (svirl/observables/vortex_detector.py): the per-cell
winding number runs on the GPU (glx_vortex_cells); the sub-cell
triangulation of the few flagged cells runs on the host."""
import ctypes

import numpy as np

from svirl_amd import _native as nat

_VREC = 14


def _find_zero(x1, y1, f1, x2, y2, f2):
    return (f2 * x1 - x2 * f1) / (f2 - f1), (f2 * y1 - y2 * f1) / (f2 - f1)


def _find_zero_line(x1, y1, f1, x2, y2, f2, x3, y3, f3, x4, y4, f4):
    """Points where the bilinear f crosses zero on the cell border
    (vortex_detector.py:230-236)."""
    xy = []
    if f2 * f1 < -1e-10:
        xy.append(_find_zero(x2, y2, f2, x1, y1, f1))
    if f3 * f2 < -1e-10:
        xy.append(_find_zero(x3, y3, f3, x2, y2, f2))
    if f4 * f3 < -1e-10:
        xy.append(_find_zero(x4, y4, f4, x3, y3, f3))
    if f1 * f4 < -1e-10:
        xy.append(_find_zero(x1, y1, f1, x4, y4, f4))
    return xy


def _find_intersection(l1x1, l1y1, l1x2, l1y2, l2x1, l2y1, l2x2, l2y2):
    """Intersection of two lines; NaN when their angle folds to 0 modulo
    pi/2 (vortex_detector.py:239-252)."""
    D = (l1x1 - l1x2) * (l2y1 - l2y2) - (l1y1 - l1y2) * (l2x1 - l2x2)
    ph = np.arctan2(D, (l1y1 - l1y2) * (l2y1 - l2y2) - (l1x1 - l1x2) * (l2x1 - l2x2))
    ph = np.mod(np.abs(ph), 0.5 * np.pi)
    if np.abs(ph) > 1e-10:
        ix = ((l1x1 * l1y2 - l1y1 * l1x2) * (l2x1 - l2x2) - (l1x1 - l1x2) * (l2x1 * l2y2 - l2y1 * l2x2)) / D
        iy = ((l1x1 * l1y2 - l1y1 * l1x2) * (l2y1 - l2y2) - (l1y1 - l1y2) * (l2x1 * l2y2 - l2y1 * l2x2)) / D
        return ix, iy, ph
    return np.nan, np.nan, ph


class VortexDetector:
    def __init__(self, _vars, params, solver):
        self.vars = _vars
        self.params = params
        self.fixed_vortices = params.fixed_vortices
        self.solver = solver
        self.cfg = _vars.cfg
        self._capacity = 1024
        self._rec = None
        self.vortex_cells = None  # (cell ids ascending, vorticity) of the last detection

    def unflatten_c(self, n):
        return n % self.cfg.Nxc, n // self.cfg.Nxc

    def flatten_c(self, i, j):
        return i + self.cfg.Nxc * j

    def _cells(self):
        """Flagged cells from the GPU winding kernel, ascending n."""
        c = self.cfg
        g = nat.Grid(int(c.Nx), int(c.Ny), float(c.dx), float(c.dy))
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
        n = count.value
        rec = self._rec.get()[:n * _VREC].reshape(n, _VREC) if n else np.zeros((0, _VREC))
        order = np.argsort(rec[:, 0], kind='stable')
        return rec[order]

    @property
    def vortices(self):
        """Vortex positions and vorticities (x, y, v) with sub-cell precision
        (vortex_detector.py:182-307)."""
        if self.solver.vortices_detected is not True:
            c = self.cfg
            rec = self._cells()
            self.vortex_cells = (rec[:, 0].astype(np.int64), rec[:, 1].copy())
            det = []
            for r in rec:
                n, v = int(r[0]), r[1]
                i, j = self.unflatten_c(n)
                x, y = c.dx * i, c.dy * j
                tp00, tpp0, tppp, tp0p = r[2] + 1j * r[3], r[4] + 1j * r[5], r[6] + 1j * r[7], r[8] + 1j * r[9]
                inta_00, inta_0p = c.dx * r[10], c.dx * r[11]
                intb_00, intb_p0 = c.dy * r[12], c.dy * r[13]
                tpsi_00 = tp00
                tpsi_p0 = tpp0 * np.exp(-1j * (0.75 * inta_00 + 0.25 * (intb_00 + inta_0p - intb_p0)))
                tpsi_pp = tppp * np.exp(-1j * (0.5 * (inta_00 + intb_p0) + 0.5 * (intb_00 + inta_0p)))
                tpsi_0p = tp0p * np.exp(-1j * (0.75 * intb_00 + 0.25 * (inta_00 + intb_p0 - inta_0p)))
                re_xy = _find_zero_line(x, y, np.real(tpsi_00), x + c.dx, y, np.real(tpsi_p0),
                                        x + c.dx, y + c.dy, np.real(tpsi_pp), x, y + c.dy, np.real(tpsi_0p))
                im_xy = _find_zero_line(x, y, np.imag(tpsi_00), x + c.dx, y, np.imag(tpsi_p0),
                                        x + c.dx, y + c.dy, np.imag(tpsi_pp), x, y + c.dy, np.imag(tpsi_0p))
                if len(re_xy) == 2 and len(im_xy) == 2:
                    ix, iy, ph = _find_intersection(re_xy[0][0], re_xy[0][1], re_xy[1][0], re_xy[1][1],
                                                    im_xy[0][0], im_xy[0][1], im_xy[1][0], im_xy[1][1])
                    if x - c.dx < ix < x + 2.0 * c.dx and y - c.dy < iy < y + 2.0 * c.dy:
                        det.append([ix, iy, np.round(v)])
            self._detected_vortices = (np.array(det, dtype=c.dtype) if det else np.zeros((0, 3), dtype=c.dtype))
            self.solver.vortices_detected = True
        d = self._detected_vortices
        return d[:, 0], d[:, 1], d[:, 2]
