"""Device runtime of one GLSolver: device selection, launch constants kept
for API compatibility, and the reductions (the roles of
svirl/parallel/startup.py and svirl/parallel/reduction.py, on libglx)."""
import ctypes

import numpy as np

from svirl_amd import _native as nat
from svirl_amd._native import DeviceArray


def intceil(k, l):
    return int(np.ceil(float(k) / float(l)))


class Startup:
    """Selects the GPU (svirl/parallel/startup.py:17-67). There is no JIT:
    the kernels are compiled ahead of time for gfx950 and take the geometry
    as arguments. ``block_size``/``grid_size`` are kept for code that reads
    them; libglx picks its own launch shapes."""

    def __init__(self, cfg):
        assert isinstance(cfg.device_id, (np.integer, int)) and cfg.device_id >= 0
        nat.init(int(cfg.device_id))
        self.cfg = cfg
        self.block_size = 128
        self.grid_size = intceil(cfg.N, self.block_size)
        self.grid_size_A = intceil(cfg.Nab, self.block_size)
        self.prec = nat.prec_of(cfg.dtype)
        self.red = None

    def get_function(self, function_name):
        """Kernels are not looked up by name here; kept for the error path
        of svirl/parallel/startup.py:74-82."""
        raise ValueError("\n Function name %s not found" % function_name)


class Reduction:
    """Scalar and vector sums on the device with FP64 accumulation
    (svirl/parallel/reduction.py:10-173)."""

    def __init__(self, par):
        self.par = par

    def _prec(self, arr):
        return nat.prec_of(arr.dtype)

    def gsum(self, ga_in, ga_out=None, N=0, block_size=0, use_gpuarray_sum=False):
        """Sum of the first N elements (all when N == 0) of a real device
        array. Returns the value, or stores it in ga_out (a 1-element
        device array) when given."""
        if ga_in is None:
            return None
        if N == 0:
            N = ga_in.size
        out = ctypes.c_double()
        bs = 0 if block_size == 0 else int(64 * intceil(block_size, 64))
        nat.call("glx_sum", self._prec(ga_in), ga_in.ptr, int(N), min(bs, 1024), ctypes.byref(out))
        if ga_out is not None:
            ga_out.set(np.array([out.value], dtype=ga_out.dtype))
            return None
        return ga_in.dtype.type(out.value) if ga_in.dtype.kind == 'f' else out.value

    def gsum_v(self, ga_in, nv, ne, block_size=0):
        """Column sums of an nv x ne row-major device array (host result)."""
        if ga_in is None:
            return None
        out = (ctypes.c_double * ne)()
        bs = 0 if block_size == 0 else int(64 * intceil(block_size, 64))
        nat.call("glx_sum_v", self._prec(ga_in), ga_in.ptr, int(nv), int(ne), min(bs, 1024), out)
        return np.array(out[:], dtype=ga_in.dtype)

    # test helpers of svirl/parallel/reduction.py:178-198
    def test_sum_v(self, a_in, nv, ne, block_size=256):
        assert 0 <= block_size <= 1024
        g = DeviceArray.from_host(np.asarray(a_in, dtype=self.par.cfg.dtype))
        return self.gsum_v(g, nv, ne, block_size=block_size)

    def test_sum(self, a_in, N, block_size=256):
        assert 0 <= block_size <= 1024
        g = DeviceArray.from_host(np.asarray(a_in, dtype=self.par.cfg.dtype))
        r = self.gsum(g, block_size=block_size)
        g.free()
        return r
