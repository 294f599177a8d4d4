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
        """The reference's kernel lookup (svirl/parallel/startup.py:74-82):
        a callable taking that kernel's arguments in its order (device
        buffers as DeviceArray / GArray / addresses, scalars as numbers;
        ``block`` and ``grid`` are accepted and ignored, libglx picks its
        own launch shapes), running the libglx entry point that computes the
        same thing. Kernels whose output is per-block partial sums laid out
        by the launch shape (sum, sum_v, x_mag2_*_sum, xmyx_*_sum) or
        per-node coefficient terms (free_energy_pseudodensity,
        free_energy_conjgrad_coef*) have no such counterpart: asking for one
        raises ValueError naming the whole-array entry point to use, as an
        unknown name does."""
        table = _kernel_table(self)
        if function_name in table:
            return _Kernel(function_name, table[function_name])
        if function_name in _WHOLE_ARRAY:
            raise ValueError("\n Function name %s not found: its result depends on the launch shape; use %s"
                             % (function_name, _WHOLE_ARRAY[function_name]))
        raise ValueError("\n Function name %s not found" % function_name)


def _addr(x):
    """Device address of a kernel argument (None / 0 -> NULL)."""
    if x is None:
        return None
    if hasattr(x, "get_d_obj"):
        return _addr(x.get_d_obj())
    if hasattr(x, "ptr"):
        return x.ptr
    v = int(x)
    return v or None


class _Kernel:
    """A named kernel of the table below, called like a compiled one."""

    def __init__(self, name, fn):
        self.name = name
        self._fn = fn

    def __call__(self, *args, block=None, grid=None, **kw):
        return self._fn(*args)


def _kernel_table(par):
    cfg = par.cfg
    prec = par.prec
    rdt = np.dtype(cfg.dtype)

    def grid_():
        return nat.Grid(int(cfg.Nx), int(cfg.Ny), float(cfg.dx), float(cfg.dy))

    def scalar_at(ptr):
        out = np.zeros(1, dtype=rdt)
        nat.call("glx_memcpy_d2h", out.ctypes.data, ptr, rdt.itemsize)
        return float(out[0])

    def axpy(n_per):
        return lambda x, y, z, alpha, N: nat.call("glx_axpy", prec, _addr(x), _addr(y), _addr(z), float(alpha),
                                                  n_per * int(N))

    def axmy(n_per):
        return lambda x, y, z, alpha, N: nat.call("glx_axmy", prec, _addr(x), _addr(y), _addr(z),
                                                  scalar_at(_addr(alpha)), n_per * int(N))

    def xpy(x, y, N):  # x += y (utils.h:118-125)
        nat.call("glx_axpy", prec, _addr(y), _addr(x), _addr(x), 1.0, int(N))

    def xmy(x, y, N):  # x -= y (utils.h:128-135): x <- (-1) y + x
        nat.call("glx_axpy", prec, _addr(y), _addr(x), _addr(x), -1.0, int(N))

    def psi_step(dt, eps, eps_map, mt, ab_abi, psi_rhs, psi, psi_next, langevin_c, jstep, rand_t, stop_eps, r2max):
        nat.call("glx_td_psi_sweep", prec, grid_(), float(dt), float(eps), _addr(eps_map), _addr(mt), _addr(ab_abi),
                 _addr(psi_rhs), _addr(psi), _addr(psi_next), float(langevin_c), int(jstep), int(rand_t),
                 float(stop_eps), _addr(r2max))

    def a_step(dt, kappa2, rho, H, mt, psi, abi_ab_rhs, ab_rhs, ab, ab_next, langevin_c, jstep, rand_t, stop_eps,
               r2max):
        nat.call("glx_td_a_sweep", prec, grid_(), float(dt), float(kappa2), float(rho), float(H), _addr(mt),
                 _addr(psi), _addr(abi_ab_rhs), _addr(ab_rhs), _addr(ab), _addr(ab_next), float(langevin_c),
                 int(jstep), int(rand_t), float(stop_eps), _addr(r2max))

    return {
        "axpy_r": axpy(1), "axpy_c": axpy(2), "axmy_r": axmy(1), "axmy_c": axmy(2),
        "xpy_r": xpy, "xmy_r": xmy,
        "iterate_order_parameter_jacobi_step": psi_step,
        "iterate_vector_potential_jacobi_step": a_step,
        "order_parameter_phase_lock": lambda psi, n, ns: nat.call("glx_td_phase_lock", prec, _addr(psi), int(n),
                                                                  _addr(ns)),
        "magnetic_field": lambda abei, ab, B: nat.call("glx_magnetic_field", prec, grid_(), _addr(abei), _addr(ab),
                                                       _addr(B)),
        "current_density": lambda kappa2, H, abei, ab, J: nat.call("glx_current_density", prec, grid_(),
                                                                   float(kappa2), float(H), _addr(abei), _addr(ab),
                                                                   _addr(J)),
        "supercurrent_density": lambda mt, psi, abei, ab, Js: nat.call("glx_supercurrent_density", prec, grid_(),
                                                                       _addr(mt), _addr(psi), _addr(abei),
                                                                       _addr(ab), _addr(Js)),
        "free_energy_jacobian_psi": lambda kappa2, eps, eps_map, H, mt, psi, abei, ab, G: nat.call(
            "glx_cg_jacobian_psi", prec, grid_(), float(kappa2), float(eps), _addr(eps_map), float(H), _addr(mt),
            _addr(psi), _addr(abei), _addr(ab), _addr(G)),
        "free_energy_jacobian_A": lambda kappa2, H, mt, psi, abei, ab, G: nat.call(
            "glx_cg_jacobian_A", prec, grid_(), float(kappa2), float(H), _addr(mt), _addr(psi), _addr(abei),
            _addr(ab), _addr(G)),
    }


# Reference kernels whose output layout follows the launch shape, and the
# whole-array entry point that gives their result.
_WHOLE_ARRAY = {
    "sum": "Reduction.gsum (glx_sum)", "sum_v": "Reduction.gsum_v (glx_sum_v)",
    "x_mag2_c_sum": "glx_pr_sums", "x_mag2_r_sum": "glx_pr_sums", "xmyx_c_sum": "glx_pr_sums",
    "xmyx_r_sum": "glx_pr_sums", "divide_scalars_positive": "glx_cg_pr_beta",
    "free_energy_pseudodensity": "glx_free_energy", "free_energy_conjgrad_coef_psi": "glx_cg_coef_psi",
    "free_energy_conjgrad_coef": "glx_cg_coef",
}


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
