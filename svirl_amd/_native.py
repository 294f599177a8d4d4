"""ctypes binding of libglx.so (include/glx.h) and a device-array handle.

The library is loaded from svirl_amd/_lib/libglx.so; importing anything that
needs it fails loudly when it is missing (there is no CPU fallback).
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("GLX_LIB", os.path.join(_HERE, "_lib", "libglx.so"))

GLX_F64, GLX_F32 = 0, 1
GLX_E_ARG = -1       # error codes of include/glx.h
GLX_E_HIP = -2
GLX_E_NOINIT = -3
GLX_E_TIMEOUT = -4
GLX_E_CAPACITY = -5

c_void_p = ctypes.c_void_p
c_int = ctypes.c_int
c_i32 = ctypes.c_int32
c_u32 = ctypes.c_uint32
c_i64 = ctypes.c_int64
c_double = ctypes.c_double
c_size_t = ctypes.c_size_t
P_i32 = ctypes.POINTER(ctypes.c_int32)
P_double = ctypes.POINTER(ctypes.c_double)


class Grid(ctypes.Structure):
    """glx_grid: Nx, Ny, dx, dy."""
    _fields_ = [("Nx", c_i32), ("Ny", c_i32), ("dx", c_double), ("dy", c_double)]


class TDParams(ctypes.Structure):
    """glx_td_params."""
    _fields_ = [
        ("dt", c_double), ("eps", c_double), ("kappa2", c_double), ("rho", c_double), ("H", c_double),
        ("stop_psi", c_double), ("stop_A", c_double), ("langevin_psi", c_double), ("langevin_A", c_double),
        ("rand_t", c_u32), ("solveA", c_i32), ("max_sweeps", c_i32), ("eqn", c_i32), ("a_src", c_void_p),
    ]


class BatchInstance(ctypes.Structure):
    """glx_batch_instance."""
    _fields_ = [("g", Grid), ("p", TDParams), ("eps_map", c_void_p), ("mt", c_void_p),
                ("psi", c_void_p * 3), ("ab", c_void_p * 3), ("psi_cur", c_i32), ("ab_cur", c_i32)]


class BatchCGInstance(ctypes.Structure):
    """glx_batch_cg_instance."""
    _fields_ = [("g", Grid), ("kappa2", c_double), ("eps", c_double), ("H", c_double), ("eps_map", c_void_p),
                ("mt", c_void_p), ("psi", c_void_p), ("ab", c_void_p), ("iters", c_i32), ("converged", c_i32),
                ("grad_rel", c_double), ("grad0", c_double), ("stalled", c_i32), ("fp64_iters", c_i32)]


# name -> (argtypes); every function returns int
_PROTOS = {
    "glx_init": [c_int],
    "glx_synchronize": [],
    "glx_device_info": [P_i32, P_i32, ctypes.c_char_p, c_int],
    "glx_malloc": [ctypes.POINTER(c_void_p), c_size_t],
    "glx_free": [c_void_p],
    "glx_memcpy_h2d": [c_void_p, c_void_p, c_size_t],
    "glx_memcpy_d2h": [c_void_p, c_void_p, c_size_t],
    "glx_memcpy_d2d": [c_void_p, c_void_p, c_size_t],
    "glx_memset": [c_void_p, c_int, c_size_t],
    "glx_timer_start": [],
    "glx_timer_stop": [ctypes.POINTER(ctypes.c_float)],
    "glx_td_psi_sweep": [c_int, Grid, c_double, c_double, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                         c_void_p, c_double, c_u32, c_u32, c_double, c_void_p],
    "glx_td_a_sweep": [c_int, Grid, c_double, c_double, c_double, c_double, c_void_p, c_void_p, c_void_p,
                       c_void_p, c_void_p, c_void_p, c_double, c_u32, c_u32, c_double, c_void_p],
    "glx_td_phase_lock": [c_int, c_void_p, c_i32, c_void_p],
    "glx_td_run": [c_int, Grid, ctypes.POINTER(TDParams), c_void_p, c_void_p, ctypes.POINTER(c_void_p), P_i32,
                   ctypes.POINTER(c_void_p), P_i32, c_void_p, c_void_p, c_i32, P_i32, c_i32, c_i32],
    "glx_td_max_grid": [c_int, c_i32, P_i32],
    "glx_td_psi_solve": [c_int, Grid, ctypes.POINTER(TDParams), c_void_p, c_void_p, c_void_p,
                         ctypes.POINTER(c_void_p), P_i32, c_void_p, P_i32],
    "glx_td_a_solve": [c_int, Grid, ctypes.POINTER(TDParams), c_void_p, c_void_p, ctypes.POINTER(c_void_p), P_i32,
                       c_void_p, P_i32],
    "glx_td_step": [c_int, Grid, ctypes.POINTER(TDParams), c_void_p, c_void_p, ctypes.POINTER(c_void_p), P_i32,
                    ctypes.POINTER(c_void_p), P_i32, c_void_p, c_void_p, P_i32],
    "glx_td_last_kernel_ms": [ctypes.POINTER(ctypes.c_float)],
    "glx_td_profile": [c_i32],
    "glx_td_force_barrier_abort": [],
    "glx_batch_td_run": [c_int, ctypes.POINTER(BatchInstance), c_i32, c_i32, P_i32, c_i32, c_i32, c_i32, c_void_p,
                         c_i32],
    "glx_batch_td_run_noalloc": [c_int, ctypes.POINTER(BatchInstance), c_i32, c_i32, P_i32, c_i32, c_i32, c_void_p,
                                 c_i32],
    "glx_batch_cg_run": [c_int, ctypes.POINTER(BatchCGInstance), c_i32, c_i32, c_double, c_i32, c_i32, c_i32],
    "glx_td_solve_ms": [ctypes.POINTER(c_double), P_i32, ctypes.POINTER(c_double), P_i32],
    "glx_td_passes": [P_i32, P_i32, P_i32, P_i32, P_i32],
    "glx_stream_probe": [c_i32, c_i64, c_i32, c_i32, ctypes.POINTER(c_double)],
    "glx_magnetic_field": [c_int, Grid, c_void_p, c_void_p, c_void_p],
    "glx_current_density": [c_int, Grid, c_double, c_double, c_void_p, c_void_p, c_void_p],
    "glx_supercurrent_density": [c_int, Grid, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p],
    "glx_free_energy": [c_int, Grid, c_double, c_double, c_void_p, c_double, c_void_p, c_void_p, c_void_p,
                        c_void_p, P_double],
    "glx_vortex_cells": [c_int, Grid, c_void_p, c_void_p, c_void_p, c_double, c_void_p, c_i32, P_i32],
    "glx_vortex_triangulate": [Grid, c_void_p, c_i32, c_void_p],
    "glx_cg_direction_step": [c_int, Grid, c_double, c_double, c_double] + [c_void_p] * 13,
    "glx_cg_update_gradient_pass": [c_int, Grid, c_double, c_double, c_void_p, c_double] + [c_void_p] * 15,
    "glx_cg_energy_wait": [P_double],
    "glx_bulk_observables": [c_int, Grid, c_double, c_double, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                             c_void_p, c_void_p],
    "glx_sum": [c_int, c_void_p, c_i64, c_i32, P_double],
    "glx_sum_v": [c_int, c_void_p, c_i64, c_i32, c_i32, P_double],
    "glx_pr_sums": [c_int, c_void_p, c_void_p, c_i64, P_double],
    "glx_cg_jacobian_psi": [c_int, Grid, c_double, c_double, c_void_p, c_double, c_void_p, c_void_p, c_void_p,
                            c_void_p, c_void_p],
    "glx_cg_jacobian_A": [c_int, Grid, c_double, c_double, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p],
    "glx_cg_coef_psi": [c_int, Grid, c_double, c_double, c_double, c_void_p, c_void_p, c_void_p, c_void_p,
                        c_void_p, P_double],
    "glx_cg_coef": [c_int, Grid, c_double, c_double, c_double, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                    c_void_p, P_double],
    "glx_axpy": [c_int, c_void_p, c_void_p, c_void_p, c_double, c_i64],
    "glx_axmy": [c_int, c_void_p, c_void_p, c_void_p, c_double, c_i64],
    "glx_cg_pr_beta": [c_int, c_void_p, c_void_p, c_i64, c_void_p],
    "glx_cg_gradient_pass": [c_int, Grid, c_double, c_double, c_void_p, c_double, c_void_p, c_void_p, c_void_p,
                             c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, ctypes.POINTER(c_double)],
    "glx_axmy_dev": [c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_i64],
    "glx_cg_line_step": [c_int, Grid, c_double, c_double, c_double, c_void_p, c_void_p, c_void_p, c_void_p,
                         c_void_p, c_void_p, c_void_p],
}

EXPORTS = tuple(_PROTOS) + ("glx_last_error",)

_lib = None


class GLXError(RuntimeError):
    """A libglx entry point returned an error code."""

    def __init__(self, name, code, message):
        super().__init__(f"{name} failed ({code}): {message}")
        self.code = code


def load():
    """Load libglx.so and declare its prototypes (no device work)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"libglx.so not found at {LIB_PATH}; build it with `python -m svirl_amd._build` "
                          "(hipcc --offload-arch=gfx950)")
    lib = ctypes.CDLL(LIB_PATH)
    for name, args in _PROTOS.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = c_int
    lib.glx_last_error.argtypes = []
    lib.glx_last_error.restype = ctypes.c_char_p
    _lib = lib
    return lib


def call(name, *args):
    """Call a libglx entry point and raise GLXError on a nonzero status."""
    lib = load()
    rc = getattr(lib, name)(*args)
    if rc != 0:
        msg = lib.glx_last_error().decode(errors="replace")
        raise GLXError(name, rc, msg)
    return rc


_initialised_device = None


def init(device=0):
    """Select the GPU and create the library stream (once per process)."""
    global _initialised_device
    if _initialised_device is None:
        call("glx_init", int(device))
        _initialised_device = int(device)
    elif _initialised_device != int(device):
        raise GLXError("glx_init", -1, f"already initialised on device {_initialised_device}")
    return _initialised_device


def device_info():
    n = c_i32(0)
    w = c_i32(0)
    buf = ctypes.create_string_buffer(256)
    call("glx_device_info", ctypes.byref(n), ctypes.byref(w), buf, 256)
    return {"num_cus": n.value, "wg_per_cu_hint": w.value, "name": buf.value.decode()}


def prec_of(dtype):
    """GLX precision code of a real or complex numpy dtype."""
    dt = np.dtype(dtype)
    if dt in (np.float64, np.complex128):
        return GLX_F64
    if dt in (np.float32, np.complex64):
        return GLX_F32
    raise TypeError(f"unsupported dtype {dt}")


def ptr(x):
    """Device address of a DeviceArray (None -> NULL)."""
    if x is None:
        return None
    return x.ptr


class DeviceArray:
    """A flat device buffer owned by Python (allocated through libglx).

    Mirrors the part of pycuda's GPUArray that svirl's host code and tests
    use: ``size``, ``dtype``, ``nbytes``, ``get()``, ``fill()``, ``copy()``.
    """

    def __init__(self, size, dtype):
        init()
        self.dtype = np.dtype(dtype)
        self.size = int(size)
        self.nbytes = self.size * self.dtype.itemsize
        p = c_void_p()
        call("glx_malloc", ctypes.byref(p), max(self.nbytes, 1))
        self.ptr = p.value

    @classmethod
    def from_host(cls, arr, dtype=None):
        a = np.ascontiguousarray(arr, dtype=dtype if dtype is not None else np.asarray(arr).dtype).ravel()
        d = cls(a.size, a.dtype)
        d.set(a)
        return d

    @classmethod
    def zeros(cls, size, dtype):
        d = cls(size, dtype)
        d.fill(0)
        return d

    def set(self, arr):
        a = np.ascontiguousarray(arr, dtype=self.dtype).ravel()
        if a.size != self.size:
            raise ValueError(f"size mismatch: device {self.size}, host {a.size}")
        call("glx_memcpy_h2d", self.ptr, a.ctypes.data, self.nbytes)

    def get(self, count=None):
        """The buffer on the host (its first `count` elements if given)."""
        n = self.size if count is None else int(count)
        if n < 0 or n > self.size:
            raise ValueError(f"count {n} outside [0, {self.size}]")
        out = np.empty(n, dtype=self.dtype)
        if n:
            call("glx_memcpy_d2h", out.ctypes.data, self.ptr, n * self.dtype.itemsize)
        return out

    def fill(self, value):
        if value == 0:
            call("glx_memset", self.ptr, 0, self.nbytes)
        else:
            self.set(np.full(self.size, value, dtype=self.dtype))

    def copy(self):
        d = DeviceArray(self.size, self.dtype)
        call("glx_memcpy_d2d", d.ptr, self.ptr, self.nbytes)
        return d

    def copy_from(self, other):
        if other.nbytes != self.nbytes:
            raise ValueError("copy_from: size mismatch")
        call("glx_memcpy_d2d", self.ptr, other.ptr, self.nbytes)

    @property
    def gpudata(self):
        return self

    def free(self):
        if self.ptr:
            call("glx_free", self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            if getattr(self, "ptr", None) and _lib is not None:
                _lib.glx_free(self.ptr)
        except Exception:
            pass

    def __repr__(self):
        return f"DeviceArray(size={self.size}, dtype={self.dtype}, ptr=0x{self.ptr or 0:x})"
