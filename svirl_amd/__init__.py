"""svirl_amd: the 2D Ginzburg-Landau solver of svirl on MI355X (gfx950).

``GLSolver`` keeps svirl's Python API (svirl/__init__.py:16-208):
``gl.vars``, ``gl.params``, ``gl.mesh``, ``gl.observables``, ``gl.solve``
(``td``, ``cg``), ``gl.vortex_detector``, ``gl.cfg``. The numerics run in
hand-written HIP kernels in libglx.so, called through ctypes. dtype float64
is the plain FP64 form; float32 is the FP32-mixed form (float fields and
stencils, FP64 reductions, residuals and line search).
"""
import numpy as np

from svirl_amd.config import Config
from svirl_amd import mesh as _mesh
from svirl_amd import observables as _obs
from svirl_amd import solvers as _solvers
from svirl_amd import vars as _vars
from svirl_amd.runtime import Reduction, Startup

__all__ = ["GLSolver"]


class GLSolver:
    """2D Ginzburg-Landau solver (time-dependent and free-energy
    minimisation) with finite or infinite GL parameter, material tiling,
    spatial linear coefficient, vortex detector and observables."""

    def __init__(self,
                 Nx=None, dx=None, Lx=None,
                 Ny=None, dy=None, Ly=None,
                 Nt=None, dt=None, T=None,
                 NtA=None, dtA=None, TA=None,
                 material_tiling=None,
                 order_parameter='random',
                 random_seed=None,
                 random_level=1.0,
                 gl_parameter=np.inf,
                 normal_conductivity=1.0,
                 linear_coefficient=1.0,
                 homogeneous_external_field=0.0,
                 external_field=0.0,
                 fixed_vortices=None,
                 fixed_vortices_correction='cell centers',
                 phase_lock_radius=None,
                 device_id=0,
                 dtype=np.float64,
                 stop_criterion_order_parameter=1e-6,
                 stop_criterion_vector_potential=1e-6,
                 order_parameter_Langevin_coefficient=0.0,
                 vector_potential_Langevin_coefficient=0.0,
                 convergence_rtol=1e-6):
        cfg = Config()
        self.dtypes = (np.float32, np.float64)
        cfg.device_id = device_id
        assert dtype in self.dtypes
        cfg.dtype = dtype
        cfg.dtype_complex = {np.float32: np.complex64, np.float64: np.complex128}[dtype]

        # geometry resolution (svirl/__init__.py:73-98)
        if Nx is not None and Lx is not None and dx is None:
            dx = float(Lx) / (Nx - 1)
        elif Nx is not None and Lx is None and dx is not None:
            Lx = float(dx) * (Nx - 1)
        elif Nx is None and Lx is not None and dx is not None:
            Nx = int(np.round(Lx / dx) + 1)
        elif Nx is not None and Lx is not None and dx is not None:
            assert np.isclose(Lx, dx * (Nx - 1))
        else:
            raise ValueError('Two out of three Nx, Lx, dx must be defined')
        if Ny is not None and Ly is not None and dy is None:
            dy = float(Ly) / (Ny - 1)
        elif Ny is not None and Ly is None and dy is not None:
            Ly = float(dy) * (Ny - 1)
        elif Ny is None and Ly is not None and dy is not None:
            Ny = int(np.round(Ly / dy) + 1)
        elif Ny is not None and Ly is not None and dy is not None:
            assert np.isclose(Ly, dy * (Ny - 1))
        else:
            raise ValueError('Two out of three Ny, Ly, and dy must be defined')
        assert isinstance(Lx, (np.floating, float, np.integer, int)) and Lx > 0.0
        assert isinstance(Ly, (np.floating, float, np.integer, int)) and Ly > 0.0
        assert isinstance(Nx, (np.integer, int)) and Nx >= 4
        assert isinstance(Ny, (np.integer, int)) and Ny >= 4

        cfg.Nx, cfg.Ny = np.int32(Nx), np.int32(Ny)
        cfg.Lx, cfg.Ly = cfg.dtype(Lx), cfg.dtype(Ly)
        cfg.dx, cfg.dy = cfg.dtype(dx), cfg.dtype(dy)
        cfg.N = int(cfg.Nx) * int(cfg.Ny)
        cfg.Nxa, cfg.Nya = cfg.Nx - 1, cfg.Ny
        cfg.Nxb, cfg.Nyb = cfg.Nx, cfg.Ny - 1
        cfg.Na, cfg.Nb = int(cfg.Nxa) * int(cfg.Nya), int(cfg.Nxb) * int(cfg.Nyb)
        cfg.Nab = cfg.Na + cfg.Nb
        cfg.Nxc, cfg.Nyc = cfg.Nx - 1, cfg.Ny - 1
        cfg.Nc = int(cfg.Nxc) * int(cfg.Nyc)
        cfg.idx, cfg.idy = 1.0 / cfg.dx, 1.0 / cfg.dy
        cfg.idx2, cfg.idy2, cfg.idxy = cfg.idx * cfg.idx, cfg.idy * cfg.idy, cfg.idx * cfg.idy
        cfg.j_dx, cfg.j_dy = 1.0j * cfg.dx, 1.0j * cfg.dy

        cfg.material_tiling = material_tiling
        cfg.order_parameter = order_parameter
        cfg.random_seed = random_seed
        cfg.random_level = random_level
        cfg.gl_parameter = gl_parameter
        cfg.linear_coefficient = linear_coefficient
        cfg.normal_conductivity = normal_conductivity
        cfg.homogeneous_external_field = homogeneous_external_field
        cfg.external_field = external_field
        cfg.fixed_vortices = fixed_vortices
        cfg.fixed_vortices_correction = fixed_vortices_correction
        cfg.phase_lock_radius = phase_lock_radius
        cfg.order_parameter_Langevin_coefficient = order_parameter_Langevin_coefficient
        cfg.vector_potential_Langevin_coefficient = vector_potential_Langevin_coefficient
        cfg.Nt, cfg.dt, cfg.T = None, None, None
        cfg.NtA, cfg.dtA, cfg.TA = None, None, None
        assert isinstance(stop_criterion_order_parameter, (np.floating, float, np.integer, int)) and stop_criterion_order_parameter > 0.0
        cfg.stop_criterion_order_parameter = cfg.dtype(stop_criterion_order_parameter)
        assert isinstance(stop_criterion_vector_potential, (np.floating, float, np.integer, int)) and stop_criterion_vector_potential > 0.0
        cfg.stop_criterion_vector_potential = cfg.dtype(stop_criterion_vector_potential)
        cfg.convergence_rtol = convergence_rtol
        self.cfg = cfg

        self.par = Startup(cfg)
        self.par.red = Reduction(self.par)
        self.mesh = _mesh.Grid(cfg)
        self.vars = _vars.Vars(self.par, self.mesh)
        self.params = _vars.Params(self.mesh, self.vars)
        self.observables = _obs.Observables(self.par, self.mesh, self.vars, self.params)
        self.solve = _solvers.Solvers(self.par, self.mesh, self.vars, self.params, self.observables)
        self.vortex_detector = _obs.VortexDetector(self.vars, self.params, self.solve)

    # flatten helpers used by the reference's tests (svirl/__init__.py:185-206)
    def flatten_a_array(self, a):
        return np.reshape(a.T, self.cfg.Na)

    def unflatten_a_array(self, a):
        return np.reshape(a, (self.cfg.Nya, self.cfg.Nxa)).T

    def flatten_b_array(self, b):
        return np.reshape(b.T, self.cfg.Nb)

    def unflatten_b_array(self, b):
        return np.reshape(b, (self.cfg.Nyb, self.cfg.Nxb)).T

    def flatten_array(self, psi):
        return np.reshape(psi.T, self.cfg.N)

    def unflatten_array(self, psi):
        return np.reshape(psi, (self.cfg.Ny, self.cfg.Nx)).T
