"""Per-solver configuration (the reference keeps these as module globals in
svirl/config.py, shared by every GLSolver; here each GLSolver owns one)."""


class Config:
    """Geometry, precision and solver parameters of one GLSolver
    (the fields of svirl/config.py and those GLSolver.__init__ adds,
    svirl/__init__.py:67-154)."""

    def __init__(self):
        self.Nx = self.dx = self.Lx = None
        self.Ny = self.dy = self.Ly = None
        self.Nz = self.dz = self.Lz = None
        self.Nt = self.dt = self.T = None
        self.NtA = self.dtA = self.TA = None
        self.material_tiling = None
        self.order_parameter = 'random'
        self.random_seed = None
        self.random_level = 1.0
        self.normal_conductivity = 1.0
        self.linear_coefficient = 1.0
        self.homogeneous_external_field = 0.0
        self.external_field = 0.0
        self.fixed_vortices = None
        self.fixed_vortices_correction = 'cell centers'
        self.phase_lock_radius = None
        self.order_parameter_Langevin_coefficient = 0.0
        self.vector_potential_Langevin_coefficient = 0.0
        self.device_id = 0
        self.dtype = None
        self.dtype_complex = None
        self.stop_criterion_order_parameter = 1e-6
        self.stop_criterion_vector_potential = 1e-6
        self.convergence_rtol = 1e-6
        self.gl_parameter = float('inf')
