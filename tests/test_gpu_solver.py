"""End-to-end parity of the drop-in GLSolver (GPU) with the FP64 oracle,
plus the reference's own tests (tests/at_*.py) re-expressed on the GPU.

Inputs are built once by the GLSolver (seeded random psi, symmetric gauge)
and handed to the oracle as the same flat arrays, so both start from
identical bits. FP64: psi and F within 1e-10 relative, vortex cells and
vorticities exact, Jacobi sweep counts exact.
"""
import numpy as np
import pytest

from oracle import gl_oracle as O

pytestmark = pytest.mark.gpu

svirl_amd = pytest.importorskip("svirl_amd")
GLSolver = svirl_amd.GLSolver


def oracle_of(gl):
    c = gl.cfg
    g = O.Geom(int(c.Nx), int(c.Ny), float(c.dx), float(c.dy))
    psi = gl.flatten_array(gl.vars.order_parameter).astype(complex)
    a, b = gl.vars.vector_potential
    ab = np.concatenate([gl.flatten_a_array(a), gl.flatten_b_array(b)]).astype(float)
    mt = None
    if gl.mesh.have_material_tiling():
        mt = np.ascontiguousarray(gl.mesh.material_tiling.T).ravel()
    em = None
    if gl.params.linear_coefficient_h() is not None:
        em = gl.flatten_array(gl.params.linear_coefficient).astype(float)
    abei = None
    if gl.params.external_irregular_vector_potential_h() is not None:
        ea, eb = gl.params.external_irregular_vector_potential
        abei = np.concatenate([gl.flatten_a_array(ea), gl.flatten_b_array(eb)]).astype(float)
    eps = float(gl.params.linear_coefficient_scalar_h()) if em is None else 1.0
    fv = gl.params.fixed_vortices
    vpi = None
    if fv._vpi is not None:
        ia, ib = fv.irregular_vector_potential
        vpi = np.concatenate([gl.flatten_a_array(ia), gl.flatten_b_array(ib)]).astype(float)
    st = O.OracleState(g, psi, ab, kappa=float(gl.params.gl_parameter), sigma=float(gl.params.normal_conductivity),
                       H=float(gl.params.homogeneous_external_field), eps=eps, eps_map=em, mt=mt, abei=abei,
                       stop_psi=float(c.stop_criterion_order_parameter), stop_A=float(c.stop_criterion_vector_potential),
                       random_t=1 if c.random_seed is None else int(c.random_seed),
                       convergence_rtol=c.convergence_rtol, vpi=vpi, lock_ns=fv._phase_lock_ns)
    return st


def gl_psi(gl):
    return gl.flatten_array(gl.vars.order_parameter)


def gl_ab(gl):
    a, b = gl.vars.vector_potential
    return np.concatenate([gl.flatten_a_array(a), gl.flatten_b_array(b)])


def rel(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


def test_c1_like_td_and_cg_kappa_inf():
    """Config 1 in miniature: kappa = inf, H = 0.1, random seeded psi, TD
    then CG (polyroots line search)."""
    gl = GLSolver(Lx=16, Ly=16, dx=0.5, dy=0.5, order_parameter='random', random_seed=1, random_level=1.0,
                  gl_parameter=np.inf, homogeneous_external_field=0.1, dtype=np.float64)
    st = oracle_of(gl)
    gl.solve.td(dt=0.1, Nt=60)
    st.td(dt=0.1, Nt=60)
    assert gl.solve._td.sweeps == st.sweeps
    assert rel(gl_psi(gl), st.psi) < 1e-10
    F = gl.observables.free_energy
    assert F == pytest.approx(st.free_energy(), rel=1e-10)
    gl.solve.cg(n_iter=40)
    st.cg(n_iter=40)
    assert len(gl.solve._cg.cg_energies) == len(st.cg_energies)
    assert np.allclose(gl.solve._cg.cg_energies, st.cg_energies, rtol=1e-10)
    assert rel(gl_psi(gl), st.psi) < 1e-9


def test_c2_like_td_finite_kappa_with_mask_and_eps_map():
    rng = np.random.default_rng(3)
    gl = GLSolver(Lx=20, Ly=16, dx=0.5, dy=0.5, order_parameter='random', random_seed=2, gl_parameter=5.0,
                  normal_conductivity=200.0, homogeneous_external_field=0.1, dtype=np.float64)
    mt = np.ones((gl.cfg.Nxc, gl.cfg.Nyc), bool)
    mt[10:16, 8:14] = False
    gl.mesh.material_tiling = mt
    gl.vars.set_order_parameter_to_zero_outside_material()
    gl.params.linear_coefficient = 1.0 + 0.2 * rng.normal(size=(gl.cfg.Nx, gl.cfg.Ny))
    st = oracle_of(gl)
    gl.solve.td(dt=0.1, Nt=40)
    st.td(dt=0.1, Nt=40)
    assert gl.solve._td.sweeps == st.sweeps
    assert rel(gl_psi(gl), st.psi) < 1e-10
    assert rel(gl_ab(gl), st.ab) < 1e-10
    assert gl.observables.free_energy == pytest.approx(st.free_energy(), rel=1e-10)
    jx, jy = gl.observables.current_density
    J = st.current_density()
    assert np.allclose(np.concatenate([gl.flatten_a_array(jx), gl.flatten_b_array(jy)]), J, atol=1e-9)
    # CG with the BFGS line search
    gl.solve.cg(n_iter=15)
    st.cg(n_iter=15)
    assert np.allclose(gl.solve._cg.cg_energies, st.cg_energies, rtol=1e-9)


def test_large_grid_td_matches_oracle():
    """Config 3's physics on a 769 x 641 grid (holes, pinning map, kappa = 5,
    H = 0.05, dt = 0.05) in FP64: long row ranges per wave, fused passes
    and a team of several dozen workgroups; sweep counts exact, fields
    within 1e-10 of the oracle after three steps."""
    from svirl_amd import workloads
    gl = GLSolver(Nx=769, Ny=641, dx=0.5, dy=0.5, order_parameter='random', random_seed=3, gl_parameter=5.0,
                  normal_conductivity=200.0, homogeneous_external_field=0.05, dtype=np.float64)
    c = gl.cfg
    gl.mesh.material_tiling = workloads.holes_mask(int(c.Nxc), int(c.Nyc), 0.5, 0.5, seed=5)
    gl.vars.set_order_parameter_to_zero_outside_material()
    gl.params.linear_coefficient = workloads.pinning_field(int(c.Nx), int(c.Ny), 0.5, seed=6)
    st = oracle_of(gl)
    gl.solve.td(dt=0.05, Nt=3)
    st.td(dt=0.05, Nt=3)
    assert gl.solve._td._default_wg() > 16
    assert gl.solve._td.sweeps == st.sweeps
    assert rel(gl_psi(gl), st.psi) < 1e-10 and rel(gl_ab(gl), st.ab) < 1e-10
    assert gl.observables.free_energy == pytest.approx(st.free_energy(), rel=1e-10)


@pytest.mark.parametrize("shape", [(4, 4), (4, 5), (5, 4)])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_smallest_grids_match_oracle(shape, dtype):
    """The smallest grids GLSolver takes (4 nodes a side, as the
    reference's constructor asserts, __init__.py:97-98): TD with A and CG
    against the oracle; FP64 to 1e-10 with exact sweep counts, FP32-mixed
    to float tolerance."""
    Nx, Ny = shape
    gl = GLSolver(Nx=Nx, Ny=Ny, dx=0.5, dy=0.5, order_parameter='random', random_seed=1, gl_parameter=2.0,
                  normal_conductivity=10.0, homogeneous_external_field=0.2, dtype=dtype)
    st = oracle_of(gl)
    gl.solve.td(dt=0.1, Nt=3)
    st.td(dt=0.1, Nt=3)
    f64 = dtype == np.float64
    tol = 1e-10 if f64 else 1e-4
    if f64:
        assert gl.solve._td.sweeps == st.sweeps
    assert rel(gl_psi(gl), st.psi) < tol and rel(gl_ab(gl), st.ab) < tol
    assert gl.observables.free_energy == pytest.approx(st.free_energy(), rel=tol, abs=tol)
    gl.solve.cg(n_iter=4)
    st.cg(n_iter=4)
    assert np.allclose(gl.solve._cg.cg_energies, st.cg_energies, rtol=1e-8 if f64 else 1e-4)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_td_langevin_matches_oracle(dtype):
    """Langevin noise on psi and A (td.py:118-154, 256-325) through the
    persistent solves, on a grid of several strips with holes: the first
    pass of each psi solve draws the noise and records the noisy right-hand
    side, the later passes read it. The first step's first pass is a full
    pass (no earlier sweep count to shorten it), so FP32 runs the Langevin
    copy of the four-sweep pass loop there, and the later steps' shortened
    first passes (21-27 sweeps per step) the one-, two- and three-sweep
    forms. Against the oracle over 4
    steps: FP64 sweep counts exact and fields within 1e-10, FP32 within
    float tolerance; with the noise off the fields differ."""
    def make():
        gl = GLSolver(Nx=130, Ny=97, dx=0.5, dy=0.5, order_parameter='random', random_seed=7, gl_parameter=5.0,
                      normal_conductivity=100.0, homogeneous_external_field=0.1, dtype=dtype)
        mt = np.ones((gl.cfg.Nxc, gl.cfg.Nyc), bool)
        mt[40:52, 30:41] = False
        gl.mesh.material_tiling = mt
        gl.vars.set_order_parameter_to_zero_outside_material()
        return gl
    gl = make()
    gl.params.order_parameter_Langevin_coefficient = 0.05
    gl.params.vector_potential_Langevin_coefficient = 0.02
    st = oracle_of(gl)
    st.langevin_psi, st.langevin_A = 0.05, 0.02
    gl.solve.td(dt=0.1, Nt=4)
    st.td(dt=0.1, Nt=4)
    f64 = dtype == np.float64
    tol = 1e-10 if f64 else 1e-4
    if f64:
        assert gl.solve._td.sweeps == st.sweeps
    assert rel(gl_psi(gl), st.psi) < tol and rel(gl_ab(gl), st.ab) < tol
    quiet = make()
    quiet.solve.td(dt=0.1, Nt=4)
    assert rel(gl_psi(quiet), st.psi) > 100 * tol


def test_fp32_large_grid_close_to_fp64():
    """The FP32-mixed form (gauge-shifted A, fused passes) against the FP64
    GPU form on Config 3's physics at 769 x 641 after 20 steps: free
    energy within 1e-6, psi and J within float-level L2 errors."""
    from svirl_amd import workloads

    def make(dt_):
        gl = GLSolver(Nx=769, Ny=641, dx=0.5, dy=0.5, order_parameter='random', random_seed=3, gl_parameter=5.0,
                      normal_conductivity=200.0, homogeneous_external_field=0.05, dtype=dt_)
        c = gl.cfg
        gl.mesh.material_tiling = workloads.holes_mask(int(c.Nxc), int(c.Nyc), 0.5, 0.5, seed=5)
        gl.vars.set_order_parameter_to_zero_outside_material()
        gl.params.linear_coefficient = workloads.pinning_field(int(c.Nx), int(c.Ny), 0.5, seed=6)
        gl.solve.td(dt=0.05, Nt=20)
        return gl
    a, b = make(np.float32), make(np.float64)
    fa, fb = a.observables.free_energy_fp64(), b.observables.free_energy_fp64()
    assert abs(fa - fb) <= 1e-6 * abs(fb)
    pa, pb = np.asarray(gl_psi(a), np.complex128), np.asarray(gl_psi(b), np.complex128)
    assert np.linalg.norm(pa - pb) <= 1e-4 * np.linalg.norm(pb)
    ja = np.concatenate([np.ravel(x) for x in a.observables.current_density]).astype(np.float64)
    jb = np.concatenate([np.ravel(x) for x in b.observables.current_density]).astype(np.float64)
    assert np.linalg.norm(ja - jb) <= 1e-2 * np.linalg.norm(jb)


def test_vortex_detector_matches_oracle():
    gl = GLSolver(Lx=24, Ly=24, dx=0.5, dy=0.5, order_parameter=1.0, gl_parameter=np.inf,
                  homogeneous_external_field=0.15, dtype=np.float64)
    gl.params.fixed_vortices.order_parameter_add_vortices([[6.2, 12.3, 17.1], [7.1, 15.2, 5.3], [1, -1, 1]],
                                                          phase=True, deep=True)
    gl.solve.td(dt=0.1, Nt=30)
    st = oracle_of(gl)
    vx, vy, vv = gl.vortex_detector.vortices
    ox, oy, ov = st.vortices()
    cells, vs = O.vortex_cells(st.g, st.psi, st.ab, st.H)
    gcells, gvs = gl.vortex_detector.vortex_cells
    assert list(gcells) == list(cells) and np.array_equal(np.round(gvs), np.round(vs))
    assert len(vx) == len(ox) and np.array_equal(vv, ov)
    assert np.allclose(vx, ox, atol=1e-10) and np.allclose(vy, oy, atol=1e-10)


@pytest.mark.parametrize("kappa", [4.0, np.inf])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_bulk_observables_through_api(kappa, dtype):
    """Observables.bulk_observables() and the fused normalcurrent_density
    return what the separate properties return, to the bit; with kappa =
    inf the current density is the supercurrent (observables.py:86-110)."""
    gl = GLSolver(Lx=14, Ly=11, dx=0.5, dy=0.5, order_parameter='random', random_seed=4, gl_parameter=kappa,
                  normal_conductivity=100.0, homogeneous_external_field=0.12, dtype=dtype)
    mt = np.ones((gl.cfg.Nxc, gl.cfg.Nyc), bool)
    mt[5:9, 4:8] = False
    gl.mesh.material_tiling = mt
    gl.solve.td(dt=0.1, Nt=5)
    B, (jx, jy), (sx, sy) = gl.observables.bulk_observables()
    assert np.array_equal(B, gl.observables.magnetic_field)
    Jx, Jy = gl.observables.current_density
    Sx, Sy = gl.observables.supercurrent_density
    assert np.array_equal(jx, Jx) and np.array_equal(jy, Jy)
    assert np.array_equal(sx, Sx) and np.array_equal(sy, Sy)
    nx, ny = gl.observables.normalcurrent_density
    assert np.array_equal(nx, Jx - Sx) and np.array_equal(ny, Jy - Sy)
    if np.isinf(kappa):
        assert np.array_equal(jx, sx) and not nx.any()


def _vpi_of(gl):
    ia, ib = gl.params.fixed_vortices.irregular_vector_potential
    return np.concatenate([gl.flatten_a_array(ia), gl.flatten_b_array(ib)])


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_td_fixed_vortices_and_phase_lock(dtype):
    """examples/_example_fixed_vortices.py scaled down: pinned vortices (an
    irregular vector potential vpi) and a phase lock, through all three eqn
    modes, against the oracle's transcription of the reference's per-step
    hooks (td.py:118-154, 205-231, 256-325), including the drift of vpi
    and, with eqn='order_parameter', of A. FP64: sweep counts exact, fields
    within 1e-10; FP32: within float tolerance."""
    gl = GLSolver(Lx=16, Ly=14, dx=0.5, dy=0.5, order_parameter='random', random_level=0.1, random_seed=2,
                  linear_coefficient=1.0, gl_parameter=1.0, normal_conductivity=10.0,
                  fixed_vortices=(np.array([4.2, 4.2, 11.7, 11.7]), np.array([4.1, 9.6, 4.1, 9.6]), [1, -1, -1, 1]),
                  phase_lock_radius=0.8, dtype=dtype)
    assert gl.params.fixed_vortices._phase_lock_ns is not None
    st = oracle_of(gl)
    f64 = dtype == np.float64
    tol = 1e-10 if f64 else 2e-4

    def check():
        if f64:
            assert gl.solve._td.sweeps == st.sweeps
        assert rel(gl_psi(gl), st.psi) < tol
        assert rel(gl_ab(gl), st.ab) < tol
        assert rel(_vpi_of(gl), st.vpi) < tol

    gl.solve.td(dt=0.1, Nt=6)
    st.td(dt=0.1, Nt=6)
    check()
    gl.solve.td(dt=0.1, Nt=3, eqn='order_parameter')
    st.td_order_parameter(dt=0.1, Nt=3)
    check()
    gl.solve.td(dt=0.1, Nt=3, eqn='vector_potential')
    st.td_vector_potential(dt=0.1, Nt=3)
    check()
    ns = gl.params.fixed_vortices._phase_lock_ns
    p = gl_psi(gl)[ns]
    assert np.all(p.imag == 0) and np.all(p.real >= 0)
    if f64:
        # the CG sees the pinned vortices through abei = vpi (as built at
        # construction, params.py:193-221) and ignores the lock (cg.py)
        st2 = oracle_of(gl)
        assert st2.abei is not None and np.any(st2.abei != 0)
        gl.solve.cg(n_iter=8)
        st2.cg(n_iter=8)
        assert np.allclose(gl.solve._cg.cg_energies, st2.cg_energies, rtol=1e-9)
        assert rel(gl_psi(gl), st2.psi) < 1e-8


def test_jacobians_finite_differences_through_api():
    """tests/at_cg_jacobians.py:111-160 on the drop-in (one-sided h=3e-9,
    atol 1e-5, rtol 1e-3 as in the reference)."""
    rng = np.random.default_rng(11)
    gl = GLSolver(Nx=9, Ny=10, dx=0.45, dy=0.48, gl_parameter=1.0, dtype=np.float64)
    gl.solve._init_cg()
    gl.params.gl_parameter = 1.0 + 3.0 * rng.random()
    gl.vars.order_parameter = 1.0
    gl.vars.randomize_order_parameter(level=0.5, seed=4)
    gl.params.homogeneous_external_field_reset = 0.01 + 0.1 * rng.random()
    gl.params.external_field = 0.01 + 0.1 * rng.random()
    h = 3e-9
    E0 = gl.observables.free_energy
    G1 = gl.unflatten_array(gl.solve._cg._free_energy_jacobian_psi.get())
    G = np.empty((gl.cfg.Nx, gl.cfg.Ny), complex)
    dpsi = np.zeros((gl.cfg.Nx, gl.cfg.Ny), complex)
    for i in range(gl.cfg.Nx):
        for j in range(gl.cfg.Ny):
            dpsi[i, j] = h
            gl.vars.order_parameter = gl.vars.order_parameter + dpsi
            re = (gl.observables.free_energy - E0) / h
            gl.vars.order_parameter = gl.vars.order_parameter - dpsi
            dpsi[i, j] = 1j * h
            gl.vars.order_parameter = gl.vars.order_parameter + dpsi
            im = (gl.observables.free_energy - E0) / h
            gl.vars.order_parameter = gl.vars.order_parameter - dpsi
            dpsi[i, j] = 0
            G[i, j] = re + 1j * im
    assert np.allclose(G1, G, atol=1e-5, rtol=1e-3)


def test_coef_psi_through_api():
    """tests/at_cg_coef_psi.py:35-57."""
    from svirl_amd.storage import GArray
    gl = GLSolver(Nx=10, Ny=9, dx=0.47, dy=0.42, gl_parameter=np.inf, homogeneous_external_field=0.1,
                  dtype=np.float64)
    dpsi = GArray(like=gl.vars.order_parameter)
    gl.solve._init_cg()
    rng = np.random.default_rng(2)
    for a, b, alpha in [[1.0, 0.0, 0.0], [0.0, 1.0, 1.0], [0.5, 0.5, 1.0]] + [
            [1.0 - bb * al, bb, al] for bb, al in rng.random((5, 2))]:
        gl.vars.order_parameter = 1.0
        gl.vars.randomize_order_parameter(level=0.5)
        psi0 = gl.vars.order_parameter
        E0 = gl.observables.free_energy
        gl.vars.order_parameter = psi0 * a
        dpsi.set_h(psi0 * b)
        dpsi.sync()
        c0, c1, c2, c3, c4 = gl.solve._cg._free_energy_conjgrad_coef_psi(dpsi.get_d_obj())
        P = c4 * alpha ** 4 + c3 * alpha ** 3 + c2 * alpha ** 2 + c1 * alpha + c0
        assert np.isclose(P, E0)


def test_reduction_api():
    """tests/at_reduction.py:20-97 through gl.par.red."""
    gl = GLSolver(Nx=32, Ny=32, dx=0.5, dy=0.5)
    rng = np.random.default_rng(5)
    for n in (1, 77, 500000):
        for bs in (64, 256, 1024):
            a = rng.random(n)
            assert np.isclose(np.sum(a), gl.par.red.test_sum(a, n, block_size=bs), atol=1e-10)
            nv = max(1, n // 5)
            v = rng.random(nv * 5)
            assert np.allclose(v.reshape(nv, 5).sum(axis=0), gl.par.red.test_sum_v(v, nv, 5, block_size=bs), atol=1e-10)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_barrier_timeout_reports_and_keeps_state(dtype):
    """The timeout path of the grid barrier, forced from the host: the run
    starts with the barrier's error flag raised (as a poll past its 5 s
    bound leaves it), every workgroup leaves at its first barrier, the call
    raises GLX_E_TIMEOUT, the solver keeps the state it had before the run,
    and the next run is the oracle's (FP64: sweep counts exact, psi to
    1e-10)."""
    from svirl_amd import _native as nat
    gl = GLSolver(Nx=300, Ny=200, dx=0.5, dy=0.5, order_parameter='random', random_seed=5, gl_parameter=2.0,
                  normal_conductivity=10.0, homogeneous_external_field=0.1, dtype=dtype)
    # 300 x 200 nodes: 5 strips x 200 rows, a persistent grid of 8 workgroups
    psi0 = np.array(gl.vars.order_parameter, copy=True)
    st = oracle_of(gl)
    nat.call("glx_td_force_barrier_abort")
    with pytest.raises(nat.GLXError) as err:
        gl.solve.td(dt=0.1, Nt=3)
    assert err.value.code == nat.GLX_E_TIMEOUT and "timed out" in str(err.value)
    assert gl.solve._td.sweeps == []
    assert np.array_equal(gl.vars.order_parameter, psi0)
    gl.solve.td(dt=0.1, Nt=3)
    st.td(dt=0.1, Nt=3)
    if dtype == np.float64:
        assert gl.solve._td.sweeps == st.sweeps
        assert rel(gl_psi(gl), st.psi) < 1e-10 and rel(gl_ab(gl), st.ab) < 1e-10
    else:
        assert rel(gl_psi(gl), st.psi) < 1e-4 and rel(gl_ab(gl), st.ab) < 1e-4


def _device_free_bytes():
    """Free device memory as the HIP runtime that libglx is linked to reports it."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so.7")
    free, total = ctypes.c_size_t(), ctypes.c_size_t()
    assert hip.hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)) == 0
    return free.value


def _material(gl, kind, rng):
    """The material patterns of tests/common.py:13-58 (apply_material_tiling)."""
    c = gl.cfg
    Nxc, Nyc = int(c.Nxc), int(c.Nyc)
    mt = np.full((Nxc, Nyc), True)
    if kind == 'empty':
        mt[:] = False
    elif kind.startswith('random'):
        n = int(int(kind[6:]) / 100 * Nxc * Nyc)
        mt[rng.integers(Nxc, size=n), rng.integers(Nyc, size=n)] = False
    elif kind[0] == 'q':
        q = int(kind[1]) - 1
        xs = slice(None, Nxc // 2) if q % 2 == 0 else slice(Nxc // 2, None)
        ys = slice(None, Nyc // 2) if q < 2 else slice(Nyc // 2, None)
        mt[xs, ys] = False
    elif kind == 'none':
        mt = None
    gl.mesh.material_tiling = mt
    gl.vars.set_order_parameter_to_zero_outside_material()


def test_constructor_destructor_cycles():
    """tests/at_destructor.py: ten solvers of random size (4..1023 nodes a
    side), kappa 1 or inf, randomized psi, external fields and a material
    pattern, each built, stepped twice and deleted; every cycle succeeds and
    the device memory they held is returned.

    The steps use sigma = 200 and dt = 0.05: the A solve is a Jacobi
    iteration that converges when dt kappa^2 / sigma is below about
    dx dy / 4, and at sigma = 1, dt = 0.1 and dx, dy in [0.2, 0.4) it
    diverges to NaN in the oracle as well (the reference algorithm's own
    behaviour, not a property of the GPU path)."""
    import gc
    rng = np.random.default_rng(2024)
    kinds = ['full', 'empty', 'q1', 'q2', 'q3', 'q4', 'none',
             'random10', 'random30', 'random50', 'random90', 'random200', 'random400', 'random800']

    # one-time device state (library init, code objects) before the count
    for dt in (np.float64, np.float32):
        gl = GLSolver(Nx=16, Ny=12, dx=0.3, dy=0.3, gl_parameter=2.0, normal_conductivity=200.0,
                      homogeneous_external_field=0.05, dtype=dt)
        gl.solve.td(dt=0.05, Nt=2)
        gl.observables.free_energy
        del gl
    gc.collect()
    free0 = _device_free_bytes()
    for i in range(10):
        Nx, Ny = (int(v) for v in rng.integers(4, 1024, size=2))
        kappa = 1.0 if rng.random() > 0.5 else np.inf
        gl = GLSolver(Nx=Nx, Ny=Ny, dx=0.2 + 0.2 * rng.random(), dy=0.2 + 0.2 * rng.random(),
                      gl_parameter=kappa, normal_conductivity=200.0,
                      dtype=np.float64 if i % 2 == 0 else np.float32)
        gl.vars.order_parameter = 1.0
        gl.vars.randomize_order_parameter(level=0.5, seed=i)
        if not np.isposinf(gl.params.gl_parameter):
            gl.params.gl_parameter = 1.0 + 3.0 * rng.random()
            gl.params.external_field = 0.01 + 0.1 * rng.random()
        gl.params.homogeneous_external_field = 0.01 + 0.1 * rng.random()
        _material(gl, kinds[rng.integers(len(kinds))], rng)
        gl.solve.td(dt=0.05, Nt=2)
        assert np.isfinite(gl.observables.free_energy), (i, Nx, Ny)
        del gl
        gc.collect()
        print(f"cycle {i}: {Nx}x{Ny}, device memory not returned so far "
              f"{(free0 - _device_free_bytes()) / 2**20:.1f} MiB", flush=True)
    leaked = free0 - _device_free_bytes()
    assert leaked <= 16 * 2**20


def test_precision_fp32_vs_fp64():
    """tests/at_precision.py in miniature: FP32-mixed and FP64 TD energies
    agree (the reference allows rtol 0.1; the FP32-mixed form is held to 1e-4)."""
    out = {}
    for dt_ in (np.float32, np.float64):
        gl = GLSolver(Lx=30, Ly=30, dx=1.0, dy=1.0, order_parameter=1.0, gl_parameter=3.6432,
                      normal_conductivity=400.0, homogeneous_external_field=0.1, dtype=dt_, convergence_rtol=1e-12)
        gl.params.fixed_vortices.order_parameter_add_vortices([15, 15], phase=True, deep=True)
        gl.mesh.material_tiling = lambda x, y: ~np.logical_and.reduce((np.abs(x - 15) < 3, np.abs(y - 15) < 3))
        gl.vars.set_order_parameter_to_zero_outside_material()
        gl.solve.td(Nt=300, dt=0.1)
        out[dt_] = gl.observables.free_energy
    assert np.isclose(out[np.float32], out[np.float64], rtol=1e-4)


def test_td_eqn_modes_and_chunking():
    """eqn='order_parameter' / 'vector_potential' and launches split into
    several persistent kernels give the reference's results."""
    gl = GLSolver(Lx=12, Ly=10, dx=0.5, dy=0.5, order_parameter='random', random_seed=5, gl_parameter=2.0,
                  normal_conductivity=20.0, homogeneous_external_field=0.2, dtype=np.float64)
    st = oracle_of(gl)
    gl.solve._init_td()
    gl.solve._td.steps_per_launch = 3
    gl.solve.td(dt=0.1, Nt=7, eqn='order_parameter')
    for _ in range(7):
        st.td_psi_solve(0.1)
    assert rel(gl_psi(gl), st.psi) < 1e-11
    gl.solve.td(dt=0.1, Nt=5, eqn='vector_potential')
    for _ in range(5):
        st.td_A_solve(0.1)
    assert rel(gl_ab(gl), st.ab) < 1e-11
    gl.solve.td(dt=0.1, Nt=8)
    st.td(dt=0.1, Nt=8)
    assert rel(gl_psi(gl), st.psi) < 1e-10 and rel(gl_ab(gl), st.ab) < 1e-10


def test_snapshot_resume_is_bit_exact(tmp_path):
    """Checkpoint after TD + CG, continue; a fresh solver loaded from the
    checkpoint continues to the same bits (psi, A, energies, sweeps)."""
    from svirl_amd import GLSolver, snapshot

    def make():
        return GLSolver(Nx=33, Ny=29, dx=0.5, dy=0.5, order_parameter='random', random_seed=4, gl_parameter=5.0,
                        normal_conductivity=100.0, homogeneous_external_field=0.15, dtype=np.float64)
    a = make()
    a.solve.td(dt=0.1, Nt=5)
    a.solve.cg(n_iter=3)
    p = str(tmp_path / "snap.npz")
    snapshot.save(a, p)
    a.solve.td(dt=0.1, Nt=4)
    a.solve.cg(n_iter=3)
    b = make()
    snapshot.load(b, p)
    b.solve.td(dt=0.1, Nt=4)
    b.solve.cg(n_iter=3)
    assert np.array_equal(a.vars.order_parameter, b.vars.order_parameter)
    aa, ab_ = a.vars.vector_potential
    ba, bb = b.vars.vector_potential
    assert np.array_equal(aa, ba) and np.array_equal(ab_, bb)
    assert a.solve._cg.cg_energies == b.solve._cg.cg_energies
    assert a.solve._td.sweeps[-4:] == b.solve._td.sweeps[-4:]


@pytest.mark.parametrize("wg", [1, 4])
def test_td_fixed_vortices_multi_workgroup(wg):
    """The fixed-source A solve and the per-step hooks on a grid of several
    strips, run by one workgroup and by a team of four (team barrier), in
    FP64 against the oracle: sweep counts exact, fields within 1e-10."""
    gl = GLSolver(Nx=131, Ny=70, dx=0.5, dy=0.5, order_parameter='random', random_level=0.3, random_seed=7,
                  gl_parameter=2.0, normal_conductivity=20.0, homogeneous_external_field=0.05,
                  fixed_vortices=([12.3, 40.1, 51.7], [10.2, 20.6, 27.9], [1, 1, -1]), phase_lock_radius=0.7,
                  dtype=np.float64)
    gl.solve._init_td()
    gl.solve._td.wg = wg
    st = oracle_of(gl)
    gl.solve.td(dt=0.1, Nt=4)
    st.td(dt=0.1, Nt=4)
    assert gl.solve._td.sweeps == st.sweeps
    assert rel(gl_psi(gl), st.psi) < 1e-10 and rel(gl_ab(gl), st.ab) < 1e-10
    assert rel(_vpi_of(gl), st.vpi) < 1e-10


def test_snapshot_resume_with_fixed_vortices(tmp_path):
    """The TD steps change vpi (td.py:256-325), so the snapshot carries it:
    a resumed run matches the uninterrupted one to the bit, vpi included."""
    from svirl_amd import GLSolver, snapshot

    def make():
        return GLSolver(Lx=12, Ly=12, dx=0.5, dy=0.5, order_parameter='random', random_level=0.2, random_seed=3,
                        gl_parameter=1.5, normal_conductivity=10.0,
                        fixed_vortices=([3.1, 8.7], [6.2, 6.3], [1, -1]), phase_lock_radius=0.6, dtype=np.float64)
    a = make()
    a.solve.td(dt=0.1, Nt=3)
    p = str(tmp_path / "snap_fv.npz")
    snapshot.save(a, p)
    a.solve.td(dt=0.1, Nt=3)
    b = make()
    snapshot.load(b, p)
    b.solve.td(dt=0.1, Nt=3)
    assert np.array_equal(a.vars.order_parameter, b.vars.order_parameter)
    assert np.array_equal(_vpi_of(a), _vpi_of(b)) and np.array_equal(gl_ab(a), gl_ab(b))
    with pytest.raises(ValueError):
        snapshot.load(GLSolver(Lx=12, Ly=12, dx=0.5, dy=0.5, gl_parameter=1.5, dtype=np.float64), p)


def test_plotter_quantities():
    """Every quantity svirl/plotter.py:18-64, 175-182 and 243 reads from a
    solver, taken the way the plotter takes it, has the plotter's shape
    and is finite, so a host-only plotting layer runs unchanged on the
    drop-in."""
    rng = np.random.default_rng(8)
    gl = GLSolver(Lx=10, Ly=8, dx=0.5, dy=0.5, order_parameter='random', random_seed=1, gl_parameter=2.0,
                  normal_conductivity=20.0, homogeneous_external_field=0.1,
                  fixed_vortices=([4.2], [3.9], [1]), dtype=np.float64)
    c = gl.cfg
    mt = np.ones((c.Nxc, c.Nyc), bool)
    mt[2:5, 2:4] = False
    gl.mesh.material_tiling = mt
    gl.params.linear_coefficient = 1.0 + 0.1 * rng.normal(size=(c.Nx, c.Ny))
    gl.solve.td(dt=0.1, Nt=3)
    node, cell, a_sh, b_sh = (c.Nx, c.Ny), (c.Nxc, c.Nyc), (c.Nxa, c.Nya), (c.Nxb, c.Nyb)
    jx, jy = gl.observables.current_density
    sx, sy = gl.observables.supercurrent_density
    nx, ny = gl.observables.normalcurrent_density
    va, vb = gl.vars.vector_potential
    ia, ib = gl.params.fixed_vortices.irregular_vector_potential
    q = {
        'material_tiling': (gl.mesh.material_tiling.astype(c.dtype), cell),
        'linear_coefficient': (gl.params.linear_coefficient, node),
        'superfluid_density': (gl.observables.superfluid_density, node),
        'order_parameter_phase': (np.angle(gl.vars.order_parameter) / (2.0 * np.pi), node),
        'op_fv_phase': (np.angle(gl.vars.order_parameter) - gl.params.fixed_vortices.fixed_vortices_phase, node),
        'magnetic_field': (gl.observables.magnetic_field - gl.params.homogeneous_external_field, cell),
        'current_density': (gl.mesh.interpolate_ab_array_to_c_array_abs(jx, jy), cell),
        'current_density_x': (jx, a_sh), 'current_density_y': (jy, b_sh),
        'supercurrent_density': (gl.mesh.interpolate_ab_array_to_c_array_abs(sx, sy), cell),
        'normalcurrent_density': (gl.mesh.interpolate_ab_array_to_c_array_abs(nx, ny), cell),
        'vector_potential_x': (va, a_sh), 'vector_potential_y': (vb, b_sh),
        'irregular_vector_potential_x': (ia, a_sh), 'irregular_vector_potential_y': (ib, b_sh),
        'fixed_vortices_phase': (gl.params.fixed_vortices.fixed_vortices_phase, node),
    }
    for name, (v, shape) in q.items():
        assert np.shape(v) == shape, name
        assert np.all(np.isfinite(v)), name
    assert np.all(q['current_density'][0] >= 0)
    extent = [-0.5 * c.dx, c.Lx + 0.5 * c.dx, -0.5 * c.dy, c.Ly + 0.5 * c.dy]
    assert np.isclose((extent[1] - extent[0]) / c.dx, c.Nx) and np.isclose((extent[3] - extent[2]) / c.dy, c.Ny)
    vx, vy, vv = gl.vortex_detector.vortices
    assert len(vx) == len(vy) == len(vv)


@pytest.mark.parametrize("kappa", [5.0, np.inf])
def test_cg_device_line_search(kappa):
    """The device line search (glx_cg_line_step, the FP32-mixed default)
    lands on the line minimum the reference's host solve finds (scipy BFGS /
    numpy polyroots): FP64 runs with either agree to 1e-8 in every energy;
    the FP32-mixed run ends within 1e-5 of the FP64 energy."""
    def make(dt_):
        gl = GLSolver(Nx=65, Ny=57, dx=0.5, dy=0.5, order_parameter='random', random_seed=7, gl_parameter=kappa,
                      normal_conductivity=100.0, homogeneous_external_field=0.1, dtype=dt_, convergence_rtol=0.0)
        gl.solve.td(dt=0.1, Nt=5)
        gl.solve._init_cg()
        return gl
    a, b, c = make(np.float64), make(np.float64), make(np.float32)
    b.solve._cg.line_search = 'device'
    assert a.solve._cg.line_search == 'host' and c.solve._cg.line_search == 'device'
    for gl in (a, b, c):
        gl.solve.cg(n_iter=30)
    ea, eb, ec = (np.array(gl.solve._cg.cg_energies, dtype=np.float64) for gl in (a, b, c))
    assert len(ea) == len(eb) == len(ec) == 30
    assert np.allclose(ea, eb, rtol=1e-8)
    assert np.all(np.diff(eb) <= 1e-12 * np.abs(eb[1:]))
    assert abs(ec[-1] - ea[-1]) <= 1e-5 * abs(ea[-1])
    # a second call continues from the directions and betas the first left
    # (the device loop enqueues each next direction step ahead of the stop
    # test; when the loop ends there, that step must leave no trace)
    for gl in (a, b):
        gl.solve.cg(n_iter=7)
    ea2, eb2 = (np.array(gl.solve._cg.cg_energies, dtype=np.float64) for gl in (a, b))
    assert len(ea2) == len(eb2) == 7 and np.allclose(ea2, eb2, rtol=1e-8)
    assert np.all(np.diff(eb2) <= 1e-12 * np.abs(eb2[1:])) and eb2[0] <= eb[-1] * (1 - 1e-14)


def test_cg_device_loop_stop_leaves_state():
    """A device-line-search run stopped by the energy test leaves the same
    directions and betas as a run capped at the same iteration: the
    direction step enqueued ahead of the stop test leaves no trace."""
    def make():
        gl = GLSolver(Nx=41, Ny=37, dx=0.5, dy=0.5, order_parameter='random', random_seed=9, gl_parameter=3.0,
                      normal_conductivity=50.0, homogeneous_external_field=0.1, dtype=np.float64,
                      convergence_rtol=1e-3)
        gl.solve.td(dt=0.1, Nt=3)
        gl.solve._init_cg()
        gl.solve._cg.line_search = 'device'
        return gl
    a = make()
    a.solve.cg(n_iter=200)
    n = len(a.solve._cg.cg_energies)
    assert 1 < n < 200                       # stopped by the energy test
    b = make()
    b.solve.cg(n_iter=n)                     # the same iterations, ended by the cap
    assert a.solve._cg.cg_energies == b.solve._cg.cg_energies
    sa, sb = a.solve._cg._state(), b.solve._cg._state()
    for k in sa:
        assert np.array_equal(sa[k], sb[k]), k
    assert np.array_equal(a.vars.order_parameter, b.vars.order_parameter)


def test_fixed_vortices_release_matches_oracle():
    """fixed_vortices_release (fixed_vortices.py:63-77) after TD steps with
    pinned vortices: psi takes exp(-i phase) of the irregular vector
    potential as it stands after the steps, vpi and the phase lock are
    dropped, and TD then continues as the oracle does without them."""
    gl = GLSolver(Nx=131, Ny=70, dx=0.5, dy=0.5, order_parameter='random', random_level=0.3, random_seed=7,
                  gl_parameter=2.0, normal_conductivity=20.0, homogeneous_external_field=0.05,
                  fixed_vortices=([12.3, 40.1, 51.7], [10.2, 20.6, 27.9], [1, 1, -1]), phase_lock_radius=0.7,
                  dtype=np.float64)
    st = oracle_of(gl)
    gl.solve.td(dt=0.1, Nt=4)
    st.td(dt=0.1, Nt=4)
    assert rel(_vpi_of(gl), st.vpi) < 1e-10
    fv = gl.params.fixed_vortices
    fv.fixed_vortices_release()
    psi_rel = O.release_fixed_vortices(st.g, st.psi, st.vpi)
    assert rel(gl_psi(gl), psi_rel) < 1e-10
    assert fv._vpi is None and fv._phase_lock_ns is None and fv.fixed_vortices_x.size == 0
    st.psi, st.vpi, st.lock_ns = psi_rel, None, None
    gl.solve.td(dt=0.1, Nt=3)
    st.td(dt=0.1, Nt=3)
    assert gl.solve._td.sweeps[-3:] == st.sweeps[-3:]
    assert rel(gl_psi(gl), st.psi) < 1e-10 and rel(gl_ab(gl), st.ab) < 1e-10


def test_startup_get_function_runs_the_reference_kernels():
    """gl.par.get_function(name) (startup.py:74-82) returns a callable with
    the reference kernel's argument order: elementwise kernels against
    numpy, the TD sweep and Jacobian against the oracle, the field against
    the observables; launch-shape-dependent kernels and unknown names raise."""
    from svirl_amd._native import DeviceArray
    gl = GLSolver(Nx=33, Ny=29, dx=0.5, dy=0.5, order_parameter='random', random_seed=3, gl_parameter=5.0,
                  normal_conductivity=100.0, homogeneous_external_field=0.15, dtype=np.float64)
    par = gl.par
    gl.vars._psi.sync_d()
    gl.vars._vp.sync_d()
    rng = np.random.default_rng(1)
    x, y = rng.normal(size=50), rng.normal(size=50)
    dx_, dy_, dz = DeviceArray.from_host(x), DeviceArray.from_host(y), DeviceArray.zeros(50, np.float64)
    par.get_function("axpy_r")(dx_, dy_, dz, np.float64(0.7), np.int32(50), block=(128, 1, 1), grid=(1, 1))
    assert np.array_equal(dz.get(), 0.7 * x + y)
    al = DeviceArray.from_host(np.array([1.5]))
    par.get_function("axmy_c")(dx_, dy_, dz, al, np.int32(25))
    assert np.array_equal(dz.get(), 1.5 * x - y)
    par.get_function("xpy_r")(dx_, dy_, np.int32(50))
    assert np.array_equal(dx_.get(), x + y)
    par.get_function("xmy_r")(dx_, dy_, np.int32(50))
    assert np.array_equal(dx_.get(), (x + y) - y)
    # magnetic field, as the observable computes it
    c = gl.cfg
    B = DeviceArray.zeros(int(c.Nxc) * int(c.Nyc), np.float64)
    par.get_function("magnetic_field")(None, gl.vars.vector_potential_h(), B)
    assert np.array_equal(np.reshape(B.get(), (c.Nyc, c.Nxc)).T, gl.observables.magnetic_field)
    # one psi Jacobi sweep and the psi Jacobian, against the oracle
    st = oracle_of(gl)
    g = st.g
    rhs = DeviceArray.from_host(st.psi.copy())
    cur = DeviceArray.from_host(st.psi.copy())
    nxt = DeviceArray.zeros(g.N, np.complex128)
    r2 = DeviceArray.zeros(1, np.int32)
    par.get_function("iterate_order_parameter_jacobi_step")(
        np.float64(0.1), np.float64(st.eps_scalar), None, None, gl.vars.vector_potential_h(), rhs, cur, nxt,
        np.float64(0.0), np.uint32(0), np.uint32(1), np.float64(1e-6), r2)
    ref_next, _, _ = O.psi_jacobi_sweep(g, 0.1, st.eps_scalar, None, None, st.ab, st.psi.copy(), st.psi, 0.0, 0, 1, 1e-6)
    assert rel(nxt.get(), ref_next) < 1e-12
    G = DeviceArray.zeros(g.N, np.complex128)
    par.get_function("free_energy_jacobian_psi")(np.float64(25.0), np.float64(st.eps_scalar), None, np.float64(0.15),
                                                 None, gl.vars.order_parameter_h(), None, gl.vars.vector_potential_h(), G)
    assert rel(G.get(), O.jacobian_psi(g, 25.0, st.eps_scalar, None, 0.15, None, st.psi, None, st.ab)) < 1e-12
    for name in ("sum", "free_energy_conjgrad_coef", "no_such_kernel"):
        with pytest.raises(ValueError):
            par.get_function(name)
