"""Batched TD (glx_batch_td_run through svirl_amd.batch.BatchTD): every
instance of a batch must end exactly where the FP64 oracle's td.py loops
take it when run alone (same sweep counts, fields within 1e-11), with
one-workgroup and multi-workgroup teams, with and without the HIP graph."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
nat = pytest.importorskip("svirl_amd._native")

from oracle import gl_oracle as O  # noqa: E402


@pytest.fixture(scope="module", autouse=True)
def _init():
    nat.init(0)


# (Nx, Ny, kappa, H, mask, eps map)
CASES = [
    (17, 17, np.inf, 0.10, False, False),
    (33, 21, 5.0, 0.20, True, True),
    (130, 65, 2.0, 0.30, False, True),
    (65, 130, 5.0, 0.05, True, False),
    (9, 12, np.inf, 0.45, False, False),
    (200, 40, 5.0, 0.15, True, True),
    (4, 4, 2.0, 0.30, True, True),       # the smallest grids GLSolver takes
    (5, 4, 5.0, 0.25, False, False),
]


def make(k, Nx, Ny, kappa, H, mask, epsmap, dtype=np.float64):
    from svirl_amd import GLSolver
    rng = np.random.default_rng(100 + k)
    dx = 0.5
    gl = GLSolver(Nx=Nx, Ny=Ny, dx=dx, dy=dx, order_parameter='random', random_seed=k + 1, gl_parameter=kappa,
                  normal_conductivity=100.0, homogeneous_external_field=H, dtype=dtype)
    c = gl.cfg
    if mask:
        mt = rng.random((int(c.Nxc), int(c.Nyc))) > 0.15
        gl.mesh.material_tiling = mt
        gl.vars.set_order_parameter_to_zero_outside_material()
    if epsmap:
        gl.params.linear_coefficient = 1.0 + 0.2 * rng.normal(size=(int(c.Nx), int(c.Ny)))
    return gl


def oracle_run(gl, dt, Nt):
    c = gl.cfg
    g = O.Geom(int(c.Nx), int(c.Ny), float(c.dx), float(c.dy))
    psi = gl.flatten_array(gl.vars.order_parameter)
    a, b = gl.vars.vector_potential
    ab = np.concatenate([gl.flatten_a_array(a), gl.flatten_b_array(b)])
    mt = None
    if gl.mesh.have_material_tiling():
        mt = np.ascontiguousarray(np.asarray(gl.mesh.material_tiling).T).ravel()
    em = None
    eps = 1.0
    lc = gl.params.linear_coefficient
    if np.ndim(lc) == 2:
        em = gl.flatten_array(lc)
    kappa = gl.params.gl_parameter
    st = O.OracleState(g, psi.astype(np.complex128), ab.astype(np.float64), kappa=kappa, sigma=100.0,
                       H=float(gl.params.homogeneous_external_field), eps=eps, eps_map=em, mt=mt)
    st.td(dt=dt, Nt=Nt)
    return st


@pytest.mark.parametrize("use_graph", [False, True])
def test_batch_fp64_matches_oracle(use_graph):
    from svirl_amd.batch import BatchTD
    sims = [make(k, *c) for k, c in enumerate(CASES)]
    refs = [oracle_run(gl, 0.1, 6) for gl in sims]
    BatchTD(sims, use_graph=use_graph, autotune=False).td(dt=0.1, Nt=6)
    for gl, st in zip(sims, refs):
        assert [tuple(x) for x in gl.solve._td.sweeps] == st.sweeps
        psi = gl.flatten_array(gl.vars.order_parameter)
        assert np.max(np.abs(psi - st.psi)) <= 1e-11
        if gl.params.solveA:
            a, b = gl.vars.vector_potential
            ab = np.concatenate([gl.flatten_a_array(a), gl.flatten_b_array(b)])
            assert np.max(np.abs(ab - st.ab)) <= 1e-11


def test_batch_equals_single_runs_and_continues():
    """Two batched calls (autotuned shape on the first) give bit for bit the
    fields of each instance run alone through gl.solve.td."""
    from svirl_amd.batch import BatchTD
    sims = [make(k, *c) for k, c in enumerate(CASES)]
    alone = [make(k, *c) for k, c in enumerate(CASES)]
    b = BatchTD(sims, use_graph=True, autotune=True, tune_steps=1)
    n1 = b.tune_need() + 1
    b.td(dt=0.1, Nt=n1)
    assert b.weights is not None
    # the tuner timed every pass depth (strip width) and several shapes
    assert {k[2] for k in b.timings} == {2, 3, 4} and len(b.timings) >= 8
    b.td(dt=0.1, Nt=3)
    assert b.shape in b.timings
    for gl, g1 in zip(sims, alone):
        g1.solve.td(dt=0.1, Nt=n1)
        g1.solve.td(dt=0.1, Nt=3)
        assert gl.solve._td.sweeps == g1.solve._td.sweeps
        assert np.array_equal(gl.vars.order_parameter, g1.vars.order_parameter)
        a, b_ = gl.vars.vector_potential
        a1, b1 = g1.vars.vector_potential
        assert np.array_equal(a, a1) and np.array_equal(b_, b1)


def test_batch_fp32_close_to_single_fp32():
    from svirl_amd.batch import BatchTD
    sims = [make(k, *c, dtype=np.float32) for k, c in enumerate(CASES)]
    alone = [make(k, *c, dtype=np.float32) for k, c in enumerate(CASES)]
    BatchTD(sims, autotune=False).td(dt=0.1, Nt=5)
    for gl, g1 in zip(sims, alone):
        g1.solve.td(dt=0.1, Nt=5)
        s0 = np.array(gl.solve._td.sweeps)
        s1 = np.array(g1.solve._td.sweeps)
        assert np.all(np.abs(s0 - s1) <= 1)
        assert np.max(np.abs(gl.vars.order_parameter - g1.vars.order_parameter)) <= 1e-5


def test_batch_barrier_timeout_keeps_states():
    """glx_batch_td_run's timeout path, forced from the host (the run starts
    with the barrier's error flag raised): the call raises GLX_E_TIMEOUT,
    every instance keeps the state it had, and the next run matches the
    oracle. The aborted call makes one step: a one-workgroup team does not
    poll the flag and steps on, and a single step writes only buffers that
    were not current."""
    from svirl_amd.batch import BatchTD
    sims = [make(k, *c) for k, c in enumerate(CASES)]
    before = [np.array(gl.vars.order_parameter, copy=True) for gl in sims]
    refs = [oracle_run(gl, 0.1, 4) for gl in sims]
    b = BatchTD(sims, use_graph=False, autotune=False)
    nat.call("glx_td_force_barrier_abort")
    with pytest.raises(nat.GLXError) as err:
        b.td(dt=0.1, Nt=1)
    assert err.value.code == nat.GLX_E_TIMEOUT
    for gl, p0 in zip(sims, before):
        assert np.array_equal(gl.vars.order_parameter, p0)
    b.td(dt=0.1, Nt=4)
    for gl, st in zip(sims, refs):
        assert [tuple(x) for x in gl.solve._td.sweeps] == st.sweeps
        assert np.max(np.abs(gl.flatten_array(gl.vars.order_parameter) - st.psi)) <= 1e-11
