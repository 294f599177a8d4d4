"""glx_batch_td_run_noalloc (no device allocation per call; the library's
fixed scratch array is its workspace) against glx_batch_td_run on the same
instances: sweep counts, psi and A bit for bit, over two calls in a row, the
second on a smaller batch (which leaves the larger call's contents in the
shared array). A run too large for the array is refused before any launch."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
nat = pytest.importorskip("svirl_amd._native")


@pytest.fixture(scope="module", autouse=True)
def _init():
    nat.init(0)


# (Nx, Ny, kappa, H, mask, eps map)
CASES = [
    (33, 21, 5.0, 0.20, True, True),
    (17, 17, np.inf, 0.10, False, False),
    (130, 65, 2.0, 0.30, False, True),
    (65, 130, 5.0, 0.05, True, False),
    (9, 12, np.inf, 0.45, False, False),
    (200, 40, 5.0, 0.15, True, True),
]


def make(k, Nx, Ny, kappa, H, mask, epsmap, dtype):
    from svirl_amd import GLSolver
    rng = np.random.default_rng(700 + k)
    gl = GLSolver(Nx=Nx, Ny=Ny, dx=0.5, dy=0.5, order_parameter='random', random_seed=k + 7, gl_parameter=kappa,
                  normal_conductivity=100.0, homogeneous_external_field=H, dtype=dtype)
    c = gl.cfg
    if mask:
        gl.mesh.material_tiling = rng.random((int(c.Nxc), int(c.Nyc))) > 0.15
        gl.vars.set_order_parameter_to_zero_outside_material()
    if epsmap:
        gl.params.linear_coefficient = 1.0 + 0.2 * rng.normal(size=(int(c.Nx), int(c.Ny)))
    return gl


def state(gl):
    a, b = gl.vars.vector_potential
    return np.array(gl.solve._td.sweeps), np.array(gl.vars.order_parameter), np.array(a), np.array(b)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_noalloc_equals_batch_td_run_twice(dtype):
    from svirl_amd.batch import BatchTD
    for cases, steps in ((CASES, 5), (CASES[:3], 3)):  # the second call is smaller
        ref = [make(k, *c, dtype=dtype) for k, c in enumerate(cases)]
        out = [make(k, *c, dtype=dtype) for k, c in enumerate(cases)]
        BatchTD(ref, use_graph=False, autotune=False).td(dt=0.1, Nt=steps)
        BatchTD(out, autotune=False, noalloc=True).td(dt=0.1, Nt=steps)
        for g0, g1 in zip(ref, out):
            for x0, x1 in zip(state(g0), state(g1)):
                assert x0.shape == x1.shape and np.array_equal(x0, x1)


def test_noalloc_refuses_runs_above_its_scratch():
    from svirl_amd.batch import BatchTD
    sims = [make(k, *c, dtype=np.float64) for k, c in enumerate(CASES[:2])]
    before = [np.array(gl.vars.order_parameter) for gl in sims]
    b = BatchTD(sims, autotune=False, noalloc=True)
    with pytest.raises(nat.GLXError) as err:
        b.td(dt=0.1, Nt=600000)  # a sweep log of 2 x 600000 x 8 bytes
    assert err.value.code == nat.GLX_E_ARG and "maximum" in str(err.value)
    for gl, p0 in zip(sims, before):
        assert np.array_equal(gl.vars.order_parameter, p0)
