"""Batched CG (glx_batch_cg_run through svirl_amd.batch.BatchCG).

The batched form runs the reference's Polak-Ribiere CG (cg.py:452-551) with
an FP64 device line search and a per-instance gradient-norm stop, so it is
checked against the FP64 oracle's pieces rather than its BFGS trajectory:
the first direction is exactly -dF/d(psi, A) (oracle Jacobians, eps map
included), the step lands on a minimum of the oracle's free energy along
that direction, energies fall, instances stop raggedly on their own test,
and FP32-mixed ends at the FP64 form's free energy."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
nat = pytest.importorskip("svirl_amd._native")

from oracle import gl_oracle as O  # noqa: E402

CASES = [(17, 13, 5.0, 0.2), (33, 29, 5.0, 0.1), (65, 40, 5.0, 0.3), (4, 5, 5.0, 0.2)]


@pytest.fixture(scope="module", autouse=True)
def _init():
    nat.init(0)


def make(k, Nx, Ny, kappa, H, dtype=np.float64, init='random'):
    from svirl_amd import GLSolver
    rng = np.random.default_rng(300 + k)
    gl = GLSolver(Nx=Nx, Ny=Ny, dx=0.5, dy=0.5, order_parameter=init, random_seed=k + 3, gl_parameter=kappa,
                  normal_conductivity=200.0, homogeneous_external_field=H, dtype=dtype)
    c = gl.cfg
    gl.mesh.material_tiling = rng.random((int(c.Nxc), int(c.Nyc))) > 0.1
    gl.vars.set_order_parameter_to_zero_outside_material()
    gl.params.linear_coefficient = 1.0 + 0.2 * rng.normal(size=(int(c.Nx), int(c.Ny)))
    return gl


def state(gl):
    c = gl.cfg
    g = O.Geom(int(c.Nx), int(c.Ny), float(c.dx), float(c.dy))
    psi = gl.flatten_array(gl.vars.order_parameter).astype(np.complex128)
    a, b = gl.vars.vector_potential
    ab = np.concatenate([gl.flatten_a_array(a), gl.flatten_b_array(b)]).astype(np.float64)
    mt = np.ascontiguousarray(np.asarray(gl.mesh.material_tiling).T).ravel()
    em = gl.flatten_array(gl.params.linear_coefficient).astype(np.float64)
    k2 = float(gl.params.gl_parameter) ** 2
    return g, psi, ab, mt, em, k2, float(gl.params.homogeneous_external_field)


def F(g, k2, em, H, mt, psi, ab):
    return O.free_energy(g, k2, 0.0, em, H, mt, psi, None, ab)


def exact_jacobian_A(g, k2, H, mt, psi, ab):
    """The oracle's Jacobian with the field part on boundary links halved
    (the exact dF/dA): the field part is the Jacobian at psi = 0."""
    jac = O.jacobian_A(g, k2, H, mt, psi, None, ab)
    field = O.jacobian_A(g, k2, H, mt, np.zeros_like(psi), None, ab)
    a = np.zeros((g.Nya, g.Nxa), bool)
    a[0, :] = a[-1, :] = True
    b = np.zeros((g.Nyb, g.Nxb), bool)
    b[:, 0] = b[:, -1] = True
    edge = np.concatenate([a.ravel(), b.ravel()])
    return jac - 0.5 * field * edge


def test_exact_jacobian_matches_finite_differences():
    """The halved boundary field part is the derivative of the oracle's F
    (central differences on a few boundary and interior links)."""
    sims = [make(0, 9, 8, 5.0, 0.2)]
    g, psi, ab, mt, em, k2, H = state(sims[0])
    J = exact_jacobian_A(g, k2, H, mt, psi, ab)
    h = 1e-6
    for k in [0, 3, g.Nxa, g.Na - 1, g.Na, g.Na + 4, g.Nab - 1]:
        e = np.zeros_like(ab)
        e[k] = h
        fd = (F(g, k2, em, H, mt, psi, ab + e) - F(g, k2, em, H, mt, psi, ab - e)) / (2 * h)
        assert abs(fd - J[k]) <= 1e-6 * (1 + abs(J[k])), (k, fd, J[k])


def test_first_iteration_is_oracle_steepest_descent_and_line_minimum():
    from svirl_amd.batch import BatchCG
    sims = [make(k, *c) for k, c in enumerate(CASES)]
    before = [state(gl) for gl in sims]
    iters, conv, _ = BatchCG(sims).cg(max_iter=1, rtol=0.0)
    assert list(iters) == [1] * len(CASES) and not conv.any()
    for gl, (g, psi0, ab0, mt, em, k2, H) in zip(sims, before):
        _, psi1, ab1, *_ = state(gl)
        d = -O.jacobian_psi(g, k2, 0.0, em, H, mt, psi0, None, ab0)
        dA = -exact_jacobian_A(g, k2, H, mt, psi0, ab0)
        dp = psi1 - psi0
        ap = np.vdot(d, dp).real / np.vdot(d, d).real
        aA = np.dot(dA, ab1 - ab0) / np.dot(dA, dA)
        assert ap > 0 and aA > 0
        assert np.linalg.norm(dp - ap * d) <= 1e-10 * np.linalg.norm(ap * d)
        assert np.linalg.norm(ab1 - ab0 - aA * dA) <= 1e-10 * np.linalg.norm(aA * dA)
        f = lambda x, y: F(g, k2, em, H, mt, psi0 + x * d, ab0 + y * dA)
        f0, fa = f(0.0, 0.0), f(ap, aA)
        assert fa < f0
        for sx in (0.9, 1.1):
            for sy in (0.9, 1.0, 1.1):
                assert fa <= f(sx * ap, sy * aA) + 1e-12 * abs(fa)


def test_falling_energy_from_random_state():
    from svirl_amd.batch import BatchCG
    sims = [make(k, *c) for k, c in enumerate(CASES)]
    f = [[F(*[state(gl)[i] for i in (0, 5, 4, 6, 3, 1, 2)])] for gl in sims]
    b = BatchCG(sims)
    for _ in range(4):
        b.cg(max_iter=25, rtol=0.0)
        for k, gl in enumerate(sims):
            f[k].append(F(*[state(gl)[i] for i in (0, 5, 4, 6, 3, 1, 2)]))
    # a small instance can reach its minimum inside the run: from there the
    # energy stays where it is (to rounding)
    for fk in f:
        assert fk[1] < fk[0], fk
        assert all(x1 <= x0 + 1e-13 * abs(x0) for x0, x1 in zip(fk, fk[1:])), fk


def test_ragged_convergence_from_uniform_state():
    """From psi = 1 (near the Meissner minimum) every instance reaches its
    own gradient-norm stop, after a different number of iterations."""
    from svirl_amd.batch import BatchCG
    sims = [make(k, *c, init=1.0) for k, c in enumerate(CASES)]
    iters, conv, grel = BatchCG(sims).cg(max_iter=3000, rtol=1e-6, iters_per_launch=37)
    assert conv.all(), (iters, grel)
    assert np.all(grel <= 1e-6)
    assert len(set(iters.tolist())) > 1, iters


def test_fp32_mixed_reaches_fp64_energy():
    from svirl_amd.batch import BatchCG
    s64 = [make(k, *c, init=1.0) for k, c in enumerate(CASES)]
    s32 = [make(k, *c, dtype=np.float32, init=1.0) for k, c in enumerate(CASES)]
    BatchCG(s64).cg(max_iter=3000, rtol=1e-6)
    BatchCG(s32).cg(max_iter=3000, rtol=1e-4)
    for a, b in zip(s32, s64):
        fa, fb = a.observables.free_energy_fp64(), b.observables.free_energy_fp64()
        assert abs(fa - fb) <= 1e-5 * abs(fb)


def test_fp32_mixed_stalls_then_fp64_finish_reaches_rtol():
    """At rtol = 1e-6 the FP32-mixed gradient stalls above the stop test;
    the stall rule stops those instances, and the FP64 finish continues each
    from where it stopped (its iteration count and the FP32 run's initial
    gradient norm) until the FP64 gradient passes the test. The result ends
    at the FP64-only run's free energy, to FP32 rounding of the state."""
    from svirl_amd.batch import BatchCG
    s32 = [make(k, *c, dtype=np.float32, init=1.0) for k, c in enumerate(CASES)]
    plain = [make(k, *c, dtype=np.float32, init=1.0) for k, c in enumerate(CASES)]
    s64 = [make(k, *c, init=1.0) for k, c in enumerate(CASES)]
    b0 = BatchCG(plain)
    it0, cv0, _ = b0.cg(max_iter=3000, rtol=1e-6, stall_iters=50)
    assert b0.stalled.any() and not (cv0 & b0.stalled).any(), (it0, cv0, b0.stalled)
    b = BatchCG(s32)
    iters, conv, grel = b.cg(max_iter=3000, rtol=1e-6, stall_iters=50, fp64_finish=True)
    assert conv.all(), (iters, grel, b.fp64_iters)
    assert np.all(grel <= 1e-6)
    # the FP32 phase is the plain stalled run, the finish adds its own iterations
    fin = b.fp64_iters > 0
    assert np.array_equal(fin, b0.stalled), (b.fp64_iters, b0.stalled)
    assert np.array_equal(iters[fin], it0[fin] + b.fp64_iters[fin]), (iters, it0, b.fp64_iters)
    assert np.array_equal(iters[~fin], it0[~fin])
    it64, cv64, _ = BatchCG(s64).cg(max_iter=3000, rtol=1e-6)
    assert cv64.all()
    for a, c in zip(s32, s64):
        fa, fc = a.observables.free_energy_fp64(), c.observables.free_energy_fp64()
        assert abs(fa - fc) <= 1e-6 * abs(fc), (fa, fc)


def test_fp64_finish_argument_checks():
    from svirl_amd.batch import BatchCG
    s64 = [make(0, *CASES[0])]
    with pytest.raises(ValueError):
        BatchCG(s64).cg(max_iter=10, rtol=1e-8, stall_iters=5, fp64_finish=True)
    s32 = [make(0, *CASES[0], dtype=np.float32)]
    with pytest.raises(ValueError):
        BatchCG(s32).cg(max_iter=10, rtol=1e-8, fp64_finish=True)
