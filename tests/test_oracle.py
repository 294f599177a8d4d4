"""Pins the FP64 oracle (oracle/gl_oracle.py) on the CPU.

Each check is chosen so that a plausible transcription mistake (a dropped
term, a wrong sign, index or transposed operand) fails it:

* per-thread loop interpreter (oracle/loops.py) vs the vectorised oracle,
  with random masks, epsilon maps and external potentials;
* closed forms: psi == 1 is a fixed point of the psi sweep; the symmetric
  gauge of H plus any pure gauge is a fixed point of the A sweep when
  psi == 0; F(psi == 1) = -Lx*Ly/2; B == H and J == 0 for the gauge;
* discrete gauge invariance of the psi sweep, F, J_s and B;
* the reference tests re-expressed: tests/at_cg_jacobians.py (finite
  differences), tests/at_cg_coef_psi.py and tests/at_cg_coef.py (polynomial
  identities), tests/at_reduction.py (sums);
* a known vortex imprint for the winding number and triangulation;
* uint32 brute force for the Langevin hash.
"""
import math

import numpy as np
import pytest

from oracle import gl_oracle as O
from oracle import loops as L


def _case(seed, Nx=None, Ny=None, mask=True, epsmap=True, abei=True):
    rng = np.random.default_rng(seed)
    Nx = int(Nx or rng.integers(6, 11))
    Ny = int(Ny or rng.integers(6, 11))
    dx, dy = 0.5 - 0.1 * rng.random(2)
    g = O.Geom(Nx, Ny, dx, dy)
    mt = (rng.random(g.Nc) > 0.3) if mask else None
    psi = (0.3 + rng.random(g.N)) * np.exp(2j * np.pi * rng.random(g.N))
    ab = 0.3 * rng.normal(size=g.Nab)
    ae = 0.2 * rng.normal(size=g.Nab) if abei else None
    em = (1.0 + 0.2 * rng.normal(size=g.N)) if epsmap else None
    return rng, g, mt, psi, ab, ae, em


def _chi_gauge(g, rng):
    """A random gauge function chi on nodes and its discrete gradient on links."""
    chi = rng.normal(size=(g.Ny, g.Nx))
    da = (chi[:, 1:] - chi[:, :-1]) / g.dx
    db = (chi[1:, :] - chi[:-1, :]) / g.dy
    return chi.ravel(), np.concatenate([da.ravel(), db.ravel()])


# ---------------------------------------------------------------------------
# loop interpreter vs vectorised restatement
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("seed", range(6))
def test_loops_match_vectorised(seed):
    rng, g, mt, psi, ab, ae, em = _case(seed, mask=seed % 3 != 0, epsmap=seed % 2 == 1, abei=seed % 2 == 0)
    rhs = psi * (1 + 0.1 * rng.random(g.N))
    a, _, r1 = O.psi_jacobi_sweep(g, 0.1, 0.9, em, mt, ab, rhs, psi, stop_eps=1e-6)
    b, r2 = L.psi_sweep(g, 0.1, 0.9, em, mt, ab, rhs, psi, 1e-6)
    assert np.array_equal(a, np.array(b)) and r1 == r2
    src = ab + 0.01 * rng.random(g.Nab)
    rA = ab + 0.02 * rng.random(g.Nab)
    a, _, r1 = O.a_jacobi_sweep(g, 0.1, 25.0, 1 / 200, 0.1, mt, psi, src, rA, ab, stop_eps=1e-8)
    b, r2 = L.a_sweep(g, 0.1, 25.0, 1 / 200, 0.1, mt, psi, src, rA, ab, 1e-8)
    assert np.array_equal(a, np.array(b)) and r1 == r2
    d = rng.normal(size=g.N) + 1j * rng.normal(size=g.N)
    dab = rng.normal(size=g.Nab)
    for k2 in (-1.0, 9.0):
        F1 = O.free_energy(g, k2, 0.9, em, 0.1, mt, psi, ae, ab)
        F2 = math.fsum(L.free_energy_terms(g, k2, 0.9, em, 0.1, mt, psi, ae, ab)) * g.dx * g.dy
        assert F1 == pytest.approx(F2, rel=1e-13, abs=1e-13)
        assert np.array_equal(O.jacobian_psi(g, k2, 0.9, em, 0.1, mt, psi, ae, ab),
                              np.array(L.jacobian_psi(g, k2, 0.9, em, 0.1, mt, psi, ae, ab)))
        c1 = O.coef_psi(g, k2, 0.9, 0.1, mt, psi, d, ae, ab)
        c2 = np.array(L.coef_terms(g, k2, 0.9, 0.1, mt, psi, d, ae, ab))
        assert np.allclose(c1, c2, rtol=1e-12, atol=1e-11)
        c1 = O.coef(g, k2, 0.9, 0.1, mt, psi, d, ae, ab, dab)
        c2 = np.array(L.coef_terms(g, k2, 0.9, 0.1, mt, psi, d, ae, ab, dab))
        assert np.allclose(c1, c2, rtol=1e-12, atol=1e-10)
    assert np.array_equal(O.jacobian_A(g, 9.0, 0.1, mt, psi, ae, ab),
                          np.array(L.jacobian_A(g, 9.0, 0.1, mt, psi, ae, ab)))
    assert np.array_equal(O.current_density(g, 9.0, 0.1, ae, ab),
                          np.array(L.current_density(g, 9.0, 0.1, ae, ab)))
    assert np.array_equal(O.supercurrent_density(g, mt, psi, ae, ab),
                          np.array(L.supercurrent_density(g, mt, psi, ae, ab)))
    assert np.array_equal(O.magnetic_field(g, ae, ab), np.array(L.magnetic_field(g, ae, ab)))


# ---------------------------------------------------------------------------
# closed forms and invariants
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("mask", [False, True])
def test_psi_one_is_fixed_point(mask):
    rng, g, mt, *_ = _case(11, mask=mask)
    psi = np.ones(g.N, dtype=complex)
    nxt, _, r = O.psi_jacobi_sweep(g, 0.1, 1.0, None, mt, np.zeros(g.Nab), psi, psi, stop_eps=1e-12)
    inside = O.mask_at_nodes(g, mt)
    assert np.array_equal(nxt[inside], psi[inside])
    assert np.all(nxt[~inside] == 0)


def test_gauge_is_fixed_point_of_A_sweep():
    rng, g, mt, *_ = _case(12, mask=True)
    H = 0.37
    A0 = O.symmetric_gauge(g, H)
    _, grad = _chi_gauge(g, rng)
    A = A0 + grad
    psi0 = np.zeros(g.N, dtype=complex)
    nxt, _, r = O.a_jacobi_sweep(g, 0.1, 25.0, 1 / 7.0, H, mt, psi0, A, A, A, stop_eps=1e-6)
    assert np.allclose(nxt, A, rtol=0, atol=1e-12)
    # a field mismatch must not be a fixed point (boundary H source is live)
    nxt2, _, _ = O.a_jacobi_sweep(g, 0.1, 25.0, 1 / 7.0, H + 0.1, mt, psi0, A, A, A, stop_eps=1e-6)
    assert np.max(np.abs(nxt2 - A)) > 1e-3


def test_free_energy_uniform_state():
    g = O.Geom(13, 9, 0.5, 0.4)
    psi = np.ones(g.N, dtype=complex)
    F = O.free_energy(g, -1.0, 1.0, None, 0.0, None, psi, None, np.zeros(g.Nab))
    assert F == pytest.approx(-0.5 * g.Lx * g.Ly, rel=1e-14)
    # kappa finite, gauge A of H: magnetic term vanishes as well
    H = 0.2
    F = O.free_energy(g, 25.0, 1.0, None, H, None, np.ones(g.N) + 0j, None, O.symmetric_gauge(g, H))
    grad = F + 0.5 * g.Lx * g.Ly
    assert grad > 0  # a uniform psi in a field costs kinetic energy


def test_gauge_field_and_current():
    g = O.Geom(12, 10, 0.5, 0.45)
    H = 0.3
    A = O.symmetric_gauge(g, H)
    assert np.allclose(O.magnetic_field(g, None, A), H, atol=1e-13)
    assert np.allclose(O.current_density(g, 9.0, H, None, A), 0.0, atol=1e-11)
    ae = O.symmetric_gauge(g, 0.1)
    assert np.allclose(O.magnetic_field(g, ae, A), H + 0.1, atol=1e-13)


@pytest.mark.parametrize("seed", range(3))
def test_gauge_invariance(seed):
    rng, g, mt, psi, ab, _, em = _case(20 + seed)
    chi, grad = _chi_gauge(g, rng)
    psi2 = psi * np.exp(1j * chi)
    ab2 = ab + grad
    rhs = psi * (1 + 0.05 * rng.random(g.N))
    n1, _, _ = O.psi_jacobi_sweep(g, 0.1, 1.0, em, mt, ab, rhs, psi, stop_eps=1e-6)
    n2, _, _ = O.psi_jacobi_sweep(g, 0.1, 1.0, em, mt, ab2, rhs * np.exp(1j * chi), psi2, stop_eps=1e-6)
    assert np.allclose(n2, n1 * np.exp(1j * chi), atol=1e-12)
    for k2 in (-1.0, 4.0):
        F1 = O.free_energy(g, k2, 1.0, em, 0.1, mt, psi, None, ab)
        F2 = O.free_energy(g, k2, 1.0, em, 0.1, mt, psi2, None, ab2)
        assert F2 == pytest.approx(F1, rel=1e-12)
    assert np.allclose(O.supercurrent_density(g, mt, psi, None, ab),
                       O.supercurrent_density(g, mt, psi2, None, ab2), atol=1e-12)
    assert np.allclose(O.magnetic_field(g, None, ab), O.magnetic_field(g, None, ab2), atol=1e-11)
    # A sweep: the supercurrent source is gauge invariant too
    a1, _, _ = O.a_jacobi_sweep(g, 0.1, 9.0, 0.01, 0.1, mt, psi, ab, ab, ab, stop_eps=1e-6)
    a2, _, _ = O.a_jacobi_sweep(g, 0.1, 9.0, 0.01, 0.1, mt, psi2, ab2, ab2, ab2, stop_eps=1e-6)
    assert np.allclose(a2 - grad, a1, atol=1e-11)


def test_td_source_matches_supercurrent_in_bulk():
    """The A sweep's J_s source equals the supercurrent observable on links
    whose Du weight is 1 (two different kernels, td.h vs observables.h)."""
    rng, g, mt, psi, ab, _, _ = _case(31, mask=False)
    zero = np.zeros(g.Nab)
    # with kappa2 = 0, dt*rho = 1, rhs = 0 the sweep returns exactly the source term
    out, _, _ = O.a_jacobi_sweep(g, 1.0, 0.0, 1.0, 0.0, None, psi, ab, zero, zero, stop_eps=1.0)
    js = O.supercurrent_density(g, None, psi, None, ab)
    a_bulk = np.zeros((g.Ny, g.Nxa), bool); a_bulk[1:-1, :] = True
    b_bulk = np.zeros((g.Nyb, g.Nx), bool); b_bulk[:, 1:-1] = True
    bulk = np.concatenate([a_bulk.ravel(), b_bulk.ravel()])
    assert np.allclose(out[bulk], js[bulk], atol=1e-14)
    assert np.allclose(out[~bulk], 2 * js[~bulk], atol=1e-14)  # Du weight 1/2 on the edge


# ---------------------------------------------------------------------------
# reference tests re-expressed
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("seed", range(3))
def test_jacobians_vs_finite_differences(seed):
    """tests/at_cg_jacobians.py:24-108 with central differences (the
    reference uses one-sided h = 3e-9, atol 1e-5, rtol 1e-3)."""
    rng = np.random.default_rng(100 + seed)
    g = O.Geom(8 + rng.integers(4), 8 + rng.integers(4), 0.5 - 0.1 * rng.random(), 0.5 - 0.1 * rng.random())
    kappa = 1.0 + 3.0 * rng.random()
    k2 = kappa ** 2
    H = 0.01 + 0.1 * rng.random()
    ab = O.symmetric_gauge(g, H)
    ae = O.symmetric_gauge(g, 0.01 + 0.1 * rng.random())
    mt = (rng.random(g.Nc) > 0.2) if seed else None
    psi = O.random_psi(g.N, 0.5, seed=seed)
    psi = np.where(O.mask_at_nodes(g, mt), psi, 0)
    em = 1.0 + 0.1 * rng.random(g.N)
    F = lambda p, a: O.free_energy(g, k2, 0.0, em, H, mt, p, ae, a)  # noqa: E731
    h = 1e-6
    jp = O.jacobian_psi(g, k2, 0.0, em, H, mt, psi, ae, ab)
    fd = np.empty(g.N, complex)
    for n in range(g.N):
        e = np.zeros(g.N, complex); e[n] = h
        re = (F(psi + e, ab) - F(psi - e, ab)) / (2 * h)
        im = (F(psi + 1j * e, ab) - F(psi - 1j * e, ab)) / (2 * h)
        fd[n] = re + 1j * im
    assert np.allclose(jp, fd, rtol=1e-6, atol=1e-7)
    jA = O.jacobian_A(g, k2, H, mt, psi, ae, ab)
    fdA = np.empty(g.Nab)
    for n in range(g.Nab):
        e = np.zeros(g.Nab); e[n] = h
        fdA[n] = (F(psi, ab + e) - F(psi, ab - e)) / (2 * h)
    assert np.allclose(jA, fdA, rtol=1e-6, atol=1e-7)


def test_jacobian_A_boundary_factor_off_gauge():
    """Away from the exact gauge the kernel's boundary links carry twice the
    magnetic part of dF/dA (cg.h:179-200); interior links stay exact."""
    rng = np.random.default_rng(7)
    g = O.Geom(9, 8, 0.5, 0.5)
    H = 0.1
    ab = O.symmetric_gauge(g, H) + 0.05 * rng.normal(size=g.Nab)
    psi = np.zeros(g.N, complex)
    F = lambda a: O.free_energy(g, 4.0, 1.0, None, H, None, psi, None, a)  # noqa: E731
    jA = O.jacobian_A(g, 4.0, H, None, psi, None, ab)
    h = 1e-6
    fd = np.array([(F(ab + h * e) - F(ab - h * e)) / (2 * h) for e in np.eye(g.Nab)])
    a_int = np.zeros((g.Ny, g.Nxa), bool); a_int[1:-1] = True
    b_int = np.zeros((g.Nyb, g.Nx), bool); b_int[:, 1:-1] = True
    interior = np.concatenate([a_int.ravel(), b_int.ravel()])
    assert np.allclose(jA[interior], fd[interior], rtol=1e-6, atol=1e-8)
    assert np.allclose(jA[~interior], 2 * fd[~interior], rtol=1e-6, atol=1e-8)


@pytest.mark.parametrize("kappa2", [-1.0, 6.0])
def test_coef_psi_polynomial_identity(kappa2):
    """tests/at_cg_coef_psi.py:35-57: P(alpha) == F(psi + alpha dpsi)."""
    rng, g, mt, psi, ab, _, _ = _case(40, epsmap=False, abei=False)
    d = 0.3 * (rng.normal(size=g.N) + 1j * rng.normal(size=g.N))
    c = O.coef_psi(g, kappa2, 0.8, 0.05, mt, psi, d, None, ab)
    for alpha in (0.0, 0.3, 1.0, -0.7, 2.5):
        P = sum(c[k] * alpha ** k for k in range(5))
        E = O.free_energy(g, kappa2, 0.8, None, 0.05, mt, psi + alpha * d, None, ab)
        assert P == pytest.approx(E, rel=1e-12, abs=1e-12)


def test_coef17_expansion():
    """tests/at_cg_coef.py:33-125: exact in alpha_psi when dA = 0, and the
    alpha_A expansion error falls from P0 to P2 to P4 as |dA| shrinks."""
    rng, g, mt, psi, ab, ae, _ = _case(41, epsmap=False)
    k2, H, eps = 5.0, 0.07, 0.9
    d = 0.3 * (rng.normal(size=g.N) + 1j * rng.normal(size=g.N))
    c = O.coef17_to_matrix(O.coef(g, k2, eps, H, mt, psi, d, ae, ab, np.zeros(g.Nab)))
    for al in (0.4, 1.3):
        E = O.free_energy(g, k2, eps, None, H, mt, psi + al * d, ae, ab)
        assert np.polynomial.polynomial.polyval2d(al, 0.0, c) == pytest.approx(E, rel=1e-12)
    dab = rng.normal(size=g.Nab)
    errs = []
    for r in (1e-1, 1e-2):
        c = O.coef17_to_matrix(O.coef(g, k2, eps, H, mt, psi, d, ae, ab, r * dab))
        al_psi, al_A = 0.42, 1.0
        E = O.free_energy(g, k2, eps, None, H, mt, psi + al_psi * d, ae, ab + al_A * r * dab)
        P = [np.polynomial.polynomial.polyval2d(al_psi, al_A, cc) for cc in
             (np.where(np.arange(5)[None, :] <= 0, c, 0), np.where(np.arange(5)[None, :] <= 2, c, 0), c)]
        e = np.abs(np.array(P) - E)
        assert e[0] > e[1] > e[2]
        errs.append(e[2])
    # 4th-order Taylor: the remainder scales like |dA|^5
    assert errs[1] < errs[0] * 1e-3


def test_pr_beta_and_reductions():
    """tests/at_reduction.py re-expressed on the oracle's helpers."""
    rng = np.random.default_rng(5)
    for n in (1, 7, 1000, 123457):
        x = rng.random(n)
        assert math.fsum(x) == pytest.approx(np.sum(x), rel=1e-12, abs=1e-10)
    g1 = rng.normal(size=50) + 1j * rng.normal(size=50)
    g0 = rng.normal(size=50) + 1j * rng.normal(size=50)
    num = np.sum((g1.conj() * (g1 - g0)).real)
    den = np.sum(np.abs(g0) ** 2)
    assert O.pr_beta(g1, g0) == pytest.approx(max(num / den, 0.0), rel=1e-12)
    assert O.pr_beta(g0, g0) == 0.0
    assert O.pr_beta(np.zeros(3), np.zeros(3)) == 0.0  # 0/0 -> fmax(NaN, 0) = 0


def test_alpha_min_roots():
    # cubic derivative of c: roots known
    c = np.array([0.0, -6.0, 5.5, -2.0 / 3.0 * 1.5 * 0, 0.25])  # P'(a) = -6 + 11 a + a^3
    a = O.alpha_psi_min(c)
    assert -6 + 11 * a + a ** 3 == pytest.approx(0.0, abs=1e-12)
    assert a >= 0
    c2 = np.zeros((5, 5)); c2[2, 0] = 1.0; c2[1, 0] = -2.0; c2[0, 2] = 3.0; c2[0, 1] = -3.0
    ap, aA = O.alpha_min(c2)
    assert ap == pytest.approx(1.0, abs=1e-6) and aA == pytest.approx(0.5, abs=1e-6)


# ---------------------------------------------------------------------------
# helpers, init, vortices
# ---------------------------------------------------------------------------

def test_rand_hash_brute_force():
    def h(s):
        s = ((s ^ 61) ^ (s >> 16)) & 0xffffffff
        s = (s * 9) & 0xffffffff
        s = s ^ (s >> 4)
        s = (s * 0x27d4eb2d) & 0xffffffff
        return s ^ (s >> 15)
    ns = [0, 1, 2, 12345, 2 ** 31 - 1, 2 ** 32 - 1, 987654321]
    assert [int(v) for v in O.rand_hash(np.array(ns, dtype=np.uint32))] == [h(s) for s in ns]
    for n, t in ((0, 1), (17, 3), (4000000000 % 2 ** 32, 77)):
        assert O.rand_1(np.uint32(n), t) == h((71 * n + 9887 * t) & 0xffffffff) / 2.0 ** 32
        assert O.rand_2(np.uint32(n), t) == h((73 * n + 9901 * t + 1) & 0xffffffff) / 2.0 ** 32


def test_residual_rule():
    assert O.scaled_residual(np.array([0.9999e-6]), 1e-6) < 10000
    assert O.scaled_residual(np.array([1.0001e-6]), 1e-6) >= 10000
    assert O.scaled_residual(np.array([5.0]), 1e-6) == 100000000
    assert O.converged(9999) and not O.converged(10000)
    assert np.float32(1e-4) * np.float32(9999) < 1 and not np.float32(1e-4) * np.float32(10000) < 1


def test_random_psi_legacy_stream():
    psi = O.random_psi(4, level=1.0, seed=0)
    assert abs(psi[0]) == pytest.approx(1.0 - 0.5488135039273248, rel=1e-15)
    assert np.all(np.abs(psi) <= 1.0)


@pytest.mark.parametrize("s", [1, -1])
def test_vortex_imprint(s):
    """A linear vortex psi = (x - x0) + 0.3 (y - y0) + i s (y - y0), A = 0,
    H = 0 winds once around the cell holding (x0, y0); the detector reports
    that cell with vorticity -s and triangulates (x0, y0) exactly. (The
    zero lines are skewed on purpose: the reference's find_intersection
    rejects exactly perpendicular lines, vortex_detector.py:88-101.)"""
    g = O.Geom(17, 15, 0.5, 0.5)
    x0, y0 = 3.3, 4.1
    x = g.dx * np.arange(g.Nx)[None, :]
    y = g.dy * np.arange(g.Ny)[:, None]
    psi = ((x - x0) + 0.3 * (y - y0) + 1j * s * (y - y0)).ravel()
    ab = np.zeros(g.Nab)
    cells, vs = O.vortex_cells(g, psi, ab, 0.0)
    assert list(cells) == [int(x0 / g.dx) + g.Nxc * int(y0 / g.dy)]
    assert np.round(vs[0]) == -s
    vx, vy, vv = O.vortex_detector(g, psi, ab, 0.0)
    assert vx[0] == pytest.approx(x0, abs=1e-12) and vy[0] == pytest.approx(y0, abs=1e-12)
    assert vv[0] == -s


# ---------------------------------------------------------------------------
# fixed vortices and phase lock (td.py:118-154, 205-216, 256-266, 318-325)
# ---------------------------------------------------------------------------

def _chi_x(g, rng):
    """A gauge function of x only: its gradient lives on the a links alone,
    which are all inside the first N entries the vpi hooks touch."""
    cx = rng.normal(size=g.Nx)
    chi = np.tile(cx, (g.Ny, 1))
    da = np.tile((cx[1:] - cx[:-1]) / g.dx, (g.Ny, 1))
    return chi.ravel(), np.concatenate([da.ravel(), np.zeros(g.Nb)])


def test_fixed_vortex_psi_hooks_gauge_covariant():
    """A psi step with vpi = grad(chi) added around the solve equals, by
    discrete gauge covariance, the plain step times exp(i chi); the
    post-processing restores A. A sign error in either hook fails it."""
    rng, g, mt, psi, ab, _, em = _case(61, Nx=9, Ny=8)
    chi, grad = _chi_x(g, rng)
    st1 = O.OracleState(g, psi, ab, kappa=np.inf, H=0.1, eps=1.0, eps_map=em, mt=mt)
    st2 = O.OracleState(g, psi * np.exp(1j * chi), ab, kappa=np.inf, H=0.1, eps=1.0, eps_map=em, mt=mt, vpi=grad)
    st1.td(dt=0.1, Nt=2)
    st2.td(dt=0.1, Nt=2)
    assert st2.sweeps == st1.sweeps
    assert np.allclose(st2.psi, st1.psi * np.exp(1j * chi), atol=1e-11)
    assert np.allclose(st2.ab, ab, rtol=0, atol=1e-14)


def test_fixed_vortex_hooks_touch_first_N_entries_only():
    """xpy_r/xmy_r run over N = number of nodes entries of the Nab-long
    buffers: vpi values beyond index N never reach the psi solve."""
    rng, g, mt, psi, ab, _, em = _case(62, Nx=8, Ny=9)
    vpi = np.zeros(g.Nab)
    vpi[g.N:] = rng.normal(size=g.Nab - g.N)
    st1 = O.OracleState(g, psi, ab, kappa=np.inf, eps_map=em, mt=mt)
    st2 = O.OracleState(g, psi, ab, kappa=np.inf, eps_map=em, mt=mt, vpi=vpi)
    st1.td(dt=0.1, Nt=2)
    st2.td(dt=0.1, Nt=2)
    assert np.array_equal(st2.psi, st1.psi) and np.array_equal(st2.ab, st1.ab)
    assert np.array_equal(st2.vpi, vpi)


def test_fixed_vortex_A_solve_fixed_source():
    """With vpi set, every A sweep reads vpi + A^n: the converged solve is a
    fixed point of the sweep map with that source (with the aliased source
    A^(2 floor(j/2)) it would be a fixed point with source A^{n+1} instead),
    and vpi + A is conserved on the first N entries across the solve."""
    rng, g, mt, psi, ab, _, _ = _case(63, Nx=9, Ny=9, mask=False)
    vpi = 0.3 * rng.normal(size=g.Nab)
    st = O.OracleState(g, psi, ab, kappa=2.0, sigma=10.0, H=0.1, stop_A=1e-13, vpi=vpi)
    st.td_A_solve(0.1)
    N = g.N
    src = vpi.copy()
    src[:N] = vpi[:N] + ab[:N]
    again, _, _ = O.a_jacobi_sweep(g, 0.1, st.kappa2, st.rho, st.H, None, psi, src, ab, st.ab, stop_eps=1e-13)
    assert np.max(np.abs(again - st.ab)) < 1e-11
    alias, _, _ = O.a_jacobi_sweep(g, 0.1, st.kappa2, st.rho, st.H, None, psi, st.ab, ab, st.ab, stop_eps=1e-13)
    assert np.max(np.abs(alias - st.ab)) > 1e-6
    assert np.allclose(st.vpi[:N] + st.ab[:N], vpi[:N] + ab[:N], rtol=0, atol=1e-13)
    assert np.array_equal(st.vpi[N:], vpi[N:])


def test_phase_lock_after_solve_and_order_parameter_hooks():
    """The lock takes |psi| at the listed nodes after each psi solve (the
    unlocked step's magnitudes, the other nodes untouched); eqn =
    'order_parameter' adds vpi once and removes it after every step
    (td.py:221-231), so A drifts by -(Nt - 1) vpi."""
    rng, g, mt, psi, ab, _, em = _case(64, Nx=8, Ny=8, mask=False)
    ns = np.array([3, 10, 27, 40])
    st1 = O.OracleState(g, psi, ab, kappa=np.inf, eps_map=em)
    st2 = O.OracleState(g, psi, ab, kappa=np.inf, eps_map=em, lock_ns=ns)
    st1.td(dt=0.1, Nt=1)
    st2.td(dt=0.1, Nt=1)
    other = np.setdiff1d(np.arange(g.N), ns)
    assert np.array_equal(st2.psi[other], st1.psi[other])
    assert np.all(st2.psi[ns].imag == 0) and np.all(st2.psi[ns].real >= 0)
    assert np.allclose(st2.psi[ns].real, np.abs(st1.psi[ns]), rtol=1e-15)
    vpi = 0.2 * rng.normal(size=g.Nab)
    st3 = O.OracleState(g, psi, ab, kappa=np.inf, eps_map=em, vpi=vpi)
    st3.td_order_parameter(dt=0.1, Nt=4)
    N = g.N
    assert np.allclose(st3.ab[:N], ab[:N] - 3 * vpi[:N], atol=1e-13)
    assert np.array_equal(st3.ab[N:], ab[N:])
