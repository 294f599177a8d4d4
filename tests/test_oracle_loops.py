"""Pins the oracle's host loops (OracleState.td_psi_solve, td_A_solve, cg)
on the CPU, against what the mathematics fixes rather than against a
re-typed copy of the loops:

* a converged Jacobi solve satisfies its backward-Euler equation: psi^{n+1}
  is a fixed point of the psi sweep with right-hand side psi^n, and on
  interior nodes it solves the textbook covariant scheme
  (psi^{n+1} - psi^n)/dt = Laplacian_A psi^{n+1} + (eps - |psi^n|^2) psi^{n+1}
  (svirl/cuda/td.h:4-133, svirl/solvers/td.py:157-218); A^{n+1} solves
  sigma (A^{n+1} - A^n)/dt = J_s(psi^{n+1}, A^{n+1}) - J(A^{n+1}) with the
  supercurrent and current observables, whose psi is the one the psi solve
  of the same step produced (td.h:310-463, td.py:252-325, 342-366);
* the A solve's source schedule A^(2 floor(j/2)) (td.py:264-303), replayed
  sweep by sweep against OracleState.td_A_solve itself;
* CG (cg.py:238-322, 452-551): the first step is a positive multiple of
  -dF/dpsi (and -dF/dA) and lands on the line minimum; every later step
  lies in span{-g_k, previous step} with the ratio that the Polak-Ribiere
  beta of the gradients at the visited states gives; energies do not rise;
  a second cg() call restarts from steepest descent when kappa = inf (the
  direction is zero-filled, cg.py:264) and keeps the previous call's
  direction and its last beta when kappa < inf (beta is recomputed only for
  i > 0, cg.py:481-483).
"""
import numpy as np
import pytest

from oracle import gl_oracle as O


def _state(seed, kappa, Nx=9, Ny=8, mask=False, epsmap=False, H=0.15, sigma=3.0, **kw):
    rng = np.random.default_rng(seed)
    g = O.Geom(Nx, Ny, 0.5, 0.45)
    mt = (rng.random(g.Nc) > 0.2) if mask else None
    psi = (0.3 + 0.7 * rng.random(g.N)) * np.exp(2j * np.pi * rng.random(g.N))
    ab = O.symmetric_gauge(g, H) + 0.2 * rng.normal(size=g.Nab)
    em = (1.0 + 0.2 * rng.normal(size=g.N)) if epsmap else None
    st = O.OracleState(g, psi, ab, kappa=kappa, sigma=sigma, H=H, eps=0.9, eps_map=em, mt=mt, **kw)
    return rng, g, st


def _interior_laplacian(g, P, ab):
    """Covariant 5-point Laplacian on interior nodes, U = exp(-i d A) on
    the link to the right / up neighbour (the sign that makes the scheme
    gauge covariant)."""
    a, b = g.a2(ab), g.b2(ab)
    c = P[1:-1, 1:-1]
    right = np.exp(-1j * g.dx * a[1:-1, 1:]) * P[1:-1, 2:]
    left = np.exp(1j * g.dx * a[1:-1, :-1]) * P[1:-1, :-2]
    up = np.exp(-1j * g.dy * b[1:, 1:-1]) * P[2:, 1:-1]
    down = np.exp(1j * g.dy * b[:-1, 1:-1]) * P[:-2, 1:-1]
    return (right - 2.0 * c + left) * g.idx2 + (up - 2.0 * c + down) * g.idy2


# ---------------------------------------------------------------------------
# TD solves
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("mask", [False, True])
def test_psi_solve_satisfies_backward_euler(mask):
    rng, g, st = _state(1, np.inf, Nx=11, Ny=10, mask=mask, epsmap=True, stop_psi=1e-12)
    psi0 = st.psi.copy()
    dt = 0.1
    st.td_psi_solve(dt)
    psi1 = st.psi
    # a fixed point of the sweep with right-hand side psi^n
    again, _, _ = O.psi_jacobi_sweep(g, dt, st.eps_scalar, st.eps_map, st.mt, st.ab, psi0, psi1, stop_eps=1e-12)
    assert np.max(np.abs(again - psi1)) < 1e-11
    # ... and not with the right-hand side psi^{n+1} (a loop that refreshed it)
    other, _, _ = O.psi_jacobi_sweep(g, dt, st.eps_scalar, st.eps_map, st.mt, st.ab, psi1, psi1, stop_eps=1e-12)
    assert np.max(np.abs(other - psi1)) > 1e-4
    if not mask:
        P0, P1 = g.psi2(psi0), g.psi2(psi1)
        E = st.eps_map.reshape(g.Ny, g.Nx)[1:-1, 1:-1]
        res = (P1[1:-1, 1:-1] - P0[1:-1, 1:-1]) / dt - _interior_laplacian(g, P1, st.ab) \
            - (E - np.abs(P0[1:-1, 1:-1]) ** 2) * P1[1:-1, 1:-1]
        assert np.max(np.abs(res)) < 1e-9 * np.max(np.abs(P1))


def test_A_solve_satisfies_backward_euler_with_new_psi():
    rng, g, st = _state(2, 2.0, Nx=10, Ny=9, sigma=40.0, stop_psi=1e-12, stop_A=1e-12)
    psi0, ab0 = st.psi.copy(), st.ab.copy()
    dt = 0.1
    st.td(dt=dt, Nt=1)
    psi1, ab1 = st.psi, st.ab
    assert not np.allclose(psi1, psi0)

    def residual(psi):
        js = O.supercurrent_density(g, None, psi, None, ab1)
        J = O.current_density(g, st.kappa2, st.H, None, ab1)
        return st.sigma * (ab1 - ab0) / dt - (js - J)

    a_int = np.zeros((g.Nya, g.Nxa), bool)
    a_int[1:-1, :] = True
    b_int = np.zeros((g.Nyb, g.Nxb), bool)
    b_int[:, 1:-1] = True
    inner = np.concatenate([a_int.ravel(), b_int.ravel()])
    r1 = residual(psi1)
    scale = np.max(np.abs(st.sigma * (ab1 - ab0) / dt))
    assert np.max(np.abs(r1[inner])) < 1e-8 * scale
    # the supercurrent source is psi^{n+1}, not psi^n
    assert np.max(np.abs(residual(psi0)[inner])) > 1e-3 * scale
    # the whole solve is a fixed point of the A sweep with source A^{n+1}
    again, _, _ = O.a_jacobi_sweep(g, dt, st.kappa2, st.rho, st.H, None, psi1, ab1, ab0, ab1, stop_eps=1e-12)
    assert np.max(np.abs(again - ab1)) < 1e-11


def test_aliasing_source_schedule_against_td_A_solve():
    """td.py:264-303: sweep j of the A solve reads its supercurrent source
    from A^(2 floor(j/2)). OracleState.td_A_solve's result equals an
    explicit replay of that schedule for its sweep count, and differs from
    replays with the source A^n or A^j."""
    rng, g, st = _state(50, 3.0, Nx=9, Ny=9, sigma=5.0, stop_A=1e-9)
    psi, ab0 = st.psi.copy(), st.ab.copy()
    nsw = st.td_A_solve(0.1)
    assert nsw >= 5

    def replay(source_of):
        A = [ab0.copy()]
        rhs = ab0.copy()
        for j in range(nsw):
            nxt, rhs, _ = O.a_jacobi_sweep(g, 0.1, st.kappa2, st.rho, st.H, None, psi, source_of(A, j), rhs, A[-1],
                                           stop_eps=1e-9)
            A.append(nxt)
        return A[-1]

    assert np.array_equal(st.ab, replay(lambda A, j: A[2 * (j // 2)]))
    assert not np.array_equal(st.ab, replay(lambda A, j: A[0]))
    assert not np.array_equal(st.ab, replay(lambda A, j: A[j]))


# ---------------------------------------------------------------------------
# CG
# ---------------------------------------------------------------------------

def _vdot(x, y):
    """Real inner product of two real or complex vectors."""
    return float(np.sum(x.real * y.real + x.imag * y.imag))


def _record(st):
    """Make st.cg record (psi, ab) after each iteration (its energy call)."""
    rec = []
    base = O.OracleState.free_energy

    def fe():
        rec.append((st.psi.copy(), st.ab.copy()))
        return base(st)

    st.free_energy = fe
    return rec


def _decompose(step, g, prev):
    """step = x (-g) + y prev by least squares; returns x, y and the
    relative residual."""
    if prev is None:
        x = -_vdot(step, g) / _vdot(g, g)
        res = step + x * g
        return x, 0.0, np.sqrt(_vdot(res, res) / _vdot(step, step))
    M = np.array([[_vdot(g, g), -_vdot(g, prev)], [-_vdot(g, prev), _vdot(prev, prev)]])
    rhs = np.array([-_vdot(step, g), _vdot(step, prev)])
    x, y = np.linalg.solve(M, rhs)
    res = step + x * g - y * prev
    return x, y, np.sqrt(_vdot(res, res) / _vdot(step, step))


def _grads(st, psi, ab):
    gp = O.jacobian_psi(st.g, st.kappa2, st.eps_scalar, st.eps_map, st.H, st.mt, psi, st.abei, ab)
    ga = O.jacobian_A(st.g, st.kappa2, st.H, st.mt, psi, st.abei, ab) if st.solveA else None
    return gp, ga


@pytest.mark.parametrize("kappa", [np.inf, 5.0])
def test_cg_first_step_is_line_minimum_along_minus_gradient(kappa):
    rng, g, st = _state(3, kappa, mask=True, convergence_rtol=0.0)
    psi0, ab0 = st.psi.copy(), st.ab.copy()
    gp, ga = _grads(st, psi0, ab0)
    st.cg(n_iter=1)
    x, _, res = _decompose(st.psi - psi0, gp, None)
    assert x > 0 and res < 1e-12
    parts = [(gp, x)]
    if st.solveA:
        xa, _, resa = _decompose(st.ab - ab0, ga, None)
        assert xa > 0 and resa < 1e-12
        parts.append((ga, xa))

    def F(s):
        p = psi0 - s[0] * x * gp
        a = ab0 - s[1] * parts[-1][1] * ga if st.solveA else ab0
        return O.free_energy(g, st.kappa2, st.eps_scalar, st.eps_map, st.H, st.mt, p, st.abei, a)

    F1 = F((1.0, 1.0))
    assert F1 < F((0.0, 0.0))
    for k in range(len(parts)):
        for f in (0.9, 1.1):
            s = [1.0, 1.0]
            s[k] = f
            assert F1 < F(tuple(s))
        # the directional derivative nearly vanishes at the step, relative to
        # its value at the start (kappa < inf: the coefficient kernel's
        # polynomial is exact in alpha_psi and 4th order in alpha_A)
        h = 1e-4

        def d(s0):
            lo, hi = list(s0), list(s0)
            lo[k] -= h
            hi[k] += h
            return (F(tuple(hi)) - F(tuple(lo))) / (2 * h)

        assert abs(d((1.0, 1.0))) < (1e-6 if k == 0 else 1e-3) * abs(d((0.0, 0.0)))


@pytest.mark.parametrize("kappa", [np.inf, 5.0])
def test_cg_energies_do_not_rise(kappa):
    rng, g, st = _state(4, kappa, mask=True, epsmap=False, convergence_rtol=0.0)
    E0 = st.free_energy()
    st.cg(n_iter=30)
    E = np.array([E0] + list(st.cg_energies))
    assert len(E) == 31
    assert np.all(np.diff(E) <= 1e-12 * np.abs(E[:-1]))
    assert E[-1] < E[0]


def _pr(gk, gprev):
    return max(_vdot(gk, gk - gprev) / _vdot(gprev, gprev), 0.0)


@pytest.mark.parametrize("kappa", [np.inf, 5.0])
def test_cg_steps_follow_polak_ribiere(kappa):
    """Each step is alpha_k (beta_k d_{k-1} - g_k): its decomposition on
    -g_k and the previous step has residual ~0, a positive -g_k part, and
    beta_k = y alpha_{k-1} / x equal to the Polak-Ribiere beta of the
    gradients at the visited states (separately for psi and A)."""
    rng, g, st = _state(5, kappa, mask=True, convergence_rtol=0.0)
    states = [(st.psi.copy(), st.ab.copy())]
    rec = _record(st)
    st.cg(n_iter=8)
    states += rec
    grads = [_grads(st, p, a) for p, a in states]
    checked = 0
    for part in ([0, 1] if st.solveA else [0]):
        alpha_prev, prev = None, None
        for k in range(len(states) - 1):
            step = states[k + 1][part] - states[k][part]
            gk = grads[k][part]
            x, y, res = _decompose(step, gk, prev)
            assert x > 0 and res < 1e-9, (part, k, res)
            if prev is not None:
                beta = y * alpha_prev / x
                want = _pr(gk, grads[k - 1][part])
                assert beta == pytest.approx(want, rel=1e-6, abs=1e-10), (part, k)
                checked += want > 1e-3
            alpha_prev, prev = x, step
    assert checked >= 4


@pytest.mark.parametrize("kappa", [np.inf, 5.0])
def test_cg_second_call_direction(kappa):
    """kappa = inf: the second call starts with steepest descent. kappa <
    inf: its first direction is beta_last d_last - g with beta_last the
    last beta of the first call (not the PR beta of the new gradient)."""
    rng, g, st = _state(6, kappa, mask=False, convergence_rtol=0.0)
    st.cg(n_iter=5)
    before = [(st.psi.copy(), st.ab.copy())]
    rec = _record(st)
    # the last two states of the first call
    st2 = _state(6, kappa, mask=False, convergence_rtol=0.0)[2]
    rec2 = _record(st2)
    st2.cg(n_iter=5)
    st.cg(n_iter=1)
    s_prev = (rec2[-1][0] - rec2[-2][0], rec2[-1][1] - rec2[-2][1])
    s_new = (rec[0][0] - before[0][0], rec[0][1] - before[0][1])
    g4 = _grads(st, *rec2[-2])   # gradient at the start of the first call's last step
    g3 = _grads(st, *rec2[-3])
    g5 = _grads(st, *before[0])  # gradient at the start of the second call
    for part in ([0, 1] if st.solveA else [0]):
        x, y, res = _decompose(s_new[part], g5[part], s_prev[part])
        assert x > 0 and res < 1e-9
        if not st.solveA:
            assert abs(y) < 1e-12 * abs(x)
        else:
            xp, _, _ = _decompose(s_prev[part], g4[part], rec2[-2][part] - rec2[-3][part])
            beta = y * xp / x
            stale = _pr(g4[part], g3[part])
            fresh = _pr(g5[part], g4[part])
            assert beta == pytest.approx(stale, rel=1e-6, abs=1e-10)
            assert abs(stale - fresh) > 1e-3 * max(stale, fresh)
