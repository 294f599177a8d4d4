"""Kernel-level parity of libglx (called through the C ABI) with the FP64
oracle, on seeded random inputs.

FP64 tolerances: elementwise results agree to a few ulp (the kernels keep the
reference's operand order and round each multiply and add separately; only
sin/cos/atan2 round differently), reductions to summation order. The
integer residual may differ by one count at most. FP32 (the FP32-mixed
form) is checked against the same oracle at float tolerance.
Grid widths cross the 63-column strip boundaries of the row-marching
kernels (Nx = 5, 63, 64, 127, 130).
"""
import ctypes

import numpy as np
import pytest

from oracle import gl_oracle as O

pytestmark = pytest.mark.gpu

nat = pytest.importorskip("svirl_amd._native")


@pytest.fixture(scope="module", autouse=True)
def _init():
    nat.init(0)


def D(a, dtype):
    return nat.DeviceArray.from_host(np.asarray(a), dtype)


def _types(prec):
    return (np.float64, np.complex128) if prec == nat.GLX_F64 else (np.float32, np.complex64)


def _tol(prec, scale=1.0):
    return (1e-12 if prec == nat.GLX_F64 else 2e-5) * scale


def case(seed, Nx, Ny, mask=True, epsmap=True):
    rng = np.random.default_rng(seed)
    dx, dy = 0.5, 0.45
    g = O.Geom(Nx, Ny, dx, dy)
    mt = (rng.random(g.Nc) > 0.25) if mask else None
    psi = (0.3 + rng.random(g.N)) * np.exp(2j * np.pi * rng.random(g.N))
    ab = O.symmetric_gauge(g, 0.3) + 0.2 * rng.normal(size=g.Nab)
    em = (1.0 + 0.2 * rng.normal(size=g.N)) if epsmap else None
    return rng, g, mt, psi, ab, em


def grid(g):
    return nat.Grid(g.Nx, g.Ny, g.dx, g.dy)


SHAPES = [(5, 4), (63, 7), (64, 9), (127, 5), (130, 33)]
PRECS = [nat.GLX_F64, nat.GLX_F32]


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("variant", [0, 1, 2])
def test_psi_sweep(prec, shape, variant):
    rs, cs = _types(prec)
    rng, g, mt, psi, ab, em = case(10 * variant + shape[0], *shape, mask=variant != 0, epsmap=variant == 2)
    rhs = psi * (1 + 0.1 * rng.random(g.N))
    lc = 0.3 if variant == 2 else 0.0
    jstep, t = 0, 7
    # the oracle sees the same rounded inputs as the kernel
    psi_r, rhs_r, ab_r = psi.astype(cs), rhs.astype(cs), ab.astype(rs)
    em_r = None if em is None else em.astype(rs)
    nxt, rhs_o, r = O.psi_jacobi_sweep(g, 0.1, 0.9, em_r, mt, ab_r.astype(float), rhs_r.astype(complex),
                                       psi_r.astype(complex), lc, jstep, t, 1e-6)
    d_psi, d_rhs, d_ab = D(psi_r, cs), D(rhs_r, cs), D(ab_r, rs)
    d_next = nat.DeviceArray.zeros(g.N, cs)
    d_em = D(em_r, rs) if em is not None else None
    d_mt = D(mt.astype(np.uint8), np.uint8) if mt is not None else None
    r2 = nat.DeviceArray.zeros(1, np.int32)
    nat.call("glx_td_psi_sweep", prec, grid(g), 0.1, 0.9, nat.ptr(d_em), nat.ptr(d_mt), d_ab.ptr, d_rhs.ptr,
             d_psi.ptr, d_next.ptr, lc, jstep, t, 1e-6, r2.ptr)
    out = d_next.get()
    assert np.max(np.abs(out - nxt)) <= _tol(prec, 10)
    if prec == nat.GLX_F64:
        assert abs(int(r2.get()[0]) - r) <= 1
        if lc > 0:
            assert np.max(np.abs(d_rhs.get() - rhs_o)) <= 1e-13
    else:
        assert abs(int(r2.get()[0]) - r) <= max(2, 2e-4 * r)


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("variant", [0, 1, 2])
def test_a_sweep(prec, shape, variant):
    rs, cs = _types(prec)
    rng, g, mt, psi, ab, _ = case(20 + 10 * variant + shape[0], *shape, mask=variant != 0, epsmap=False)
    src = (ab + 0.01 * rng.random(g.Nab)).astype(rs)
    rhs = (ab + 0.02 * rng.random(g.Nab)).astype(rs)
    psi_r, ab_r = psi.astype(cs), ab.astype(rs)
    lc = 0.2 if variant == 2 else 0.0
    k2, rho, H = 25.0, 1 / 20.0, 0.3
    nxt, rhs_o, r = O.a_jacobi_sweep(g, 0.1, k2, rho, H, mt, psi_r.astype(complex), src.astype(float),
                                     rhs.astype(float), ab_r.astype(float), lc, 0, 5, 1e-6)
    d_psi, d_src, d_rhs, d_ab = D(psi_r, cs), D(src, rs), D(rhs, rs), D(ab_r, rs)
    d_mt = D(mt.astype(np.uint8), np.uint8) if mt is not None else None
    # variant 1 exercises the aliased source: the source buffer is the output buffer
    if variant == 1:
        d_next = D(src, rs)
        d_src_arg = d_next
    else:
        d_next = nat.DeviceArray.zeros(g.Nab, rs)
        d_src_arg = d_src
    r2 = nat.DeviceArray.zeros(1, np.int32)
    nat.call("glx_td_a_sweep", prec, grid(g), 0.1, k2, rho, H, nat.ptr(d_mt), d_psi.ptr, d_src_arg.ptr, d_rhs.ptr,
             d_ab.ptr, d_next.ptr, lc, 0, 5, 1e-6, r2.ptr)
    out = d_next.get()
    assert np.max(np.abs(out - nxt)) <= _tol(prec, 20)
    if prec == nat.GLX_F64:
        assert abs(int(r2.get()[0]) - r) <= 1
        if lc > 0:
            assert np.max(np.abs(d_rhs.get() - rhs_o)) <= 1e-13


def _td_params(dt, eps, k2, rho, H, stop_psi, stop_A, solveA, eqn=0, rand_t=1):
    p = nat.TDParams()
    p.dt, p.eps, p.kappa2, p.rho, p.H = dt, eps, k2, rho, H
    p.stop_psi, p.stop_A = stop_psi, stop_A
    p.langevin_psi = p.langevin_A = 0.0
    p.rand_t, p.solveA, p.max_sweeps, p.eqn = rand_t, 1 if solveA else 0, 1024, eqn
    return p


@pytest.mark.parametrize("shape,kappa,mask,epsmap,nsteps,wg", [
    ((33, 29), 5.0, True, True, 12, 0),
    ((70, 41), np.inf, False, False, 15, 0),
    ((70, 41), np.inf, False, False, 15, 3),
    ((130, 17), 3.0, True, False, 6, 0),
    ((9, 9), 2.0, False, True, 25, 1),
    ((2, 2), 2.0, False, False, 4, 0),   # the smallest grids the C ABI takes:
    ((3, 2), 2.0, False, True, 4, 0),    # one cell; odd A sweep counts, so the
    ((2, 3), np.inf, False, False, 4, 0),  # three-sweep A pass runs too
])
def test_td_run_matches_oracle(shape, kappa, mask, epsmap, nsteps, wg):
    """Persistent multi-step kernel vs td.py's loops: identical sweep
    counts per step and the same fields (FP64)."""
    rng, g, mt, psi, ab, em = case(99 + shape[0], *shape, mask=mask, epsmap=epsmap)
    ab = O.symmetric_gauge(g, 0.25)
    psi = np.where(O.mask_at_nodes(g, mt), psi, 0)
    st = O.OracleState(g, psi, ab, kappa=kappa, sigma=100.0, H=0.25, eps=1.0, eps_map=em, mt=mt)
    st.td(dt=0.1, Nt=nsteps)
    P = [D(psi, np.complex128), nat.DeviceArray.zeros(g.N, np.complex128), nat.DeviceArray.zeros(g.N, np.complex128)]
    Q = [D(ab, np.float64), nat.DeviceArray.zeros(g.Nab, np.float64), nat.DeviceArray.zeros(g.Nab, np.float64)]
    Pp = (ctypes.c_void_p * 3)(*[b.ptr for b in P])
    Qp = (ctypes.c_void_p * 3)(*[b.ptr for b in Q])
    pc, qc = ctypes.c_int32(0), ctypes.c_int32(0)
    sw = (ctypes.c_int32 * (2 * nsteps))()
    d_em = D(em, np.float64) if em is not None else None
    d_mt = D(mt.astype(np.uint8), np.uint8) if mt is not None else None
    prm = _td_params(0.1, 0.0 if em is not None else 1.0, st.kappa2, st.rho, 0.25, 1e-6, 1e-6, st.solveA)
    nat.call("glx_td_run", nat.GLX_F64, grid(g), ctypes.byref(prm), nat.ptr(d_em), nat.ptr(d_mt), Pp,
             ctypes.byref(pc), Qp, ctypes.byref(qc), None, None, nsteps, sw, wg, 256)
    sweeps = np.frombuffer(sw, dtype=np.int32).reshape(nsteps, 2)
    assert [tuple(x) for x in sweeps.tolist()] == st.sweeps
    assert np.max(np.abs(P[pc.value].get() - st.psi)) <= 1e-11
    assert np.max(np.abs(Q[qc.value].get() - st.ab)) <= 1e-11


def _td_run_f32(g, psi, ab, em, mt, prm, nsteps, wg=0):
    P = [D(psi, np.complex64), nat.DeviceArray.zeros(g.N, np.complex64), nat.DeviceArray.zeros(g.N, np.complex64)]
    Q = [D(ab, np.float32), nat.DeviceArray.zeros(g.Nab, np.float32), nat.DeviceArray.zeros(g.Nab, np.float32)]
    Pp = (ctypes.c_void_p * 3)(*[b.ptr for b in P])
    Qp = (ctypes.c_void_p * 3)(*[b.ptr for b in Q])
    pc, qc = ctypes.c_int32(0), ctypes.c_int32(0)
    sw = (ctypes.c_int32 * (2 * nsteps))()
    d_em = D(em, np.float32) if em is not None else None
    d_mt = D(mt.astype(np.uint8), np.uint8) if mt is not None else None
    nat.call("glx_td_run", nat.GLX_F32, grid(g), ctypes.byref(prm), nat.ptr(d_em), nat.ptr(d_mt), Pp,
             ctypes.byref(pc), Qp, ctypes.byref(qc), None, None, nsteps, sw, wg, 256)
    sweeps = np.frombuffer(sw, dtype=np.int32).reshape(nsteps, 2).copy()
    return P[pc.value].get().astype(np.complex128), Q[qc.value].get().astype(np.float64), sweeps


@pytest.mark.parametrize("shape,kappa,mask,epsmap,nsteps", [
    ((33, 29), 5.0, True, True, 12),
    ((70, 41), np.inf, False, False, 15),
    ((130, 17), 3.0, True, False, 6),
])
def test_td_run_fp32_mixed(shape, kappa, mask, epsmap, nsteps):
    """FP32-mixed persistent run (A stored relative to the symmetric gauge
    inside the run) against the FP64 oracle: fields within float
    resolution, sweep counts within one of the FP64 counts."""
    rng, g, mt, psi, ab, em = case(99 + shape[0], *shape, mask=mask, epsmap=epsmap)
    ab = O.symmetric_gauge(g, 0.25) + 0.05 * rng.normal(size=g.Nab)
    psi = np.where(O.mask_at_nodes(g, mt), psi, 0)
    psi32 = psi.astype(np.complex64).astype(np.complex128)
    ab32 = ab.astype(np.float32).astype(np.float64)
    em32 = em.astype(np.float32).astype(np.float64) if em is not None else None
    st = O.OracleState(g, psi32, ab32, kappa=kappa, sigma=100.0, H=0.25, eps=1.0, eps_map=em32, mt=mt)
    st.td(dt=0.1, Nt=nsteps)
    prm = _td_params(0.1, 0.0 if em is not None else 1.0, st.kappa2, st.rho, 0.25, 1e-6, 1e-6, st.solveA)
    p, a, sweeps = _td_run_f32(g, psi32, ab32, em32, mt, prm, nsteps)
    ref = np.array(st.sweeps)
    assert np.all(np.abs(sweeps - ref) <= 1), (sweeps.tolist(), st.sweeps)
    assert np.max(np.abs(p - st.psi)) <= 2e-5
    assert np.max(np.abs(a - st.ab)) <= 2e-5


def test_td_run_fp32_psi_only_leaves_A_untouched():
    """kappa = inf (no A solve): the FP32 run uses the plain form, so the
    vector potential comes back bit for bit."""
    rng, g, mt, psi, ab, em = case(5, 70, 41, mask=False, epsmap=False)
    ab32 = O.symmetric_gauge(g, 0.25).astype(np.float32).astype(np.float64)
    prm = _td_params(0.1, 1.0, -1.0, 1.0, 0.25, 1e-6, 1e-6, False)
    p, a, sweeps = _td_run_f32(g, psi.astype(np.complex64).astype(np.complex128), ab32, None, None, prm, 3)
    assert np.array_equal(a, ab32)
    assert np.all(sweeps[:, 1] == 0) and np.all(sweeps[:, 0] > 0)


def test_td_run_fp32_large_potential():
    """|A| ~ 26 at the edges (float ulp 1.9e-6, above the 1e-6 stop
    threshold): the FP32-mixed run still converges in a few sweeps per step
    and tracks the FP64 oracle."""
    g = O.Geom(513, 41, 0.5, 0.5)
    H = 0.4
    psi = O.random_psi(g.N, 1.0, seed=5)
    ab = O.symmetric_gauge(g, H)
    st = O.OracleState(g, psi, ab, kappa=5.0, sigma=100.0, H=H)
    st.td(dt=0.1, Nt=4)
    psi32 = psi.astype(np.complex64).astype(np.complex128)
    prm = _td_params(0.1, 1.0, st.kappa2, st.rho, H, 1e-6, 1e-6, True)
    p, a, sweeps = _td_run_f32(g, psi32, ab.astype(np.float32).astype(np.float64), None, None, prm, 4)
    ref = np.array(st.sweeps)
    assert np.all(sweeps[:, 1] <= ref[:, 1] + 1), (sweeps.tolist(), st.sweeps)
    assert np.all(np.abs(sweeps[:, 0] - ref[:, 0]) <= 1), (sweeps.tolist(), st.sweeps)
    assert np.max(np.abs(p - st.psi)) <= 5e-5
    assert np.max(np.abs(a - st.ab)) <= 2e-5 * (1 + np.max(np.abs(st.ab)))


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("shape", [(5, 4), (64, 9), (130, 33)])
def test_observables(prec, shape):
    rs, cs = _types(prec)
    rng, g, mt, psi, ab, em = case(7 + shape[0], *shape)
    ae = (0.1 * rng.normal(size=g.Nab)).astype(rs)
    psi_r, ab_r = psi.astype(cs), ab.astype(rs)
    em_r = em.astype(rs)
    d_psi, d_ab, d_ae, d_em = D(psi_r, cs), D(ab_r, rs), D(ae, rs), D(em_r, rs)
    d_mt = D(mt.astype(np.uint8), np.uint8)
    P, A, AE, EM = psi_r.astype(complex), ab_r.astype(float), ae.astype(float), em_r.astype(float)
    tol = 1e-12 if prec == nat.GLX_F64 else 2e-4
    for k2 in (-1.0, 16.0):
        F = ctypes.c_double()
        nat.call("glx_free_energy", prec, grid(g), k2, 0.0, d_em.ptr, 0.2, d_mt.ptr, d_psi.ptr, d_ae.ptr, d_ab.ptr,
                 ctypes.byref(F))
        Fo = O.free_energy(g, k2, 0.0, EM, 0.2, mt, P, AE, A)
        assert F.value == pytest.approx(Fo, rel=tol, abs=tol)
    out = nat.DeviceArray.zeros(g.Nab, rs)
    nat.call("glx_current_density", prec, grid(g), 16.0, 0.2, d_ae.ptr, d_ab.ptr, out.ptr)
    J = O.current_density(g, 16.0, 0.2, AE, A)
    assert np.max(np.abs(out.get() - J)) <= tol * 100 * (1 + np.max(np.abs(J)))
    nat.call("glx_supercurrent_density", prec, grid(g), d_mt.ptr, d_psi.ptr, d_ae.ptr, d_ab.ptr, out.ptr)
    assert np.max(np.abs(out.get() - O.supercurrent_density(g, mt, P, AE, A))) <= tol * 10
    Bd = nat.DeviceArray.zeros(g.Nc, rs)
    nat.call("glx_magnetic_field", prec, grid(g), d_ae.ptr, d_ab.ptr, Bd.ptr)
    assert np.max(np.abs(Bd.get() - O.magnetic_field(g, AE, A))) <= tol * 100


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("shape", [(5, 4), (64, 9), (130, 33)])
@pytest.mark.parametrize("mask,abei", [(True, True), (False, False)])
def test_bulk_observables_equal_separate_kernels(prec, shape, mask, abei):
    """The fused B, J, J_s pass gives the three separate kernels' values to
    the bit (and so the oracle's, by test_observables), with any output
    left out."""
    rs, cs = _types(prec)
    rng, g, mt, psi, ab, em = case(17 + shape[0], *shape, mask=mask)
    ae = (0.1 * rng.normal(size=g.Nab)).astype(rs) if abei else None
    d_psi, d_ab = D(psi.astype(cs), cs), D(ab.astype(rs), rs)
    d_ae = D(ae, rs) if abei else None
    d_mt = D(mt.astype(np.uint8), np.uint8) if mask else None
    k2, H = 16.0, 0.2
    B0, J0, S0 = (nat.DeviceArray.zeros(g.Nc, rs), nat.DeviceArray.zeros(g.Nab, rs),
                  nat.DeviceArray.zeros(g.Nab, rs))
    nat.call("glx_magnetic_field", prec, grid(g), nat.ptr(d_ae), d_ab.ptr, B0.ptr)
    nat.call("glx_current_density", prec, grid(g), k2, H, nat.ptr(d_ae), d_ab.ptr, J0.ptr)
    nat.call("glx_supercurrent_density", prec, grid(g), nat.ptr(d_mt), d_psi.ptr, nat.ptr(d_ae), d_ab.ptr, S0.ptr)
    B1, J1, S1 = (nat.DeviceArray.zeros(g.Nc, rs), nat.DeviceArray.zeros(g.Nab, rs),
                  nat.DeviceArray.zeros(g.Nab, rs))
    nat.call("glx_bulk_observables", prec, grid(g), k2, H, nat.ptr(d_mt), d_psi.ptr, nat.ptr(d_ae), d_ab.ptr,
             B1.ptr, J1.ptr, S1.ptr)
    assert np.array_equal(B1.get(), B0.get())
    assert np.array_equal(J1.get(), J0.get())
    assert np.array_equal(S1.get(), S0.get())
    # J alone (psi may be NULL then); the other buffers stay untouched
    J2, S2 = nat.DeviceArray.zeros(g.Nab, rs), nat.DeviceArray.zeros(g.Nab, rs)
    nat.call("glx_bulk_observables", prec, grid(g), k2, H, None, None, nat.ptr(d_ae), d_ab.ptr, None, J2.ptr, None)
    assert np.array_equal(J2.get(), J0.get()) and not S2.get().any()
    with pytest.raises(nat.GLXError):
        nat.call("glx_bulk_observables", prec, grid(g), k2, H, None, None, None, d_ab.ptr, None, None, S2.ptr)


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("shape", [(6, 5), (64, 9), (130, 33)])
def test_cg_kernels(prec, shape):
    rs, cs = _types(prec)
    rng, g, mt, psi, ab, em = case(3 + shape[0], *shape)
    ae = (0.1 * rng.normal(size=g.Nab)).astype(rs)
    d = (0.3 * (rng.normal(size=g.N) + 1j * rng.normal(size=g.N))).astype(cs)
    dab = (0.05 * rng.normal(size=g.Nab)).astype(rs)
    psi_r, ab_r, em_r = psi.astype(cs), ab.astype(rs), em.astype(rs)
    P, A, AE, EM = psi_r.astype(complex), ab_r.astype(float), ae.astype(float), em_r.astype(float)
    d_psi, d_ab, d_ae, d_em, d_d, d_dab = D(psi_r, cs), D(ab_r, rs), D(ae, rs), D(em_r, rs), D(d, cs), D(dab, rs)
    d_mt = D(mt.astype(np.uint8), np.uint8)
    tol = 1e-12 if prec == nat.GLX_F64 else 3e-4
    k2, H = 9.0, 0.15
    G = nat.DeviceArray.zeros(g.N, cs)
    nat.call("glx_cg_jacobian_psi", prec, grid(g), k2, 0.0, d_em.ptr, H, d_mt.ptr, d_psi.ptr, d_ae.ptr, d_ab.ptr, G.ptr)
    jo = O.jacobian_psi(g, k2, 0.0, EM, H, mt, P, AE, A)
    assert np.max(np.abs(G.get() - jo)) <= tol * 10 * (1 + np.max(np.abs(jo)))
    GA = nat.DeviceArray.zeros(g.Nab, rs)
    nat.call("glx_cg_jacobian_A", prec, grid(g), k2, H, d_mt.ptr, d_psi.ptr, d_ae.ptr, d_ab.ptr, GA.ptr)
    jA = O.jacobian_A(g, k2, H, mt, P, AE, A)
    assert np.max(np.abs(GA.get() - jA)) <= tol * 10 * (1 + np.max(np.abs(jA)))
    c5 = (ctypes.c_double * 5)()
    nat.call("glx_cg_coef_psi", prec, grid(g), k2, 0.9, H, d_mt.ptr, d_psi.ptr, d_d.ptr, d_ae.ptr, d_ab.ptr, c5)
    co = O.coef_psi(g, k2, 0.9, H, mt, P, d.astype(complex), AE, A)
    assert np.allclose(np.array(c5[:]), co, rtol=tol * 10, atol=tol * 10 * np.max(np.abs(co)))
    c17 = (ctypes.c_double * 17)()
    nat.call("glx_cg_coef", prec, grid(g), k2, 0.9, H, d_mt.ptr, d_psi.ptr, d_d.ptr, d_ae.ptr, d_ab.ptr, d_dab.ptr, c17)
    co = O.coef(g, k2, 0.9, H, mt, P, d.astype(complex), AE, A, dab.astype(float))
    assert np.allclose(np.array(c17[:]), co, rtol=tol * 10, atol=tol * 10 * np.max(np.abs(co)))
    # Polak-Ribiere sums and axpy / axmy
    x = (rng.normal(size=g.N) + 1j * rng.normal(size=g.N)).astype(cs)
    y = (rng.normal(size=g.N) + 1j * rng.normal(size=g.N)).astype(cs)
    dx_, dy_ = D(x, cs), D(y, cs)
    out = (ctypes.c_double * 2)()
    nat.call("glx_pr_sums", prec, dx_.ptr, dy_.ptr, 2 * g.N, out)
    X, Y = x.astype(complex), y.astype(complex)
    assert out[0] == pytest.approx(np.sum(X.real * (X.real - Y.real) + X.imag * (X.imag - Y.imag)), rel=1e-6 if prec else 1e-12)
    assert out[1] == pytest.approx(np.sum(np.abs(Y) ** 2), rel=1e-6 if prec else 1e-12)
    z = nat.DeviceArray.zeros(g.N, cs)
    nat.call("glx_axpy", prec, dx_.ptr, dy_.ptr, z.ptr, 0.7, 2 * g.N)
    assert np.allclose(z.get(), 0.7 * X + Y, rtol=1e-6, atol=1e-6)
    nat.call("glx_axmy", prec, dx_.ptr, dy_.ptr, z.ptr, 0.7, 2 * g.N)
    assert np.allclose(z.get(), 0.7 * X - Y, rtol=1e-6, atol=1e-6)


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("shape", [(6, 5), (64, 9), (130, 33)])
@pytest.mark.parametrize("withA", [True, False])
def test_cg_direction_step_equals_separate_passes(prec, shape, withA):
    """glx_cg_direction_step (direction update fused into the coefficient
    pass) against glx_axmy_dev followed by glx_cg_coef(_psi): the same new
    directions and coefficients (FP64 to the bit); the device form gives
    glx_cg_line_step's alphas on the separately updated direction and leaves
    the state alone."""
    rs, cs = _types(prec)
    rng, g, mt, psi, ab, em = case(41 + shape[0], *shape)
    ae = (0.1 * rng.normal(size=g.Nab)).astype(rs)
    dold = (0.3 * (rng.normal(size=g.N) + 1j * rng.normal(size=g.N))).astype(cs)
    jac = (0.2 * (rng.normal(size=g.N) + 1j * rng.normal(size=g.N))).astype(cs)
    daold = (0.05 * rng.normal(size=g.Nab)).astype(rs)
    jacA = (0.04 * rng.normal(size=g.Nab)).astype(rs)
    k2, H, eps = (9.0 if withA else -1.0), 0.15, 0.9
    beta = nat.DeviceArray.from_host(np.array([0.37, 0.61, 0.0, 0.0]), np.float64)
    d_mt = D(mt.astype(np.uint8), np.uint8)
    d_ae = D(ae, rs)
    nc = 17 if withA else 5

    def run(fused, device):
        d_psi, d_ab = D(psi.astype(cs), cs), D(ab.astype(rs), rs)
        d_do, d_jac, d_dao, d_jA = D(dold, cs), D(jac, cs), D(daold, rs), D(jacA, rs)
        out_d, out_a = nat.DeviceArray.zeros(g.N, cs), nat.DeviceArray.zeros(g.Nab, rs)
        alpha = nat.DeviceArray.zeros(2, np.float64)
        c = None if device else (ctypes.c_double * nc)()
        if fused:
            nat.call("glx_cg_direction_step", prec, grid(g), k2, eps, H, d_mt.ptr, d_psi.ptr, d_ae.ptr, d_ab.ptr,
                     d_do.ptr, d_jac.ptr, out_d.ptr, d_dao.ptr if withA else None, d_jA.ptr if withA else None,
                     out_a.ptr if withA else None, beta.ptr, c, None if c is not None else alpha.ptr)
            dn, dan = out_d, out_a
        else:
            nat.call("glx_axmy_dev", prec, d_do.ptr, d_jac.ptr, d_do.ptr, beta.ptr, 2 * g.N)
            if withA:
                nat.call("glx_axmy_dev", prec, d_dao.ptr, d_jA.ptr, d_dao.ptr, ctypes.c_void_p(beta.ptr + 8), g.Nab)
            dn, dan = d_do, d_dao
            if device:
                nat.call("glx_cg_line_step", prec, grid(g), k2, eps, H, d_mt.ptr, d_psi.ptr, dn.ptr, d_ae.ptr,
                         d_ab.ptr, dan.ptr if withA else None, alpha.ptr)
            elif withA:
                nat.call("glx_cg_coef", prec, grid(g), k2, eps, H, d_mt.ptr, d_psi.ptr, dn.ptr, d_ae.ptr, d_ab.ptr,
                         dan.ptr, c)
            else:
                nat.call("glx_cg_coef_psi", prec, grid(g), k2, eps, H, d_mt.ptr, d_psi.ptr, dn.ptr, d_ae.ptr,
                         d_ab.ptr, c)
        return (dn.get(), dan.get() if withA else None, None if c is None else np.array(c[:]),
                alpha.get(), d_psi.get(), d_ab.get())

    for device in (False, True):
        a, b = run(True, device), run(False, device)
        f64 = prec == nat.GLX_F64
        same = (lambda x, y: np.array_equal(x, y)) if f64 else (lambda x, y: np.allclose(x, y, rtol=1e-6, atol=1e-6))
        assert same(a[0], b[0])
        if withA:
            assert same(a[1], b[1])
        if not device:
            assert np.allclose(a[2], b[2], rtol=1e-12 if f64 else 1e-5, atol=0 if f64 else 1e-5 * np.max(np.abs(b[2])))
        else:
            assert np.allclose(a[3], b[3], rtol=1e-10 if f64 else 1e-4)
            assert np.array_equal(a[4], psi.astype(cs)) and np.array_equal(a[5], ab.astype(rs))
    with pytest.raises(nat.GLXError):
        dd = D(dold, cs)
        nat.call("glx_cg_direction_step", prec, grid(g), k2, eps, H, None, dd.ptr, None, D(ab.astype(rs), rs).ptr,
                 dd.ptr, D(jac, cs).ptr, dd.ptr, None, None, None, beta.ptr, (ctypes.c_double * 5)(), None)


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("shape", [(6, 5), (64, 9), (130, 33)])
@pytest.mark.parametrize("withA", [True, False])
def test_cg_update_gradient_pass_equals_separate_passes(prec, shape, withA):
    """glx_cg_update_gradient_pass against glx_axpy_dev-style updates
    (glx_axpy with the same alphas) followed by glx_cg_gradient_pass: the
    same updated state, Jacobians, energy and staged betas (FP64 to the
    bit); the inputs are left untouched."""
    rs, cs = _types(prec)
    rng, g, mt, psi, ab, em = case(51 + shape[0], *shape)
    ae = (0.1 * rng.normal(size=g.Nab)).astype(rs)
    dpsi = (0.3 * (rng.normal(size=g.N) + 1j * rng.normal(size=g.N))).astype(cs)
    dab = (0.05 * rng.normal(size=g.Nab)).astype(rs)
    jprev = (0.2 * (rng.normal(size=g.N) + 1j * rng.normal(size=g.N))).astype(cs)
    jAprev = (0.04 * rng.normal(size=g.Nab)).astype(rs)
    k2, H, eps = (9.0 if withA else -1.0), 0.15, 0.0
    al = np.array([0.23, -0.41]) if withA else np.array([0.23, 0.0])
    d_mt, d_ae, d_em = D(mt.astype(np.uint8), np.uint8), D(ae, rs), D(em.astype(rs), rs)
    d_al = D(al, np.float64)

    def run(fused):
        d_psi, d_ab = D(psi.astype(cs), cs), D(ab.astype(rs), rs)
        d_dp, d_da = D(dpsi, cs), D(dab, rs)
        jac, jacA = nat.DeviceArray.zeros(g.N, cs), nat.DeviceArray.zeros(g.Nab, rs)
        jp, jAp = D(jprev, cs), D(jAprev, rs)
        beta = nat.DeviceArray.zeros(2, np.float64)
        F = ctypes.c_double()
        if fused:
            p_out, a_out = nat.DeviceArray.zeros(g.N, cs), nat.DeviceArray.zeros(g.Nab, rs)
            nat.call("glx_cg_update_gradient_pass", prec, grid(g), k2, eps, d_em.ptr, H, d_mt.ptr, d_psi.ptr, d_ae.ptr,
                     d_ab.ptr, d_dp.ptr, d_da.ptr if withA else None, d_al.ptr, p_out.ptr,
                     a_out.ptr if withA else None, jac.ptr, jp.ptr, jacA.ptr if withA else None,
                     jAp.ptr if withA else None, beta.ptr, ctypes.byref(F))
            assert np.array_equal(d_psi.get(), psi.astype(cs)) and np.array_equal(d_ab.get(), ab.astype(rs))
            st_psi, st_ab = p_out.get(), (a_out.get() if withA else d_ab.get())
        else:
            nat.call("glx_axpy", prec, d_dp.ptr, d_psi.ptr, d_psi.ptr, float(rs(al[0])), 2 * g.N)
            if withA:
                nat.call("glx_axpy", prec, d_da.ptr, d_ab.ptr, d_ab.ptr, float(rs(al[1])), g.Nab)
            nat.call("glx_cg_gradient_pass", prec, grid(g), k2, eps, d_em.ptr, H, d_mt.ptr, d_psi.ptr, d_ae.ptr,
                     d_ab.ptr, jac.ptr, jp.ptr, jacA.ptr if withA else None, jAp.ptr if withA else None, beta.ptr,
                     ctypes.byref(F))
            st_psi, st_ab = d_psi.get(), d_ab.get()
        return st_psi, st_ab, jac.get(), (jacA.get() if withA else None), F.value, beta.get()

    a, b = run(True), run(False)
    f64 = prec == nat.GLX_F64
    for x, y in zip(a, b):
        if x is None:
            continue
        if f64:
            assert np.array_equal(x, y)
        else:
            assert np.allclose(x, y, rtol=1e-5, atol=1e-5 * (1 + np.max(np.abs(y))))


@pytest.mark.parametrize("n", [1, 5, 1000, 123457, 2048 * 2048 - 3])
def test_reductions(n):
    """tests/at_reduction.py: sum and 5-vector sums vs numpy, atol 1e-10."""
    rng = np.random.default_rng(n)
    x = rng.random(n)
    out = ctypes.c_double()
    dxv = D(x, np.float64)  # keep the device array alive across the call
    nat.call("glx_sum", nat.GLX_F64, dxv.ptr, n, 0, ctypes.byref(out))
    assert np.isclose(out.value, np.sum(x), atol=1e-10, rtol=1e-13)
    nv = max(1, n // 5)
    v = rng.random(nv * 5)
    o5 = (ctypes.c_double * 5)()
    dv = D(v, np.float64)
    nat.call("glx_sum_v", nat.GLX_F64, dv.ptr, nv, 5, 0, o5)
    assert np.allclose(np.array(o5[:]), v.reshape(nv, 5).sum(axis=0), atol=1e-10, rtol=1e-13)


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("n", [1, 1000, 300001])
def test_pr_beta_device(prec, n):
    """glx_cg_pr_beta leaves max(sum g(g-gp)/sum gp^2, 0) in device memory
    (the oracle's pr_beta, FP64 sums, rounded to the working precision);
    glx_axmy_dev applies it as the reference's axmy does; 0/0 gives 0."""
    rs, _ = _types(prec)
    rng = np.random.default_rng(n)
    g1 = rng.normal(size=n).astype(rs)
    g0 = (g1 + 0.3 * rng.normal(size=n)).astype(rs)
    d1, d0 = D(g1, rs), D(g0, rs)
    beta = nat.DeviceArray.zeros(1, np.float64)
    nat.call("glx_cg_pr_beta", prec, d1.ptr, d0.ptr, n, beta.ptr)
    want = float(rs(O.pr_beta(g1.astype(np.float64), g0.astype(np.float64))))
    got = float(beta.get()[0])
    assert got == pytest.approx(want, rel=1e-12 if prec == nat.GLX_F64 else 1e-6)
    dirv = rng.normal(size=n).astype(rs)
    dd = D(dirv, rs)
    nat.call("glx_axmy_dev", prec, dd.ptr, d1.ptr, dd.ptr, beta.ptr, n)
    assert np.allclose(dd.get(), rs(got) * dirv - g1, rtol=1e-6, atol=1e-6)
    z = nat.DeviceArray.zeros(n, rs)
    nat.call("glx_cg_pr_beta", prec, z.ptr, z.ptr, n, beta.ptr)
    assert beta.get()[0] == 0.0


def test_vortex_cells_match_oracle():
    """Winding numbers: flagged cells and rounded vorticities bit-exact."""
    g = O.Geom(70, 66, 0.5, 0.5)
    x = g.dx * np.arange(g.Nx)[None, :]
    y = g.dy * np.arange(g.Ny)[:, None]
    psi = np.ones((g.Ny, g.Nx), complex)
    for (x0, y0, s) in [(5.3, 6.1, 1), (20.2, 11.7, -1), (30.05, 25.4, 1), (12.6, 29.9, 2)]:
        psi = psi * ((x - x0) + 0.2 * (y - y0) + 1j * s * (y - y0)) ** (1 if abs(s) == 1 else 2)
    psi = psi.ravel() / np.max(np.abs(psi))
    H = 0.05
    ab = O.symmetric_gauge(g, H) * 0.0
    cells, vs = O.vortex_cells(g, psi, ab, 0.0)
    cap = 64
    rec = nat.DeviceArray.zeros(cap * 14, np.float64)
    cnt = ctypes.c_int32()
    d_psi, d_ab = D(psi, np.complex128), D(ab, np.float64)
    nat.call("glx_vortex_cells", nat.GLX_F64, grid(g), d_psi.ptr, d_ab.ptr, None, 0.0,
             rec.ptr, cap, ctypes.byref(cnt))
    r = rec.get()[:cnt.value * 14].reshape(-1, 14)
    assert np.all(np.diff(r[:, 0]) > 0)  # ascending cell id, as the reference loops
    assert list(r[:, 0].astype(int)) == list(cells)
    assert np.array_equal(np.round(r[:, 1]), np.round(vs))
    assert np.max(np.abs(r[:, 1] - vs)) < 1e-12


@pytest.mark.parametrize("prec", PRECS)
def test_vortex_cells_large_random_field(prec):
    """Winding numbers of a random field at 1100 x 1300 nodes: 18 strips of
    63 cells per cell row and 1299 cell rows, so the per-row ranks span many
    strips and the row-offset scan runs over more rows than one pass of its
    workgroup covers. About a fifth of the cells are flagged. Against the
    oracle (FP32 inputs widened exactly, and H rounded to float as svirl's
    Params stores it in a float32 run, params.py:132): the same cells in
    ascending id, vorticities within 1e-12, and every record's corner psi
    and link values those of its own cell, which pins where each record was
    written."""
    rng = np.random.default_rng(21)
    g = O.Geom(1100, 1300, 0.5, 0.5)
    rt, ct = _types(prec)
    psi = (rng.standard_normal(g.N) + 1j * rng.standard_normal(g.N)).astype(ct)
    ab = (0.8 * rng.standard_normal(g.Nab)).astype(rt)
    H = 0.05
    psi64, ab64 = psi.astype(np.complex128), ab.astype(np.float64)
    cells, vs = O.vortex_cells(g, psi64, ab64, float(rt(H)))
    assert 0.1 * g.Nc < len(cells) < 0.4 * g.Nc
    cap = len(cells) + 16
    rec = nat.DeviceArray.zeros(cap * 14, np.float64)
    cnt = ctypes.c_int32()
    d_psi, d_ab = D(psi, ct), D(ab, rt)
    print(f"vortex cells {g.Nx}x{g.Ny} prec {prec}", flush=True)
    nat.call("glx_vortex_cells", prec, grid(g), d_psi.ptr, d_ab.ptr, None, H, rec.ptr, cap, ctypes.byref(cnt))
    assert cnt.value == len(cells)
    r = rec.get()[:cnt.value * 14].reshape(-1, 14)
    n = r[:, 0].astype(np.int64)
    assert np.array_equal(n, cells)
    assert np.max(np.abs(r[:, 1] - vs)) < 1e-12
    i, j = n % g.Nxc, n // g.Nxc
    Nxa, Na = g.Nx - 1, (g.Nx - 1) * g.Ny
    corners = [i + g.Nx * j, i + 1 + g.Nx * j, i + 1 + g.Nx * (j + 1), i + g.Nx * (j + 1)]
    for k, m in enumerate(corners):
        assert np.array_equal(r[:, 2 + 2 * k], psi64.real[m])
        assert np.array_equal(r[:, 3 + 2 * k], psi64.imag[m])
    links = [i + Nxa * j, i + Nxa * (j + 1), Na + i + g.Nx * j, Na + i + 1 + g.Nx * j]
    for k, m in enumerate(links):
        assert np.array_equal(r[:, 10 + k], ab64[m])


@pytest.mark.parametrize("with_abi", [False, True])
def test_vortex_triangulate_matches_oracle(with_abi):
    """Sub-cell triangulation on the device against the oracle's transcription
    of vortex_detector.py:75-147: same kept cells in the same order, equal
    rounded vorticities, positions within 1e-10. Twenty seeded vortices of
    both signs in a symmetric-gauge field, so the corner rotations matter."""
    rng = np.random.default_rng(5)
    g = O.Geom(131, 121, 0.5, 0.45)
    x = g.dx * np.arange(g.Nx)[None, :]
    y = g.dy * np.arange(g.Ny)[:, None]
    psi = np.ones((g.Ny, g.Nx), complex)
    for _ in range(20):
        # |u| ~ distance near the core and ~1 away from it, so the corner
        # values of a vortex cell clear the reference's 1e-10 sign-change test
        x0, y0 = rng.uniform(2, g.dx * (g.Nx - 3)), rng.uniform(2, g.dy * (g.Ny - 3))
        s = rng.choice([-1, 1])
        u = (x - x0) + 0.15 * (y - y0) + 1j * s * (y - y0)
        psi = psi * u / np.sqrt(np.abs(u) ** 2 + 1.0)
    psi = psi.ravel()
    H = 0.02
    ab = O.symmetric_gauge(g, H)
    abi = 0.01 * rng.standard_normal(g.Nab) if with_abi else None
    ox, oy, ov = O.vortex_detector(g, psi, ab, H, abi)
    assert len(ox) >= 10
    cells, vs = O.vortex_cells(g, psi, ab, H)
    cap = 256
    rec = nat.DeviceArray.zeros(cap * 14, np.float64)
    xyv = nat.DeviceArray.zeros(cap * 4, np.float64)
    cnt = ctypes.c_int32()
    d_psi, d_ab = D(psi, np.complex128), D(ab, np.float64)
    d_abi = D(abi, np.float64) if with_abi else None
    nat.call("glx_vortex_cells", nat.GLX_F64, grid(g), d_psi.ptr, d_ab.ptr, d_abi.ptr if with_abi else None, H,
             rec.ptr, cap, ctypes.byref(cnt))
    n = cnt.value
    assert n == len(cells)
    nat.call("glx_vortex_triangulate", grid(g), rec.ptr, n, xyv.ptr)
    r = rec.get()[:n * 14].reshape(n, 14)
    assert list(r[:, 0].astype(int)) == list(cells)  # in the reference's order already
    out = xyv.get()[:n * 4].reshape(n, 4)
    assert set(np.unique(out[:, 3])) <= {0.0, 1.0}
    kept = out[out[:, 3] == 1.0]
    assert len(kept) == len(ox)
    assert np.array_equal(kept[:, 2], ov)
    assert np.max(np.abs(kept[:, 0] - ox)) < 1e-10 and np.max(np.abs(kept[:, 1] - oy)) < 1e-10
    # no records: no launch, no error
    nat.call("glx_vortex_triangulate", grid(g), rec.ptr, 0, xyv.ptr)
    with pytest.raises(nat.GLXError):
        nat.call("glx_vortex_triangulate", grid(g), None, 3, xyv.ptr)


def test_vortex_triangulate_every_cell_of_random_field():
    """The triangulation run on every cell of a random field (records built
    on the host, in shuffled order), so the rejections are exercised too:
    border zero counts other than two, near-parallel lines and points
    outside the 3x3 neighbourhood. Kept cells, vorticities and positions
    against the oracle's triangulate (vortex_detector.py:75-147)."""
    rng = np.random.default_rng(9)
    g = O.Geom(41, 37, 0.5, 0.5)
    psi = rng.standard_normal(g.N) + 1j * rng.standard_normal(g.N)
    ab = 0.8 * rng.standard_normal(g.Nab)
    vs = rng.uniform(-2.4, 2.4, g.Nc)
    cells = np.arange(g.Nc)
    ox = O.triangulate(g, cells, vs, psi, g.a2(ab).T, g.b2(ab).T)
    i, j = cells % g.Nxc, cells // g.Nxc
    q = i + g.Nx * j
    rec = np.empty((g.Nc, 14))
    rec[:, 0], rec[:, 1] = cells, vs
    for k, off in enumerate((0, 1, 1 + g.Nx, g.Nx)):
        rec[:, 2 + 2 * k], rec[:, 3 + 2 * k] = psi[q + off].real, psi[q + off].imag
    rec[:, 10], rec[:, 11] = ab[i + g.Nxa * j], ab[i + g.Nxa * (j + 1)]
    rec[:, 12], rec[:, 13] = ab[g.Na + i + g.Nxb * j], ab[g.Na + i + 1 + g.Nxb * j]
    perm = rng.permutation(g.Nc)
    d_rec = D(rec[perm].ravel(), np.float64)
    xyv = nat.DeviceArray.zeros(4 * g.Nc, np.float64)
    nat.call("glx_vortex_triangulate", grid(g), d_rec.ptr, g.Nc, xyv.ptr)
    out = np.empty((g.Nc, 4))
    out[perm] = xyv.get().reshape(g.Nc, 4)
    kept = out[out[:, 3] == 1.0]
    assert 0 < len(kept) < g.Nc
    assert len(kept) == len(ox)
    assert np.array_equal(kept[:, 2], ox[:, 2])
    assert np.max(np.abs(kept[:, :2] - ox[:, :2])) < 1e-10


def test_capacity_error():
    g = O.Geom(16, 16, 0.5, 0.5)
    x = g.dx * np.arange(g.Nx)[None, :]
    y = g.dy * np.arange(g.Ny)[:, None]
    psi = ((x - 2.3) + 0.1 * (y - 2.1) + 1j * (y - 2.1)) * ((x - 5.2) + 0.1 * (y - 5.6) + 1j * (y - 5.6))
    rec = nat.DeviceArray.zeros(14, np.float64)
    cnt = ctypes.c_int32()
    d_psi, d_ab = D(psi.ravel(), np.complex128), D(np.zeros(g.Nab), np.float64)
    with pytest.raises(nat.GLXError) as e:
        nat.call("glx_vortex_cells", nat.GLX_F64, grid(g), d_psi.ptr, d_ab.ptr, None, 0.0, rec.ptr, 1,
                 ctypes.byref(cnt))
    assert e.value.code == nat.GLX_E_CAPACITY and cnt.value == 2


def test_argument_validation():
    """Host-side validation of the C ABI: each call is rejected with
    GLX_E_ARG and a message before anything is launched (no kernel runs)."""
    def rejects(name, *args):
        with pytest.raises(nat.GLXError) as e:
            nat.call(name, *args)
        assert e.value.code == nat.GLX_E_ARG, (name, e.value)
        assert str(e.value)

    F = ctypes.c_double()
    x = nat.DeviceArray.zeros(64, np.float64)
    big = nat.Grid(40000, 40000, 0.5, 0.5)   # more than 2^30 nodes
    for g in (big, nat.Grid(1, 8, 0.5, 0.5), nat.Grid(8, 8, 0.0, 0.5), nat.Grid(8, 8, 0.5, -1.0)):
        rejects("glx_free_energy", nat.GLX_F64, g, -1.0, 1.0, None, 0.0, None, x.ptr, None, x.ptr, ctypes.byref(F))
    g = nat.Grid(8, 8, 0.5, 0.5)
    rejects("glx_free_energy", nat.GLX_F64, g, -1.0, 1.0, None, 0.0, None, None, None, x.ptr, ctypes.byref(F))
    rejects("glx_free_energy", 7, g, -1.0, 1.0, None, 0.0, None, x.ptr, None, x.ptr, ctypes.byref(F))  # precision
    out = ctypes.c_double()
    rejects("glx_sum", nat.GLX_F64, x.ptr, 64, 100, ctypes.byref(out))   # block size not a multiple of 64
    rejects("glx_sum", nat.GLX_F64, x.ptr, -1, 0, ctypes.byref(out))
    # persistent TD: bad buffer index, negative step count, grid above the co-resident limit
    p = _td_params(0.1, 1.0, -1.0, 1.0, 0.0, 1e-6, 1e-6, False)
    P = [nat.DeviceArray.zeros(g.Nx * g.Ny, np.complex128) for _ in range(3)]
    Q = [nat.DeviceArray.zeros(2 * (g.Nx - 1) * g.Ny, np.float64) for _ in range(3)]
    Pp = (ctypes.c_void_p * 3)(*[b.ptr for b in P])
    Qp = (ctypes.c_void_p * 3)(*[b.ptr for b in Q])
    sw = (ctypes.c_int32 * 8)()
    for pc, qc, nsteps, wg in ((3, 0, 1, 0), (0, -1, 1, 0), (0, 0, -1, 0), (0, 0, 1, 1 << 20)):
        rejects("glx_td_run", nat.GLX_F64, g, ctypes.byref(p), None, None, Pp, ctypes.byref(ctypes.c_int32(pc)), Qp,
                ctypes.byref(ctypes.c_int32(qc)), None, None, nsteps, sw, wg, 256)
    p.a_src = Q[1].ptr  # a fixed source may not be one of the rotating buffers
    p.solveA = 1
    p.kappa2 = 4.0
    rejects("glx_td_run", nat.GLX_F64, g, ctypes.byref(p), None, None, Pp, ctypes.byref(ctypes.c_int32(0)), Qp,
            ctypes.byref(ctypes.c_int32(0)), None, None, 1, sw, 0, 256)
    # fused CG passes: outputs may not alias their inputs
    d, j = nat.DeviceArray.zeros(g.Nx * g.Ny, np.complex128), nat.DeviceArray.zeros(g.Nx * g.Ny, np.complex128)
    beta = nat.DeviceArray.zeros(4, np.float64)
    rejects("glx_cg_update_gradient_pass", nat.GLX_F64, g, -1.0, 1.0, None, 0.0, None, P[0].ptr, None, Q[0].ptr,
            d.ptr, None, beta.ptr, P[0].ptr, None, j.ptr, j.ptr, None, None, beta.ptr, ctypes.byref(F))
    rejects("glx_vortex_triangulate", big, x.ptr, 1, x.ptr)
    rejects("glx_bulk_observables", nat.GLX_F64, g, 4.0, 0.1, None, None, None, Q[0].ptr, None, None, x.ptr)
