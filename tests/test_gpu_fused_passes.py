"""The fused multi-sweep passes of the persistent solves (psi: passes of up to
three sweeps; A: an opening three-sweep pass, pairs, single sweeps) against
the FP64 oracle's sweep-by-sweep iteration over runs whose sweep counts move
up and down: every step's sweep counts exact, fields within 1e-11. The
schedules cover a three-sweep A pass that does not converge (the solve
continues from the A^2 it stored), one whose second sweep already meets the
stop rule (redone as a pair), and psi passes shortened by the previous
step's count."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
nat = pytest.importorskip("svirl_amd._native")

from oracle import gl_oracle as O  # noqa: E402

# (kappa, H, dt, sigma, steps): A sweep counts of the oracle on this grid
# include 5, 4, 3, 3, 3, 4 (first case) and 3, ..., 3, 2 (second case)
CASES = [(5.0, 0.1, 0.02, 50.0, 12), (5.0, 0.1, 0.05, 200.0, 12), (1.0, 0.2, 0.5, 20.0, 8)]


@pytest.fixture(scope="module", autouse=True)
def _init():
    nat.init(0)


def D(a, dtype):
    return nat.DeviceArray.from_host(np.asarray(a), dtype)


def gpu_run(g, psi0, ab0, kappa, sigma, H, dt, nsteps, wg):
    N = g.N
    P = [D(psi0, np.complex128)] + [nat.DeviceArray.zeros(N, np.complex128) for _ in range(2)]
    Q = [D(ab0, np.float64)] + [nat.DeviceArray.zeros(ab0.size, np.float64) for _ in range(2)]
    Pp = (ctypes.c_void_p * 3)(*[b.ptr for b in P])
    Qp = (ctypes.c_void_p * 3)(*[b.ptr for b in Q])
    pc, qc = ctypes.c_int32(0), ctypes.c_int32(0)
    prm = nat.TDParams()
    prm.dt, prm.eps, prm.kappa2, prm.rho, prm.H = dt, 1.0, kappa * kappa, 1.0 / sigma, H
    prm.stop_psi, prm.stop_A = 1e-6, 1e-6
    prm.langevin_psi = prm.langevin_A = 0.0
    prm.rand_t, prm.solveA, prm.max_sweeps, prm.eqn = 1, 1, 1024, 0
    sw = (ctypes.c_int32 * (2 * nsteps))()
    nat.call("glx_td_run", nat.GLX_F64, nat.Grid(g.Nx, g.Ny, g.dx, g.dy), ctypes.byref(prm), None, None, Pp,
             ctypes.byref(pc), Qp, ctypes.byref(qc), None, None, nsteps, sw, wg, 256)
    return P[pc.value].get(), Q[qc.value].get(), [tuple(x) for x in np.frombuffer(sw, dtype=np.int32).reshape(nsteps, 2)]


@pytest.mark.parametrize("case", range(len(CASES)))
@pytest.mark.parametrize("wg", [1, 8])
def test_fused_passes_match_sweep_by_sweep(case, wg):
    kappa, H, dt, sigma, nsteps = CASES[case]
    g = O.Geom(150, 130, 0.5, 0.5)
    psi0 = O.random_psi(g.N, 1.0, seed=5)
    ab0 = O.symmetric_gauge(g, H) if case < 2 else O.symmetric_gauge(g, 0.0)
    st = O.OracleState(g, psi0.copy(), ab0.copy(), kappa=kappa, sigma=sigma, H=H)
    st.td(dt=dt, Nt=nsteps)
    psi, ab, sw = gpu_run(g, psi0, ab0, kappa, sigma, H, dt, nsteps, wg)
    assert sw == st.sweeps
    assert np.max(np.abs(psi - st.psi)) <= 1e-11
    assert np.max(np.abs(ab - st.ab)) <= 1e-11


def test_fp32_fused_psi_solve_equals_single_sweeps():
    """FP32: the persistent psi solve (fused passes of up to three sweeps,
    shared link factors and diagonals) gives the sweep counts of single
    sweeps (glx_td_psi_sweep) driven by the reference's host loop and stop
    rule (td.py:157-218), and their fields to a few FP32 ulps (FP32 code is
    built with FMA contraction, which the compiler applies per code path)."""
    rng = np.random.default_rng(11)
    g = O.Geom(150, 130, 0.5, 0.5)
    H, dt, nsteps = 0.1, 0.05, 6
    psi0 = O.random_psi(g.N, 1.0, seed=9).astype(np.complex64)
    ab = O.symmetric_gauge(g, H).astype(np.float32)
    mt = (rng.random(g.Nxc * g.Nyc) > 0.1).astype(np.uint8)
    em = (1.0 + 0.2 * rng.normal(size=g.N)).astype(np.float32)
    grid = nat.Grid(g.Nx, g.Ny, g.dx, g.dy)
    d_ab, d_mt, d_em = D(ab, np.float32), D(mt, np.uint8), D(em, np.float32)
    # persistent solve, psi equation only
    P = [D(psi0, np.complex64)] + [nat.DeviceArray.zeros(g.N, np.complex64) for _ in range(2)]
    Q = [d_ab] + [nat.DeviceArray.zeros(g.Nab, np.float32) for _ in range(2)]
    Pp = (ctypes.c_void_p * 3)(*[b.ptr for b in P])
    Qp = (ctypes.c_void_p * 3)(*[b.ptr for b in Q])
    pc, qc = ctypes.c_int32(0), ctypes.c_int32(0)
    prm = nat.TDParams()
    prm.dt, prm.eps, prm.kappa2, prm.rho, prm.H = dt, 1.0, 25.0, 1.0 / 50.0, H
    prm.stop_psi, prm.stop_A = 1e-6, 1e-6
    prm.langevin_psi = prm.langevin_A = 0.0
    prm.rand_t, prm.solveA, prm.max_sweeps, prm.eqn = 1, 1, 1024, 1
    sw = (ctypes.c_int32 * (2 * nsteps))()
    nat.call("glx_td_run", nat.GLX_F32, grid, ctypes.byref(prm), d_em.ptr, d_mt.ptr, Pp, ctypes.byref(pc), Qp,
             ctypes.byref(qc), None, None, nsteps, sw, 8, 256)
    fused = P[pc.value].get()
    fused_sw = [int(x) for x in np.frombuffer(sw, dtype=np.int32).reshape(nsteps, 2)[:, 0]]
    # single sweeps and the host stop rule
    rhs = D(psi0, np.complex64)
    bufs = [nat.DeviceArray.zeros(g.N, np.complex64), nat.DeviceArray.zeros(g.N, np.complex64)]
    r2 = nat.DeviceArray.zeros(1, np.int32)
    counts = []
    for step in range(nsteps):
        cur = rhs
        j = 0
        while True:
            r2.fill(0)
            nxt = bufs[j % 2]
            nat.call("glx_td_psi_sweep", nat.GLX_F32, grid, dt, 1.0, d_em.ptr, d_mt.ptr, d_ab.ptr, rhs.ptr, cur.ptr,
                     nxt.ptr, 0.0, j, 1 + step, 1e-6, r2.ptr)
            cur = nxt
            j += 1
            if int(r2.get()[0]) < 10000 or j >= 1024:
                break
        counts.append(j)
        rhs.set(cur.get())
    assert fused_sw == counts
    assert np.max(np.abs(fused - rhs.get())) <= 1e-6
