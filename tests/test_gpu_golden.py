"""The GPU build against the committed golden fixtures (tests/golden/*.npz):
sweep counts exact, fields within 1e-11 (FP64)."""
import ctypes
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
nat = pytest.importorskip("svirl_amd._native")

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    return np.load(os.path.join(HERE, name + ".npz"), allow_pickle=False)


def D(a, dtype):
    return nat.DeviceArray.from_host(np.asarray(a), dtype)


@pytest.fixture(scope="module", autouse=True)
def _init():
    nat.init(0)


def run_td(f, nsteps, kappa2, rho, H, eps, em, mt, max_sweeps=1024, stop_A=1e-6, eqn=0, ab0=None):
    g = nat.Grid(int(f["Nx"]), int(f["Ny"]), float(f["dx"]), float(f["dy"]))
    N = int(f["Nx"]) * int(f["Ny"])
    psi0 = f["psi0"] if "psi0" in f.files else f["psi"]
    ab0 = f["ab0"] if ab0 is None else ab0
    P = [D(psi0, np.complex128)] + [nat.DeviceArray.zeros(N, np.complex128) for _ in range(2)]
    Q = [D(ab0, np.float64)] + [nat.DeviceArray.zeros(ab0.size, np.float64) for _ in range(2)]
    Pp = (ctypes.c_void_p * 3)(*[b.ptr for b in P])
    Qp = (ctypes.c_void_p * 3)(*[b.ptr for b in Q])
    pc, qc = ctypes.c_int32(0), ctypes.c_int32(0)
    prm = nat.TDParams()
    prm.dt, prm.eps, prm.kappa2, prm.rho, prm.H = float(f["dt"]), eps, kappa2, rho, H
    prm.stop_psi, prm.stop_A = 1e-6, stop_A
    prm.langevin_psi = prm.langevin_A = 0.0
    prm.rand_t, prm.solveA, prm.max_sweeps, prm.eqn = 1, 1 if kappa2 > 0 else 0, max_sweeps, eqn
    sw = (ctypes.c_int32 * (2 * nsteps))()
    d_em = D(em, np.float64) if em is not None else None
    d_mt = D(np.asarray(mt).astype(np.uint8), np.uint8) if mt is not None else None
    nat.call("glx_td_run", nat.GLX_F64, g, ctypes.byref(prm), nat.ptr(d_em), nat.ptr(d_mt), Pp, ctypes.byref(pc),
             Qp, ctypes.byref(qc), None, None, nsteps, sw, 0, 256)
    return P[pc.value].get(), Q[qc.value].get(), np.frombuffer(sw, dtype=np.int32).reshape(nsteps, 2)


def test_td_kinf_17():
    f = load("td_kinf_17")
    psi, _, sw = run_td(f, int(f["Nt"]), -1.0, 1.0, float(f["H"]), 1.0, None, None)
    assert np.array_equal(sw, f["sweeps"])
    assert np.max(np.abs(psi - f["psi"])) < 1e-11


def test_td_k5_mask():
    f = load("td_k5_mask")
    psi, ab, sw = run_td(f, int(f["Nt"]), float(f["kappa"]) ** 2, 1.0 / float(f["sigma"]), float(f["H"]), 0.0,
                         f["eps_map"], f["mt"])
    assert np.array_equal(sw, f["sweeps"])
    assert np.max(np.abs(psi - f["psi"])) < 1e-11
    assert np.max(np.abs(ab - f["ab"])) < 1e-11


def test_a_alias_schedule_persistent():
    """Six A sweeps of one solve in the persistent kernel (A only, capped at
    6 sweeps) land on the fixture's sixth iterate."""
    f = load("a_alias_9")
    _, ab, sw = run_td(f, 1, float(f["kappa2"]), float(f["rho"]), float(f["H"]), 1.0, None, None,
                       max_sweeps=6, stop_A=1e-14, eqn=2)
    assert sw[0, 1] == 6
    assert np.max(np.abs(ab - f["iterates"][5])) < 1e-12


def test_a_alias_schedule_single_sweeps():
    """The same six sweeps through glx_td_a_sweep with the reference's
    ping-pong buffers (the source buffer is the original A buffer)."""
    f = load("a_alias_9")
    g = nat.Grid(9, 9, 0.5, 0.5)
    X = D(f["ab0"], np.float64)
    Y = nat.DeviceArray.zeros(X.size, np.float64)
    rhs = X.copy()
    psi = D(f["psi"], np.complex128)
    r2 = nat.DeviceArray.zeros(1, np.int32)
    cur, nxt = X, Y
    for j in range(6):
        r2.fill(0)
        nat.call("glx_td_a_sweep", nat.GLX_F64, g, float(f["dt"]), float(f["kappa2"]), float(f["rho"]),
                 float(f["H"]), None, psi.ptr, X.ptr, rhs.ptr, cur.ptr, nxt.ptr, 0.0, j, 1, 1e-14, r2.ptr)
        assert np.max(np.abs(nxt.get() - f["iterates"][j])) < 1e-12
        assert abs(int(r2.get()[0]) - int(f["residuals"][j])) <= 1
        cur, nxt = nxt, cur


def test_observables_fixture():
    f = load("observables")
    g = nat.Grid(int(f["Nx"]), int(f["Ny"]), float(f["dx"]), float(f["dy"]))
    psi, ab, abei, em = D(f["psi"], np.complex128), D(f["ab"], np.float64), D(f["abei"], np.float64), D(f["eps_map"], np.float64)
    mt = D(f["mt"].astype(np.uint8), np.uint8)
    F = ctypes.c_double()
    nat.call("glx_free_energy", nat.GLX_F64, g, float(f["kappa2"]), 0.0, em.ptr, float(f["H"]), mt.ptr, psi.ptr,
             abei.ptr, ab.ptr, ctypes.byref(F))
    assert F.value == pytest.approx(float(f["F"]), rel=1e-12)
    out = nat.DeviceArray.zeros(f["J"].size, np.float64)
    nat.call("glx_current_density", nat.GLX_F64, g, float(f["kappa2"]), float(f["H"]), abei.ptr, ab.ptr, out.ptr)
    assert np.allclose(out.get(), f["J"], rtol=1e-12, atol=1e-11)
    nat.call("glx_supercurrent_density", nat.GLX_F64, g, mt.ptr, psi.ptr, abei.ptr, ab.ptr, out.ptr)
    assert np.allclose(out.get(), f["Js"], rtol=1e-12, atol=1e-13)
    B = nat.DeviceArray.zeros(f["B"].size, np.float64)
    nat.call("glx_magnetic_field", nat.GLX_F64, g, abei.ptr, ab.ptr, B.ptr)
    assert np.allclose(B.get(), f["B"], rtol=1e-12, atol=1e-12)


def test_cg_kinf_13_through_api():
    from svirl_amd import GLSolver
    f = load("cg_kinf_13")
    gl = GLSolver(Nx=13, Ny=13, dx=0.5, dy=0.5, order_parameter=1.0, gl_parameter=np.inf,
                  homogeneous_external_field=float(f["H"]))
    gl.vars.order_parameter = gl.unflatten_array(f["psi0"])
    gl.solve.cg(n_iter=8)
    assert np.allclose(gl.solve._cg.cg_energies, f["energies"], rtol=1e-11)
    assert np.max(np.abs(gl.flatten_array(gl.vars.order_parameter) - f["psi"])) < 1e-10


def test_solve_and_step_entry_points_match_td_run():
    """glx_td_psi_solve then glx_td_a_solve, and glx_td_step, give bit for
    bit the state and sweep counts of one glx_td_run step (FP64)."""
    f = load("td_k5_mask")
    psi_ref, ab_ref, sw_ref = run_td(f, 1, float(f["kappa"]) ** 2, 1.0 / float(f["sigma"]), float(f["H"]), 0.0,
                                     f["eps_map"], f["mt"])
    g = nat.Grid(int(f["Nx"]), int(f["Ny"]), float(f["dx"]), float(f["dy"]))
    N = int(f["Nx"]) * int(f["Ny"])
    prm = nat.TDParams()
    prm.dt, prm.eps, prm.kappa2, prm.rho, prm.H = float(f["dt"]), 0.0, float(f["kappa"]) ** 2, 1.0 / float(f["sigma"]), float(f["H"])
    prm.stop_psi, prm.stop_A = 1e-6, 1e-6
    prm.langevin_psi = prm.langevin_A = 0.0
    prm.rand_t, prm.solveA, prm.max_sweeps, prm.eqn = 1, 1, 1024, 0
    em = D(f["eps_map"], np.float64)
    mt = D(f["mt"].astype(np.uint8), np.uint8)
    # psi solve, then A solve
    P = [D(f["psi0"], np.complex128)] + [nat.DeviceArray.zeros(N, np.complex128) for _ in range(2)]
    Q = [D(f["ab0"], np.float64)] + [nat.DeviceArray.zeros(f["ab0"].size, np.float64) for _ in range(2)]
    Pp = (ctypes.c_void_p * 3)(*[b.ptr for b in P])
    Qp = (ctypes.c_void_p * 3)(*[b.ptr for b in Q])
    pc, qc, sp, sa = ctypes.c_int32(0), ctypes.c_int32(0), ctypes.c_int32(), ctypes.c_int32()
    nat.call("glx_td_psi_solve", nat.GLX_F64, g, ctypes.byref(prm), em.ptr, mt.ptr, Q[0].ptr, Pp, ctypes.byref(pc),
             None, ctypes.byref(sp))
    nat.call("glx_td_a_solve", nat.GLX_F64, g, ctypes.byref(prm), mt.ptr, P[pc.value].ptr, Qp, ctypes.byref(qc),
             None, ctypes.byref(sa))
    assert (sp.value, sa.value) == tuple(sw_ref[0])
    assert np.array_equal(P[pc.value].get(), psi_ref)
    assert np.array_equal(Q[qc.value].get(), ab_ref)
    # one step
    P2 = [D(f["psi0"], np.complex128)] + [nat.DeviceArray.zeros(N, np.complex128) for _ in range(2)]
    Q2 = [D(f["ab0"], np.float64)] + [nat.DeviceArray.zeros(f["ab0"].size, np.float64) for _ in range(2)]
    Pp2 = (ctypes.c_void_p * 3)(*[b.ptr for b in P2])
    Qp2 = (ctypes.c_void_p * 3)(*[b.ptr for b in Q2])
    pc2, qc2 = ctypes.c_int32(0), ctypes.c_int32(0)
    sw = (ctypes.c_int32 * 2)()
    nat.call("glx_td_step", nat.GLX_F64, g, ctypes.byref(prm), em.ptr, mt.ptr, Pp2, ctypes.byref(pc2), Qp2,
             ctypes.byref(qc2), None, None, sw)
    assert tuple(sw) == tuple(sw_ref[0])
    assert np.array_equal(P2[pc2.value].get(), psi_ref)
    assert np.array_equal(Q2[qc2.value].get(), ab_ref)
