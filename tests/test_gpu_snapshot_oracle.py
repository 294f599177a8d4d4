"""Snapshot resume against the oracle: a solver loaded from a checkpoint
continues the way the reference's own objects would have continued.

The reference carries hidden state between calls: TD._random_t, the
counter of the Langevin noise, advanced once per psi solve and once per A
solve (td.py:35-37, 206, 308), and the CG search directions and betas,
which persist across cg() calls (cg.py:54-71; with kappa < inf the first
iteration of a call reuses the previous call's direction and beta,
cg.py:481-483). The oracle models the reference's objects as one
OracleState that simply keeps running; the GPU side stops, writes a
snapshot, and a fresh solver loads it and runs on. FP64 throughout: fields
within 1e-10 (TD) / 1e-9 (CG) relative, sweep counts exact, energies within
1e-10.

The oracle's side is pinned too: a new OracleState built from nothing but
the items a snapshot holds (psi, A, the noise counter, the CG directions
and betas), taken from the oracle's own state, continues to the same bits
as the uninterrupted one, so those items are the whole of the state the
restated loops carry.
"""
import numpy as np
import pytest

from oracle import gl_oracle as O

pytestmark = pytest.mark.gpu

svirl_amd = pytest.importorskip("svirl_amd")
GLSolver = svirl_amd.GLSolver

from tests.test_gpu_solver import gl_ab, gl_psi, oracle_of, rel  # noqa: E402

LP, LA = 0.05, 0.02  # Langevin coefficients of psi and A


def make(langevin=True):
    gl = GLSolver(Nx=33, Ny=29, dx=0.5, dy=0.5, order_parameter='random', random_seed=4, gl_parameter=5.0,
                  normal_conductivity=100.0, homogeneous_external_field=0.15, dtype=np.float64, convergence_rtol=0.0)
    if langevin:
        gl.params.order_parameter_Langevin_coefficient = LP
        gl.params.vector_potential_Langevin_coefficient = LA
    return gl


def resumed(ref, d):
    """An OracleState with ref's parameters and, of its state, only what a
    snapshot holds: psi, A, the noise counter and the CG directions and
    betas (keys as the snapshot file names them)."""
    st = O.OracleState(ref.g, d["psi"], d["ab"], kappa=ref.kappa, sigma=ref.sigma, H=ref.H, eps=ref.eps,
                       eps_map=ref.eps_map, mt=ref.mt, abei=ref.abei, stop_psi=ref.stop_psi, stop_A=ref.stop_A,
                       langevin_psi=ref.langevin_psi, langevin_A=ref.langevin_A,
                       random_t=int(d["td_random_t"]) if "td_random_t" in d else 1,
                       convergence_rtol=ref.convergence_rtol)
    if "cg_dir_psi" in d:
        st._cg_init()
        st.dir_psi = np.array(d["cg_dir_psi"], dtype=np.complex128)
        st.beta_psi, st.beta_A = (float(x) for x in np.asarray(d["cg_beta"])[:2])
        if "cg_dir_A" in d:
            st.dir_A = np.array(d["cg_dir_A"], dtype=np.float64)
    return st


def oracle_from_snapshot(ref, path):
    with np.load(path, allow_pickle=False) as f:
        return resumed(ref, {k: f[k] for k in f.files})


def own_snapshot(st):
    """The same items taken from the oracle's own state."""
    d = {"psi": st.psi.copy(), "ab": st.ab.copy(), "td_random_t": st.random_t}
    if st._cg_ready:
        d.update({"cg_dir_psi": st.dir_psi.copy(), "cg_beta": np.array([st.beta_psi, st.beta_A]),
                  "cg_dir_A": st.dir_A.copy()})
    return d


def test_td_with_noise_resumes_like_the_oracle(tmp_path):
    from svirl_amd import snapshot
    a = make()
    st = oracle_of(a)
    st.langevin_psi, st.langevin_A = LP, LA
    a.solve.td(dt=0.1, Nt=3)
    st.td(dt=0.1, Nt=3)
    assert a.solve._td.sweeps == st.sweeps
    assert rel(gl_psi(a), st.psi) < 1e-10 and rel(gl_ab(a), st.ab) < 1e-10
    p = str(tmp_path / "td.npz")
    snapshot.save(a, p)
    with np.load(p, allow_pickle=False) as f:
        # the noise counter moved once per psi solve and once per A solve
        assert int(f["td_random_t"]) == st.random_t == 4 + 2 * 3
        assert np.array_equal(f["psi"], gl_psi(a)) and np.array_equal(f["ab"], gl_ab(a))
    b = make()
    snapshot.load(b, p)
    b.solve.td(dt=0.1, Nt=4)
    st_own = resumed(st, own_snapshot(st))
    st2 = oracle_from_snapshot(st, p)
    st.td(dt=0.1, Nt=4)
    st_own.td(dt=0.1, Nt=4)
    st2.td(dt=0.1, Nt=4)
    # the oracle resumed from its own snapshot items is the uninterrupted
    # oracle to the bit; resumed from the GPU's snapshot, to FP64 rounding
    assert np.array_equal(st_own.psi, st.psi) and np.array_equal(st_own.ab, st.ab)
    assert st_own.random_t == st.random_t and st_own.sweeps == st.sweeps[-4:]
    assert rel(st2.psi, st.psi) < 1e-10 and st2.sweeps == st.sweeps[-4:]
    assert b.solve._td.sweeps[-4:] == st.sweeps[-4:]
    assert rel(gl_psi(b), st.psi) < 1e-10 and rel(gl_ab(b), st.ab) < 1e-10
    # a solver that lost the counter draws other noise
    c = make()
    c.vars.order_parameter = a.vars.order_parameter
    c.vars.vector_potential = a.vars.vector_potential
    c.solve.td(dt=0.1, Nt=4)
    assert rel(gl_psi(c), st.psi) > 1e-6


def test_cg_resumes_like_the_oracle(tmp_path):
    """kappa < inf: the second cg() call continues with the first call's
    direction and its last beta (not a steepest-descent restart)."""
    from svirl_amd import snapshot
    a = make(langevin=False)
    a.solve.td(dt=0.1, Nt=3)
    st = oracle_of(a)
    a.solve.cg(n_iter=4)
    st.cg(n_iter=4)
    assert np.allclose(a.solve._cg.cg_energies, st.cg_energies, rtol=1e-10)
    p = str(tmp_path / "cg.npz")
    snapshot.save(a, p)
    b = make(langevin=False)
    snapshot.load(b, p)
    b.solve.cg(n_iter=5)
    st_own = resumed(st, own_snapshot(st))
    st2 = oracle_from_snapshot(st, p)
    st.cg(n_iter=5)
    st_own.cg(n_iter=5)
    st2.cg(n_iter=5)
    assert st_own.cg_energies == st.cg_energies and np.array_equal(st_own.psi, st.psi)
    assert np.allclose(st2.cg_energies, st.cg_energies, rtol=1e-10)
    assert np.allclose(b.solve._cg.cg_energies, st.cg_energies, rtol=1e-10)
    assert rel(gl_psi(b), st.psi) < 1e-9 and rel(gl_ab(b), st.ab) < 1e-9
    # a restart from steepest descent (directions dropped) lands elsewhere
    with np.load(p, allow_pickle=False) as f:
        assert np.any(f["cg_beta"][:2] > 0.0)
    st3 = oracle_from_snapshot(st, p)
    st3.dir_psi[:] = 0.0
    st3.dir_A[:] = 0.0
    st3.beta_psi = st3.beta_A = 0.0
    st3.cg(n_iter=5)
    assert abs(st3.cg_energies[0] - st.cg_energies[0]) > 1e-12 * abs(st.cg_energies[0])
