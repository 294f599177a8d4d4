"""Writes the golden fixtures in this directory from the FP64 oracle.

The reference holds no data files or known-answer vectors (every one of its
tests draws random inputs at run time), so these fixtures are produced by
the oracle (oracle/gl_oracle.py), after it passed the checks in
tests/test_oracle.py. They freeze small, seeded cases so that the GPU build
and the oracle itself are checked against the same stored bits:

  td_kinf_17.npz     17x17 nodes, kappa = inf, H = 0.1, 6 TD steps
  td_k5_mask.npz     21x19 nodes, kappa = 5, holes + eps map, 5 TD steps
  a_alias_9.npz      9x9 nodes, the first 6 A sweeps of one solve (the
                     A^(2*floor(j/2)) source schedule of td.py:264-303)
  observables.npz    F, J, J_s, B of a random 13x11 state with abei
  cg_kinf_13.npz     13x13 nodes, kappa = inf, 8 CG iterations

and the benchmark configurations at their own sizes (seeded inputs built as
svirl_amd.workloads builds them for the GPU):

  c1_129.npz         Config 1: 129^2, kappa = inf, H = 0.1, TD dt = 0.1 for
                     Nt = 1000 steps, then CG n_iter = 1000 (the reference's
                     energy stop rule ends it earlier): the whole state
  c2_1025.npz        Config 2: 1025^2, kappa = 5, sigma = 200, H = 0.1,
                     3 TD steps then 5 CG iterations: sweep counts, energies
                     and sampled rows of psi, a, b and J
  c3_4097.npz        Config 3: 4097^2 with holes and pinning map, kappa = 5,
                     sigma = 200, H = 0.05, 2 TD steps of dt = 0.05: the same
                     (about four minutes and 10 GB of memory)
  c4_sample.npz      the four smallest Config 4 instances after 3 TD steps
  c5_sample.npz      two Config 5 instances after one exact steepest-descent
                     line-minimum step

Run: python tests/golden/make_golden.py [name ...]   (no name: the small
cases and c1_129, c2_1025)
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import gl_oracle as O  # noqa: E402


def td_kinf_17():
    g = O.Geom(17, 17, 0.5, 0.5)
    psi0 = O.random_psi(g.N, 1.0, seed=1)
    ab0 = O.symmetric_gauge(g, 0.1)
    st = O.OracleState(g, psi0, ab0, kappa=np.inf, H=0.1)
    st.td(dt=0.1, Nt=6)
    return dict(Nx=17, Ny=17, dx=0.5, dy=0.5, H=0.1, dt=0.1, Nt=6, psi0=psi0, ab0=ab0,
                psi=st.psi, sweeps=np.array(st.sweeps), F=st.free_energy())


def td_k5_mask():
    rng = np.random.default_rng(7)
    g = O.Geom(21, 19, 0.5, 0.5)
    mt = np.ones((g.Nyc, g.Nxc), bool)
    mt[6:11, 8:13] = False
    mt = mt.ravel()
    em = 1.0 + 0.2 * rng.normal(size=g.N)
    psi0 = np.where(O.mask_at_nodes(g, mt), O.random_psi(g.N, 1.0, seed=2), 0)
    ab0 = O.symmetric_gauge(g, 0.2)
    st = O.OracleState(g, psi0, ab0, kappa=5.0, sigma=50.0, H=0.2, eps_map=em, mt=mt)
    st.td(dt=0.1, Nt=5)
    return dict(Nx=21, Ny=19, dx=0.5, dy=0.5, kappa=5.0, sigma=50.0, H=0.2, dt=0.1, Nt=5, mt=mt, eps_map=em,
                psi0=psi0, ab0=ab0, psi=st.psi, ab=st.ab, sweeps=np.array(st.sweeps), F=st.free_energy())


def a_alias_9():
    rng = np.random.default_rng(9)
    g = O.Geom(9, 9, 0.5, 0.5)
    psi = (0.5 + 0.5 * rng.random(g.N)) * np.exp(2j * np.pi * rng.random(g.N))
    ab0 = O.symmetric_gauge(g, 0.3) + 0.05 * rng.normal(size=g.Nab)
    kappa2, rho, H, dt = 9.0, 1 / 5.0, 0.3, 0.1
    A = [ab0.copy()]
    rhs = ab0.copy()
    rs = []
    for j in range(6):
        src = A[2 * (j // 2)]
        nxt, rhs, r = O.a_jacobi_sweep(g, dt, kappa2, rho, H, None, psi, src, rhs, A[-1], stop_eps=1e-14)
        A.append(nxt)
        rs.append(r)
    return dict(Nx=9, Ny=9, dx=0.5, dy=0.5, kappa2=kappa2, rho=rho, H=H, dt=dt, psi=psi, ab0=ab0,
                iterates=np.array(A[1:]), residuals=np.array(rs))


def observables():
    rng = np.random.default_rng(13)
    g = O.Geom(13, 11, 0.45, 0.5)
    mt = rng.random(g.Nc) > 0.2
    psi = (0.3 + rng.random(g.N)) * np.exp(2j * np.pi * rng.random(g.N))
    ab = O.symmetric_gauge(g, 0.2) + 0.1 * rng.normal(size=g.Nab)
    abei = O.symmetric_gauge(g, 0.05)
    em = 1.0 + 0.1 * rng.normal(size=g.N)
    k2, H = 16.0, 0.2
    return dict(Nx=13, Ny=11, dx=0.45, dy=0.5, kappa2=k2, H=H, mt=mt, psi=psi, ab=ab, abei=abei, eps_map=em,
                F=O.free_energy(g, k2, 0.0, em, H, mt, psi, abei, ab),
                J=O.current_density(g, k2, H, abei, ab),
                Js=O.supercurrent_density(g, mt, psi, abei, ab),
                B=O.magnetic_field(g, abei, ab))


def cg_kinf_13():
    g = O.Geom(13, 13, 0.5, 0.5)
    psi0 = O.random_psi(g.N, 0.5, seed=8)
    ab0 = O.symmetric_gauge(g, 0.1)
    st = O.OracleState(g, psi0, ab0, kappa=np.inf, H=0.1)
    st.cg(n_iter=8)
    return dict(Nx=13, Ny=13, dx=0.5, dy=0.5, H=0.1, psi0=psi0, ab0=ab0, psi=st.psi,
                energies=np.array(st.cg_energies))


def sample_rows(Ny, k):
    """k rows spread over [0, Ny - 1], always with both edges and their
    neighbours (the rows the boundary terms touch)."""
    rows = set(np.linspace(0, Ny - 1, k - 4).round().astype(int).tolist()) | {0, 1, Ny - 2, Ny - 1}
    return np.array(sorted(rows))


def rows_of(g, rows, psi, ab, J):
    """psi, a, J_a at the sampled rows and b, J_b at the sampled rows below
    the top row (b links exist for rows 0..Ny-2)."""
    rb = rows[rows <= g.Ny - 2]
    return dict(psi_rows=g.psi2(psi)[rows], a_rows=g.a2(ab)[rows], b_rows=g.b2(ab)[rb],
                Ja_rows=g.a2(J)[rows], Jb_rows=g.b2(J)[rb], rows=rows, rows_b=rb)


def config_state(Nx, dx, H, kappa, sigma, seed, mask_seed=None, eps_seed=None):
    """The initial oracle state of a seeded benchmark instance: random psi
    (vars.py:93-109 stream), symmetric gauge of H, and optionally the holes
    mask and pinning field of svirl_amd.workloads, flattened as GLSolver
    flattens them; psi zeroed outside the material."""
    from svirl_amd import workloads
    g = O.Geom(Nx, Nx, dx, dx)
    mt = em = None
    if mask_seed is not None:
        mt = np.ascontiguousarray(workloads.holes_mask(g.Nxc, g.Nyc, dx, dx, seed=mask_seed).T).ravel()
    if eps_seed is not None:
        em = np.ascontiguousarray(workloads.pinning_field(Nx, Nx, dx, seed=eps_seed).T).ravel()
    psi = O.random_psi(g.N, 1.0, seed=seed)
    if mt is not None:
        psi = np.where(O.mask_at_nodes(g, mt), psi, 0)
    st = O.OracleState(g, psi, O.symmetric_gauge(g, H), kappa=kappa, sigma=sigma, H=H, eps_map=em, mt=mt,
                       random_t=seed)
    return g, st


def c1_129():
    g, st = config_state(129, 0.5, 0.1, np.inf, 1.0, seed=1)
    st.td(dt=0.1, Nt=1000)
    out = dict(Nx=129, dx=0.5, H=0.1, dt=0.1, Nt=1000, sweeps=np.array(st.sweeps), psi_td=st.psi.copy(),
               F_td=st.free_energy(), J_td=st.current_density())
    cells, vs = O.vortex_cells(g, st.psi, st.ab, st.H)
    out.update(cells_td=cells, vort_td=np.round(vs))
    st.cg(n_iter=1000)
    cells, vs = O.vortex_cells(g, st.psi, st.ab, st.H)
    out.update(psi_cg=st.psi, cg_energies=np.array(st.cg_energies), cells_cg=cells, vort_cg=np.round(vs))
    return out


def c2_1025():
    g, st = config_state(1025, 0.5, 0.1, 5.0, 200.0, seed=2)
    rows = sample_rows(g.Ny, 16)
    st.td(dt=0.1, Nt=3)
    out = dict(Nx=1025, dx=0.5, H=0.1, kappa=5.0, sigma=200.0, dt=0.1, Nt=3, sweeps=np.array(st.sweeps),
               F_td=st.free_energy())
    out.update({k + "_td": v for k, v in rows_of(g, rows, st.psi, st.ab, st.current_density()).items()})
    st.cg(n_iter=5)
    out.update(cg_energies=np.array(st.cg_energies))
    out.update({k + "_cg": v for k, v in rows_of(g, rows, st.psi, st.ab, st.current_density()).items()})
    return out


def c3_4097():
    g, st = config_state(4097, 0.5, 0.05, 5.0, 200.0, seed=3, mask_seed=30, eps_seed=31)
    rows = sample_rows(g.Ny, 10)
    st.td(dt=0.05, Nt=2)
    out = dict(Nx=4097, dx=0.5, H=0.05, kappa=5.0, sigma=200.0, dt=0.05, Nt=2, sweeps=np.array(st.sweeps),
               F_td=st.free_energy())
    out.update({k + "_td": v for k, v in rows_of(g, rows, st.psi, st.ab, st.current_density()).items()})
    return out


def c4_sample():
    """The four smallest instances of Config 4 (workloads.config4_specs),
    3 TD steps of dt = 0.1 each: sweep counts, F, the whole psi and A, and
    the current density (J, or J_s at kappa = inf)."""
    from svirl_amd import workloads
    specs = workloads.config4_specs()
    pick = sorted(range(len(specs)), key=lambda k: (specs[k]["Nx"] * specs[k]["Ny"], k))[:4]
    out = dict(index=np.array(pick), dt=0.1, Nt=3)
    for n, k in enumerate(pick):
        s = specs[k]
        g = O.Geom(s["Nx"], s["Ny"], s["dx"], s["dx"])
        em = np.ascontiguousarray(workloads.pinning_field(s["Nx"], s["Ny"], s["dx"], seed=s["seed"]).T).ravel()
        st = O.OracleState(g, O.random_psi(g.N, 1.0, seed=s["seed"]), O.symmetric_gauge(g, s["H"]),
                           kappa=s["kappa"], sigma=s["sigma"], H=s["H"], eps_map=em, random_t=s["seed"])
        st.td(dt=0.1, Nt=3)
        out.update({f"sweeps_{n}": np.array(st.sweeps), f"F_{n}": st.free_energy(), f"psi_{n}": st.psi,
                    f"ab_{n}": st.ab, f"J_{n}": st.current_density()})
    return out


def c5_sample():
    """Two instances of Config 5 (workloads.config5_specs; 257^2, kappa = 5)
    after one exact steepest-descent step: the direction is -dF/d(psi, A)
    (the oracle's Jacobians with the field part on boundary links halved,
    which makes dF/dA exact) and the step minimises the oracle's F along it
    (Nelder-Mead on (alpha_psi, alpha_A), then polished to 1e-13 in F).
    Stores F before and at the minimum, the two step lengths and the
    current density J at the minimum."""
    import scipy.optimize
    from svirl_amd import workloads
    specs = workloads.config5_specs()
    pick = [0, 1023]
    out = dict(index=np.array(pick))
    for n, k in enumerate(pick):
        s = specs[k]
        g = O.Geom(s["Nx"], s["Ny"], s["dx"], s["dx"])
        em = np.ascontiguousarray(workloads.pinning_field(s["Nx"], s["Ny"], s["dx"], seed=s["seed"]).T).ravel()
        psi = O.random_psi(g.N, 1.0, seed=s["seed"])
        ab = O.symmetric_gauge(g, s["H"])
        k2, H = s["kappa"] ** 2, s["H"]
        d = -O.jacobian_psi(g, k2, 0.0, em, H, None, psi, None, ab)
        jac = O.jacobian_A(g, k2, H, None, psi, None, ab)
        field = O.jacobian_A(g, k2, H, None, np.zeros_like(psi), None, ab)
        ea = np.zeros((g.Nya, g.Nxa), bool)
        ea[0, :] = ea[-1, :] = True
        eb = np.zeros((g.Nyb, g.Nxb), bool)
        eb[:, 0] = eb[:, -1] = True
        dA = -(jac - 0.5 * field * np.concatenate([ea.ravel(), eb.ravel()]))

        def f(x):
            return O.free_energy(g, k2, 0.0, em, H, None, psi + x[0] * d, None, ab + x[1] * dA)

        F0 = f((0.0, 0.0))
        # scale the step lengths so that both start at a comparable size
        s0 = np.array([1.0 / np.sqrt(np.vdot(d, d).real), 1.0 / np.sqrt(np.dot(dA, dA))]) * 0.1
        r = scipy.optimize.minimize(lambda y: f(y * s0), x0=np.ones(2), method='Nelder-Mead',
                                    options=dict(xatol=1e-12, fatol=1e-13 * abs(F0), maxiter=4000))
        a1 = r.x * s0
        out.update({f"F0_{n}": F0, f"F1_{n}": float(r.fun), f"alpha_{n}": a1,
                    f"J1_{n}": O.current_density(g, k2, H, None, ab + a1[1] * dA)})
    return out


CASES = {"td_kinf_17": td_kinf_17, "td_k5_mask": td_k5_mask, "a_alias_9": a_alias_9,
         "observables": observables, "cg_kinf_13": cg_kinf_13, "c1_129": c1_129, "c2_1025": c2_1025,
         "c3_4097": c3_4097, "c4_sample": c4_sample, "c5_sample": c5_sample}
SMALL = ["td_kinf_17", "td_k5_mask", "a_alias_9", "observables", "cg_kinf_13"]  # seconds each
DEFAULT = [k for k in CASES if k != "c3_4097"]


def main(names):
    for name in names or DEFAULT:
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **CASES[name]())
        print("wrote", name, flush=True)


if __name__ == "__main__":
    main(sys.argv[1:])
