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

Run: python tests/golden/make_golden.py
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


CASES = {"td_kinf_17": td_kinf_17, "td_k5_mask": td_k5_mask, "a_alias_9": a_alias_9,
         "observables": observables, "cg_kinf_13": cg_kinf_13}


def main():
    for name, fn in CASES.items():
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **fn())
        print("wrote", name)


if __name__ == "__main__":
    main()
