"""Solver facade: gl.solve.td and gl.solve.cg (svirl/solvers/solvers.py)."""
import numpy as np

from .cg import CG
from .td import TD


class Solvers:
    def __init__(self, par, mesh, _vars, params, observables):
        self.par = par
        self.mesh = mesh
        self.vars = _vars
        self.params = params
        self.observables = observables
        self._td = None
        self._cg = None
        self.__detected = False

    def td(self, dt=0.1, Nt=1000, T=None, eqn=None):
        """Time-dependent GL for Nt steps of dt (or until T)
        (solvers.py:41-58)."""
        self.__solve_td(dt=dt, Nt=Nt, T=T, eqn=eqn)
        self.vortices_detected = False

    def cg(self, n_iter=1000):
        """Non-linear conjugate-gradient minimisation of the GL free energy
        (solvers.py:61-71)."""
        self.__solve_cg(n_iter=n_iter)
        self.vortices_detected = False

    @property
    def vortices_detected(self):
        return self.__detected

    @vortices_detected.setter
    def vortices_detected(self, status):
        assert isinstance(status, bool)
        self.__detected = status

    def _init_td(self):
        if self._td is None:
            self._td = TD(self.par, self.mesh, self.vars, self.params, self.observables)

    def _init_cg(self):
        if self._cg is None:
            self._cg = CG(self.par, self.mesh, self.vars, self.params, self.observables)

    def __solve_td(self, dt=None, Nt=None, T=None, eqn=None):
        assert isinstance(Nt, (np.integer, int))
        assert isinstance(dt, (np.floating, float, np.integer, int))
        if eqn is not None:
            assert eqn in ['order_parameter', 'vector_potential']
        if T is not None:
            assert isinstance(T, (np.floating, float, np.integer, int))
        self._init_td()
        self._td._solve(dt=dt, Nt=Nt, T=T, eqn=eqn)

    def __solve_cg(self, n_iter):
        self._init_cg()
        self._cg._solve(n_iter)
