"""Test infrastructure only: an FP64 NumPy restatement of svirl's 2D
Ginzburg-Landau kernels and host loops.

Nothing in the shipped package (``svirl_amd``) imports this package. Only
``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` use it, and only as the checker / CPU baseline.

Modules
-------
gl_oracle : vectorised restatement of every kernel in svirl/cuda/*.h and of
            the host loops in svirl/solvers/{td,cg}.py,
            svirl/observables/*.py, svirl/vars/*.py (cited per function).
loops     : literal per-thread loop interpreter of the same kernels, for tiny
            grids; it pins the vectorised restatement's index handling.
"""
