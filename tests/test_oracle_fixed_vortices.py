"""The oracle's fixed-vortex phase and release (fixed_vortices.py:63-77,
130-137) pinned by what the mathematics fixes: the phase is a line
integral of the irregular vector potential, so it recovers a potential it
was built from (a discrete gradient) exactly, and a single vortex's angle
field through its branch cut; releasing onto a uniform psi leaves a winding
of the vortex's vorticity in the vortex's cell and nowhere else."""
import numpy as np

from oracle import gl_oracle as O


def geom(Nx=23, Ny=17, dx=0.5, dy=0.4):
    return O.Geom(Nx, Ny, dx, dy)


def node_xy(g):
    i, j = np.meshgrid(np.arange(g.Nx), np.arange(g.Ny))
    return g.dx * i, g.dy * j  # [j, i]


def vpi_of_scalar(g, chi):
    """a = d chi / dx on x links, b = d chi / dy on y links ([j, i] views)."""
    a = np.diff(chi, axis=1) / g.dx
    b = np.diff(chi, axis=0) / g.dy
    return np.concatenate([a.ravel(), b.ravel()])


def test_phase_recovers_a_gradient_potential():
    g = geom()
    rng = np.random.default_rng(5)
    chi = rng.normal(size=(g.Ny, g.Nx)).cumsum(axis=0).cumsum(axis=1) * 0.1
    ph = O.fixed_vortices_phase(g, vpi_of_scalar(g, chi))
    assert np.allclose(ph, (chi - chi[0, 0]).ravel(), rtol=0, atol=1e-12)
    # a wrong sign or a transposed operand fails
    assert not np.allclose(O.fixed_vortices_phase(g, -vpi_of_scalar(g, chi)), (chi - chi[0, 0]).ravel(), atol=1e-6)


def test_phase_recovers_a_vortex_angle_through_its_branch_cut():
    """The irregular potential of one vortex as the solver builds it
    (fixed_vortices.py:187-194: v/dx times the x differences of the angle
    about the vortex, v/dy times the y differences, angle taken relative
    to node (0, 0)); the line integral telescopes to that angle, jumps of
    2 pi across the cut included."""
    g = geom()
    x, y = node_xy(g)
    for (x0, y0, v) in ((5.3, 3.1, 1), (2.05, 4.45, -2)):
        ang = np.arctan2(y - y0, x - x0)
        ang = ang - ang[0, 0]
        vpi = v * vpi_of_scalar(g, ang)
        assert np.allclose(O.fixed_vortices_phase(g, vpi), (v * ang).ravel(), atol=1e-12)


def test_release_leaves_the_vortex_winding_in_its_cell():
    g = geom()
    x, y = node_xy(g)
    x0, y0 = 5.25, 3.0  # inside cell (10, 7)
    ang = np.arctan2(y - y0, x - x0)
    ang = ang - ang[0, 0]
    # (|v| = 1: a doubled winding steps by about pi between the corners of
    # its own cell, where the detector's per-edge wrap is ambiguous)
    for v in (1, -1):
        psi = O.release_fixed_vortices(g, np.ones(g.N, complex), v * vpi_of_scalar(g, ang))
        w = O.winding_numbers(g, psi, np.zeros(g.Nab), 0.0)
        expect = np.zeros((g.Ny - 1, g.Nx - 1))
        expect[int(y0 / g.dy), int(x0 / g.dx)] = v
        assert np.allclose(w, expect, atol=1e-9), (v, np.argwhere(np.abs(w) > 1e-6))
