"""FP64 NumPy restatement of svirl's 2D Ginzburg-Landau solver.

TEST INFRASTRUCTURE ONLY: the checker and the CPU baseline. Nothing under
``svirl_amd/`` imports it.

Every function restates one reference kernel or host routine; the docstring
cites the reference file:line it follows (paths relative to the reference
repository root). Arrays use the reference's flat device layouts:

* psi  : complex, node n = i + Nx*j                 (svirl/cuda/common.h:76-81)
* ab   : real, [a (i + Nxa*j), b at Na + i + Nxb*j]  (svirl/cuda/td.h:91-100)
* mt   : bool per cell, i + Nxc*j, or None           (svirl/cuda/td.h:53-56)
* eps  : scalar, or a per-node real map (eps_map), or None

Internally every kernel is evaluated on 2-D views indexed [j, i] (row j,
column i), with the reference's per-thread arithmetic order kept term by
term so that the FP64 results agree with a per-thread evaluation up to the
rounding of cos/sin/atan2.
"""
import math

import numpy as np

TWO_PI = 2.0 * np.pi


# ---------------------------------------------------------------------------
# geometry
# ---------------------------------------------------------------------------

class Geom:
    """Grid sizes and the kernel-side constants (svirl/__init__.py:100-123,
    svirl/cuda/td.h:34-35, observables.h:14,51-53).

    The kernels recompute their constants as ``1.0/(dx*dx)`` etc. from the
    JIT-templated ``dx``; that is what is stored here.
    """

    def __init__(self, Nx, Ny, dx, dy):
        self.Nx, self.Ny = int(Nx), int(Ny)
        self.dx, self.dy = float(dx), float(dy)
        self.N = self.Nx * self.Ny
        self.Nxa, self.Nya = self.Nx - 1, self.Ny
        self.Nxb, self.Nyb = self.Nx, self.Ny - 1
        self.Na = self.Nxa * self.Nya
        self.Nb = self.Nxb * self.Nyb
        self.Nab = self.Na + self.Nb
        self.Nxc, self.Nyc = self.Nx - 1, self.Ny - 1
        self.Nc = self.Nxc * self.Nyc
        self.Lx = self.dx * (self.Nx - 1)
        self.Ly = self.dy * (self.Ny - 1)
        self.idx = 1.0 / self.dx
        self.idy = 1.0 / self.dy
        self.idx2 = 1.0 / (self.dx * self.dx)
        self.idy2 = 1.0 / (self.dy * self.dy)
        self.idxy = 1.0 / (self.dx * self.dy)

    # 2-D views ------------------------------------------------------------
    def psi2(self, psi):
        return np.asarray(psi).reshape(self.Ny, self.Nx)

    def a2(self, ab):
        return np.asarray(ab)[:self.Na].reshape(self.Ny, self.Nxa)

    def b2(self, ab):
        return np.asarray(ab)[self.Na:].reshape(self.Nyb, self.Nx)

    def a_node(self, ab):
        """a on the node grid: [j, i] = a(i, j) for i < Nx-1, else 0."""
        out = np.zeros((self.Ny, self.Nx))
        if ab is not None:
            out[:, :self.Nxa] = self.a2(ab)
        return out

    def b_node(self, ab):
        """b on the node grid: [j, i] = b(i, j) for j < Ny-1, else 0."""
        out = np.zeros((self.Ny, self.Nx))
        if ab is not None:
            out[:self.Nyb, :] = self.b2(ab)
        return out

    def flat_ab(self, a_n, b_n):
        return np.concatenate([a_n[:, :self.Nxa].ravel(), b_n[:self.Nyb, :].ravel()])


def shift(X, dj, di):
    """Y[j, i] = X[j+dj, i+di] where that index exists, else 0."""
    Ny, Nx = X.shape
    Y = np.zeros_like(X)
    js0, js1 = max(0, -dj), min(Ny, Ny - dj)
    is0, is1 = max(0, -di), min(Nx, Nx - di)
    if js1 > js0 and is1 > is0:
        Y[js0:js1, is0:is1] = X[js0 + dj:js1 + dj, is0 + di:is1 + di]
    return Y


def node_cells(g, mt):
    """Material flags of the four cells around every node (td.h:48-57).

    Returns (mm, mp, pm, pp), each bool [j, i]: cells (i-1,j-1), (i-1,j),
    (i,j-1), (i,j), False where the cell lies outside the grid.
    """
    C = np.zeros((g.Ny + 1, g.Nx + 1), dtype=bool)
    if mt is None:
        C[1:g.Ny, 1:g.Nx] = True
    else:
        C[1:g.Ny, 1:g.Nx] = np.asarray(mt, dtype=bool).reshape(g.Nyc, g.Nxc)
    return (C[0:g.Ny, 0:g.Nx], C[1:g.Ny + 1, 0:g.Nx],
            C[0:g.Ny, 1:g.Nx + 1], C[1:g.Ny + 1, 1:g.Nx + 1])


def _f(b):
    return np.where(b, 1.0, 0.0)


def _eps_field(g, eps, eps_map):
    if eps_map is not None:
        return np.asarray(eps_map, dtype=np.float64).reshape(g.Ny, g.Nx)
    return float(eps)


# ---------------------------------------------------------------------------
# device helpers (svirl/cuda/common.h)
# ---------------------------------------------------------------------------

def rand_hash(s):
    """Thomas Wang hash on uint32 (common.h:36-46)."""
    s = np.asarray(s, dtype=np.uint32)
    with np.errstate(over='ignore'):
        s = (s ^ np.uint32(61)) ^ (s >> np.uint32(16))
        s = s * np.uint32(9)
        s = s ^ (s >> np.uint32(4))
        s = s * np.uint32(0x27d4eb2d)
        s = s ^ (s >> np.uint32(15))
    return s


def rand_1(n, t):
    """common.h:49-55: hash(71 n + 9887 t) / 2^32."""
    with np.errstate(over='ignore'):
        m = np.asarray(n, dtype=np.uint32) * np.uint32(71) + np.uint32(9887) * np.uint32(t)
    return 2.0 ** -32 * rand_hash(m).astype(np.float64)


def rand_2(n, t):
    """common.h:58-63: hash(73 n + 9901 t + 1) / 2^32."""
    with np.errstate(over='ignore'):
        m = (np.asarray(n, dtype=np.uint32) * np.uint32(73)
             + np.uint32(9901) * np.uint32(t) + np.uint32(1))
    return 2.0 ** -32 * rand_hash(m).astype(np.float64)


def _u_times(ph, z):
    """U(ph)*z with U(ph) = exp(-i ph) = (cos ph, -sin ph) (common.h:28-33),
    multiplied the pycuda-complex way: (ur*x - ui*y, ur*y + ui*x)."""
    ur = np.cos(ph)
    ui = -np.sin(ph)
    x, y = z.real, z.imag
    return ur * x - ui * y, ur * y + ui * x


def js(psi0, ph, psi1):
    """Im(conj(psi0) U(ph) psi1) as written in common.h:66-73."""
    x0, y0 = psi0.real, psi0.imag
    x1, y1 = psi1.real, psi1.imag
    return (x0 * y1 - y0 * x1) * np.cos(ph) - (x0 * x1 + y0 * y1) * np.sin(ph)


def g_grad(psi0, ph, psi1):
    """|psi1 U(ph) - psi0|^2 (observables.h:238-247)."""
    x1, y1 = psi1.real, psi1.imag
    c, s = np.cos(ph), np.sin(ph)
    re = x1 * c + y1 * s - psi0.real
    im = y1 * c - x1 * s - psi0.imag
    return re * re + im * im


def grad_c(psi0, ph, psi1):
    """psi1 U(ph) - psi0 as a complex (cg.h:304-311)."""
    x1, y1 = psi1.real, psi1.imag
    c, s = np.cos(ph), np.sin(ph)
    return (x1 * c + y1 * s - psi0.real) + 1j * (y1 * c - x1 * s - psi0.imag)


def scaled_residual(r, stop_eps):
    """int32(min(1e4 r / stop_eps, 1e8)) (td.h:124-132, 541-543)."""
    with np.errstate(over='ignore', invalid='ignore'):
        rr = 1.0e4 * r / stop_eps
    rr = np.where(rr > 1.0e8, 1.0e8, rr)
    # a NaN residual converts to 0 (the GPU float-to-int conversion of NaN)
    rr = np.nan_to_num(rr, nan=0.0)
    return int(np.max(np.trunc(rr))) if np.size(rr) else 0


def converged(r2max):
    """td.py:198-202: 1e-4 * float(r2max) < 1.0 (in real_t; identical to
    r2max < 10000 in both float32 and float64)."""
    return 1.0e-4 * float(r2max) < 1.0


# ---------------------------------------------------------------------------
# TDGL kernels (svirl/cuda/td.h)
# ---------------------------------------------------------------------------

def psi_jacobi_sweep(g, dt, eps, eps_map, mt, ab, psi_rhs, psi,
                     langevin_c=0.0, jstep=0, rand_t=1, stop_eps=1e-6):
    """One backward-Euler Jacobi sweep for psi (td.h:4-133).

    Returns (psi_next, psi_rhs_after, r2max_int). psi_rhs_after differs from
    psi_rhs only when Langevin noise is added at jstep 0 (td.h:66-70).
    ``ab`` is the regular vector potential (TD passes vars._vp only,
    td.py:120-142).
    """
    mm, mp, pm, pp = node_cells(g, mt)
    inside = mm | mp | pm | pp
    P = g.psi2(psi)
    R = g.psi2(psi_rhs).astype(np.complex128).copy()
    if jstep == 0 and langevin_c > 1.0e-32:
        n = np.arange(g.N, dtype=np.uint32).reshape(g.Ny, g.Nx)
        noise = (rand_1(n, rand_t) - 0.5) + 1j * (rand_2(n, rand_t) - 0.5)
        R = np.where(inside, R + langevin_c * noise, R)
    rm0 = _f(mm | mp); rp0 = _f(pm | pp); r0m = _f(mm | pm); r0p = _f(mp | pp)
    ig = 1.0
    E = _eps_field(g, eps, eps_map)
    a_n, b_n = g.a_node(ab), g.b_node(ab)

    def term(cond, ph, Pn):
        re, im = _u_times(ph, Pn)
        return np.where(cond, re, 0.0), np.where(cond, im, 0.0)

    m0 = term(mm | mp, -g.dx * shift(a_n, 0, -1), shift(P, 0, -1))
    p0 = term(pm | pp, g.dx * a_n, shift(P, 0, 1))
    q0m = term(mm | pm, -g.dy * shift(b_n, -1, 0), shift(P, -1, 0))
    q0p = term(mp | pp, g.dy * b_n, shift(P, 1, 0))

    diag = 1.0 + dt * (R.real * R.real + R.imag * R.imag - E
                       + ig * (g.idx2 * (rm0 + rp0) + g.idy2 * (r0m + r0p)))
    dtig = dt * ig
    num_re = R.real + dtig * (g.idx2 * (m0[0] + p0[0]) + g.idy2 * (q0m[0] + q0p[0]))
    num_im = R.imag + dtig * (g.idx2 * (m0[1] + p0[1]) + g.idy2 * (q0m[1] + q0p[1]))
    with np.errstate(invalid='ignore', divide='ignore'):
        nre = np.where(inside, num_re / diag, 0.0)
        nim = np.where(inside, num_im / diag, 0.0)
    r = np.maximum(np.abs(nre - P.real), np.abs(nim - P.imag))
    nxt = (nre + 1j * nim).ravel()
    return nxt, R.ravel(), scaled_residual(r, stop_eps)


def a_jacobi_sweep(g, dt, kappa2, rho, H, mt, psi, abi_ab_rhs, ab_rhs, ab,
                   langevin_c=0.0, jstep=0, rand_t=1, stop_eps=1e-6):
    """One Jacobi sweep of sigma dA/dt = J_s - kappa^2 curl curl A
    (td.h:310-463). Returns (ab_next, ab_rhs_after, r2max_int).

    ``abi_ab_rhs`` is the supercurrent source buffer (its values as they
    were before this sweep wrote anything).
    """
    mm, mp, pm, pp = node_cells(g, mt)
    P = g.psi2(psi)
    dt_rho = dt * rho
    dt_rho_kappa2 = dt_rho * kappa2
    rhs = np.array(ab_rhs, dtype=np.float64, copy=True)
    if jstep == 0 and langevin_c > 1.0e-32:
        na = np.arange(g.Na, dtype=np.uint32)
        nb = np.arange(g.Na, g.Nab, dtype=np.uint32)
        rhs[:g.Na] += langevin_c * (rand_1(na, rand_t) - 0.5)
        rhs[g.Na:] += langevin_c * (rand_2(nb, rand_t) - 0.5)
    a = g.a2(ab); b = g.b2(ab)
    src_a = g.a2(abi_ab_rhs); src_b = g.b2(abi_ab_rhs)
    ra = g.a2(rhs); rb = g.b2(rhs)

    # ---- a links: i < Nx-1, all j (td.h:366-410)
    Nxa, Ny = g.Nxa, g.Ny
    j = np.arange(Ny)[:, None] * np.ones((1, Nxa))
    rh = np.where(j == 0, 2.0 * kappa2 * H * g.idy,
                  np.where(j == Ny - 1, -2.0 * kappa2 * H * g.idy, 0.0))
    dd = np.where((j == 0) | (j == Ny - 1), 2.0, 1.0)
    cond = (pm | pp)[:, :Nxa]
    jl = np.where(cond, g.idx * js(P[:, :Nxa], g.dx * src_a, P[:, 1:]), 0.0)
    bn = g.b_node(ab)  # [j, i] = b(i, j), zero row at j = Ny-1
    T1 = np.zeros((Ny, Nxa))
    T1[1:] = g.idy2 * a[:-1] - g.idxy * bn[:-1, :Nxa] + g.idxy * bn[:-1, 1:]
    T2 = np.zeros((Ny, Nxa))
    T2[:-1] = g.idy2 * a[1:] + g.idxy * bn[:-1, :Nxa] - g.idxy * bn[:-1, 1:]
    a_next = ((ra + dt_rho * (jl + rh) + dt_rho_kappa2 * dd * (T1 + T2))
              / (1.0 + 2.0 * dt_rho_kappa2 * g.idy2))
    res_a = np.abs(a_next - a)

    # ---- b links: all i, j < Ny-1 (td.h:412-456)
    Nx, Nyb = g.Nx, g.Nyb
    i = np.ones((Nyb, 1)) * np.arange(Nx)[None, :]
    rh = np.where(i == 0, -2.0 * kappa2 * H * g.idx,
                  np.where(i == Nx - 1, 2.0 * kappa2 * H * g.idx, 0.0))
    dd = np.where((i == 0) | (i == Nx - 1), 2.0, 1.0)
    cond = (mp | pp)[:Nyb, :]
    jl = np.where(cond, g.idy * js(P[:Nyb], g.dy * src_b, P[1:]), 0.0)
    an = g.a_node(ab)  # [j, i] = a(i, j), zero column at i = Nx-1
    T1 = np.zeros((Nyb, Nx))
    T1[:, 1:] = g.idx2 * b[:, :-1] - g.idxy * an[:Nyb, :-1] + g.idxy * an[1:, :-1]
    T2 = np.zeros((Nyb, Nx))
    T2[:, :-1] = g.idx2 * b[:, 1:] + g.idxy * an[:Nyb, :-1] - g.idxy * an[1:, :-1]
    b_next = ((rb + dt_rho * (jl + rh) + dt_rho_kappa2 * dd * (T1 + T2))
              / (1.0 + 2.0 * dt_rho_kappa2 * g.idx2))
    res_b = np.abs(b_next - b)

    r = max(scaled_residual(res_a, stop_eps) if res_a.size else 0,
            scaled_residual(res_b, stop_eps) if res_b.size else 0)
    return np.concatenate([a_next.ravel(), b_next.ravel()]), rhs, r


def phase_lock(psi, lock_ns):
    """td.h:295-307: psi[n] <- |psi[n]| at the listed nodes."""
    out = np.array(psi, copy=True)
    out[lock_ns] = np.abs(out[lock_ns])
    return out


# ---------------------------------------------------------------------------
# observables (svirl/cuda/observables.h)
# ---------------------------------------------------------------------------

def _ab_sum(g, abei, ab):
    """abei + ab with NULL meaning absent, as the kernels add them."""
    tot = np.zeros(g.Nab)
    if abei is not None:
        tot = tot + np.asarray(abei, dtype=np.float64)
    if ab is not None:
        tot = tot + np.asarray(ab, dtype=np.float64)
    return tot


def _curl(g, v):
    """idx*(b[i+1,j]-b[i,j]) - idy*(a[i,j+1]-a[i,j]) per cell, [j, i]."""
    a, b = g.a2(v), g.b2(v)
    return g.idx * (b[:, 1:] - b[:, :-1]) - g.idy * (a[1:, :] - a[:-1, :])


def magnetic_field(g, abei, ab):
    """observables.h:4-37: B per cell from abei and ab (each optional)."""
    B = np.zeros((g.Nyc, g.Nxc))
    if abei is not None:
        B = B + _curl(g, abei)
    if ab is not None:
        B = B + _curl(g, ab)
    return B.ravel()


def _link_curlcurl(g, H, abei, ab):
    """The ghost-cell curl-curl accumulator 'c' on every a and b link, in
    the kernels' accumulation order (observables.h:66-155, cg.h:179-290):
    H source, abei diagonal (2/dd), ab diagonal (2), then for the lower
    neighbour abei (unscaled) and ab (scaled by dd), then the upper one.
    Returns (ca [j, i] (Ny, Nxa), cb [j, i] (Nyb, Nx)).
    """
    Ny, Nxa, Nx, Nyb = g.Ny, g.Nxa, g.Nx, g.Nyb
    srcs = [(v, v is ab) for v in (abei, ab) if v is not None]
    # a links
    j = np.arange(Ny)[:, None] * np.ones((1, Nxa))
    dd = np.where((j == 0) | (j == Ny - 1), 2.0, 1.0)
    c = np.where(j == 0, 0.0 - 2.0 * H * g.idy, np.where(j == Ny - 1, 0.0 + 2.0 * H * g.idy, 0.0))
    for v, reg in srcs:
        a = g.a2(v)
        c = c + ((2.0 * g.idy2) * a if reg else 2.0 / dd * g.idy2 * a)
    for v, reg in srcs:
        a = g.a2(v); bn = g.b_node(v)
        t1 = np.zeros((Ny, Nxa))
        t1[1:] = -g.idy2 * a[:-1] + g.idxy * bn[:-1, :Nxa] - g.idxy * bn[:-1, 1:]
        c = np.where(j > 0, c + (dd * t1 if reg else dd / dd * t1), c)
    for v, reg in srcs:
        a = g.a2(v); bn = g.b_node(v)
        t2 = np.zeros((Ny, Nxa))
        t2[:-1] = -g.idy2 * a[1:] - g.idxy * bn[:-1, :Nxa] + g.idxy * bn[:-1, 1:]
        c = np.where(j < Ny - 1, c + (dd * t2 if reg else dd / dd * t2), c)
    ca = c
    # b links
    i = np.ones((Nyb, 1)) * np.arange(Nx)[None, :]
    dd = np.where((i == 0) | (i == Nx - 1), 2.0, 1.0)
    c = np.where(i == 0, 0.0 + 2.0 * H * g.idx, np.where(i == Nx - 1, 0.0 - 2.0 * H * g.idx, 0.0))
    for v, reg in srcs:
        b = g.b2(v)
        c = c + ((2.0 * g.idx2) * b if reg else 2.0 / dd * g.idx2 * b)
    for v, reg in srcs:
        b = g.b2(v); an = g.a_node(v)
        t1 = np.zeros((Nyb, Nx))
        t1[:, 1:] = -g.idx2 * b[:, :-1] + g.idxy * an[:Nyb, :-1] - g.idxy * an[1:, :-1]
        c = np.where(i > 0, c + (dd * t1 if reg else dd / dd * t1), c)
    for v, reg in srcs:
        b = g.b2(v); an = g.a_node(v)
        t2 = np.zeros((Nyb, Nx))
        t2[:, :-1] = -g.idx2 * b[:, 1:] - g.idxy * an[:Nyb, :-1] + g.idxy * an[1:, :-1]
        c = np.where(i < Nx - 1, c + (dd * t2 if reg else dd / dd * t2), c)
    return ca, c


def current_density(g, kappa2, H, abei, ab):
    """observables.h:40-156: kappa^2 (curl B - H) on links with the
    doubled-ghost boundary; the abei terms are not scaled by dd."""
    ca, cb = _link_curlcurl(g, H, abei, ab)
    return np.concatenate([(kappa2 * ca).ravel(), (kappa2 * cb).ravel()])


def supercurrent_density(g, mt, psi, abei, ab):
    """observables.h:159-235: r_Du * id * js on links touching material."""
    mm, mp, pm, pp = node_cells(g, mt)
    P = g.psi2(psi)
    tot = _ab_sum(g, abei, ab)
    A, B = g.a2(tot), g.b2(tot)
    condx = (pm | pp)[:, :g.Nxa]
    rp0 = 0.5 * (_f(pm) + _f(pp))[:, :g.Nxa]
    jx = np.where(condx, rp0 * g.idx * js(P[:, :-1], g.dx * A, P[:, 1:]), 0.0)
    condy = (mp | pp)[:g.Nyb, :]
    r0p = 0.5 * (_f(mp) + _f(pp))[:g.Nyb, :]
    jy = np.where(condy, r0p * g.idy * js(P[:-1, :], g.dy * B, P[1:, :]), 0.0)
    return np.concatenate([jx.ravel(), jy.ravel()])


def _du_weights(mm, mp, pm, pp):
    rm0 = 0.5 * (_f(mm) + _f(mp)); rp0 = 0.5 * (_f(pm) + _f(pp))
    r0m = 0.5 * (_f(mm) + _f(pm)); r0p = 0.5 * (_f(mp) + _f(pp))
    gw = 0.25 * (rm0 + rp0 + r0m + r0p)
    return rm0, rp0, r0m, r0p, gw


def free_energy_density(g, kappa2, eps, eps_map, H, mt, psi, abei, ab):
    """Per-node contribution e[j, i] of free_energy_pseudodensity
    (observables.h:250-357), before the dx*dy factor."""
    mm, mp, pm, pp = node_cells(g, mt)
    inside = mm | mp | pm | pp
    rm0, rp0, r0m, r0p, gw = _du_weights(mm, mp, pm, pp)
    P = g.psi2(psi)
    E = _eps_field(g, eps, eps_map)
    p2 = P.real * P.real + P.imag * P.imag
    e = np.where(inside, gw * (0.5 * p2 - E) * p2, 0.0)
    tot = _ab_sum(g, abei, ab)
    an, bn = g.a_node(tot), g.b_node(tot)
    ex = np.where(inside & (pm | pp), rp0 * g.idx2 * g_grad(P, g.dx * an, shift(P, 0, 1)), 0.0)
    e = e + ex
    ey = np.where(inside & (mp | pp), r0p * g.idy2 * g_grad(P, g.dy * bn, shift(P, 1, 0)), 0.0)
    e = e + ey
    if kappa2 > 0.0:
        dB = np.full((g.Nyc, g.Nxc), -H)
        if abei is not None:
            dB = dB + _curl(g, abei)
        if ab is not None:
            dB = dB + _curl(g, ab)
        em = np.zeros((g.Ny, g.Nx))
        em[:g.Nyc, :g.Nxc] = kappa2 * dB * dB
        e = e + em
    return e


def free_energy(g, kappa2, eps, eps_map, H, mt, psi, abei, ab):
    """observables.py:124-149 + observables.h:250-362: dx*dy*sum(e)."""
    e = free_energy_density(g, kappa2, eps, eps_map, H, mt, psi, abei, ab)
    return g.dx * g.dy * math.fsum(e.ravel())


# ---------------------------------------------------------------------------
# CG kernels (svirl/cuda/cg.h)
# ---------------------------------------------------------------------------

def _g_grad_jac_psi(psi0, ph, psi1):
    """cg.h:4-12: 2 (psi0 - psi1 U(ph))."""
    x1, y1 = psi1.real, psi1.imag
    c, s = np.cos(ph), np.sin(ph)
    return 2.0 * (psi0.real - (x1 * c + y1 * s)) + 1j * 2.0 * (psi0.imag - (y1 * c - x1 * s))


def jacobian_psi(g, kappa2, eps, eps_map, H, mt, psi, abei, ab):
    """free_energy_jacobian_psi (cg.h:15-121): dF/d(Re psi, Im psi) * dxdy,
    all four links; returns the value the kernel adds to a zeroed buffer."""
    mm, mp, pm, pp = node_cells(g, mt)
    inside = mm | mp | pm | pp
    rm0, rp0, r0m, r0p, gw = _du_weights(mm, mp, pm, pp)
    P = g.psi2(psi)
    E = _eps_field(g, eps, eps_map)
    p = P.real * P.real + P.imag * P.imag - E
    jac = np.where(inside, 2.0 * gw * p * P.real, 0.0) + 1j * np.where(inside, 2.0 * gw * p * P.imag, 0.0)
    tot = _ab_sum(g, abei, ab)
    an, bn = g.a_node(tot), g.b_node(tot)
    for cond, r, idd, ph, Pn in (
            (mm | mp, rm0, g.idx2, -g.dx * shift(an, 0, -1), shift(P, 0, -1)),
            (pm | pp, rp0, g.idx2, g.dx * an, shift(P, 0, 1)),
            (mm | pm, r0m, g.idy2, -g.dy * shift(bn, -1, 0), shift(P, -1, 0)),
            (mp | pp, r0p, g.idy2, g.dy * bn, shift(P, 1, 0))):
        t = _g_grad_jac_psi(P, ph, Pn)
        jac = jac + np.where(inside & cond, r * idd * t.real, 0.0) + 1j * np.where(inside & cond, r * idd * t.imag, 0.0)
    return (g.dx * g.dy * jac).ravel()


def jacobian_A(g, kappa2, H, mt, psi, abei, ab):
    """free_energy_jacobian_A (cg.h:124-301): 2 dxdy [kappa^2 curlcurl
    (ghost BC) - r_Du id js] on each link; the value added to a zeroed buffer."""
    mm, mp, pm, pp = node_cells(g, mt)
    P = g.psi2(psi)
    ca, cb = _link_curlcurl(g, H, abei, ab)
    ca, cb = ca * kappa2, cb * kappa2
    tot = _ab_sum(g, abei, ab)
    A, B = g.a2(tot), g.b2(tot)
    condx = (pm | pp)[:, :g.Nxa]
    rp0 = 0.5 * (_f(pm) + _f(pp))[:, :g.Nxa]
    ca = np.where(condx, ca + (-rp0 * g.idx * js(P[:, :-1], g.dx * A, P[:, 1:])), ca)
    condy = (mp | pp)[:g.Nyb, :]
    r0p = 0.5 * (_f(mp) + _f(pp))[:g.Nyb, :]
    cb = np.where(condy, cb + (-r0p * g.idy * js(P[:-1, :], g.dy * B, P[1:, :])), cb)
    return np.concatenate([(2.0 * g.dx * g.dy * ca).ravel(), (2.0 * g.dx * g.dy * cb).ravel()])


def _ph_links(g, abei, ab):
    """Link phases as the coefficient kernels form them: ph = 0;
    ph += d*abei; ph += d*ab (cg.h:413-416, 599-602). Returns node-grid
    arrays (ph_a [j, i] for a(i, j), ph_b [j, i] for b(i, j))."""
    pa = np.zeros((g.Ny, g.Nx)); pb = np.zeros((g.Ny, g.Nx))
    for v in (abei, ab):
        if v is not None:
            pa = pa + g.dx * g.a_node(v)
            pb = pb + g.dy * g.b_node(v)
    return pa, pb


def coef_psi(g, kappa2, eps, H, mt, psi, dpsi, abei, ab):
    """free_energy_conjgrad_coef_psi (cg.h:314-474): c0..c4 of
    F(psi + alpha dpsi) at fixed A. The kernel reads only the scalar
    epsilon (no epsilon_spatial lookup) and the magnetic term uses ab alone.
    """
    mm, mp, pm, pp = node_cells(g, mt)
    inside = mm | mp | pm | pp
    rm0, rp0, r0m, r0p, gw = _du_weights(mm, mp, pm, pp)
    P = g.psi2(psi); D = g.psi2(dpsi)
    p2 = P.real * P.real + P.imag * P.imag
    d2 = D.real * D.real + D.imag * D.imag
    tw = 2.0 * (P.real * D.real + P.imag * D.imag)
    eps = float(eps)
    c = [np.zeros((g.Ny, g.Nx)) for _ in range(5)]
    c[4] = np.where(inside, gw * 0.5 * d2 * d2, 0.0)
    c[3] = np.where(inside, gw * tw * d2, 0.0)
    c[2] = np.where(inside, gw * (-eps * d2 + 0.5 * tw * tw + p2 * d2), 0.0)
    c[1] = np.where(inside, gw * tw * (p2 - eps), 0.0)
    c[0] = np.where(inside, gw * (0.5 * p2 - eps) * p2, 0.0)
    pha, phb = _ph_links(g, abei, ab)
    for cond, r, idd, ph, Pn, Dn in (
            (pm | pp, rp0, g.idx2, pha, shift(P, 0, 1), shift(D, 0, 1)),
            (mp | pp, r0p, g.idy2, phb, shift(P, 1, 0), shift(D, 1, 0))):
        m = inside & cond
        c[2] = c[2] + np.where(m, r * idd * g_grad(D, ph, Dn), 0.0)
        z = np.conj(grad_c(P, ph, Pn)) * grad_c(D, ph, Dn)
        c[1] = c[1] + np.where(m, r * idd * 2.0 * z.real, 0.0)
        c[0] = c[0] + np.where(m, r * idd * g_grad(P, ph, Pn), 0.0)
    if kappa2 > 0.0:
        dB = _curl(g, ab) - H
        c[0][:g.Nyc, :g.Nxc] = c[0][:g.Nyc, :g.Nxc] + kappa2 * dB * dB
    dxdy = g.dx * g.dy
    return np.array([math.fsum(ci.ravel()) * dxdy for ci in c])


def coef(g, kappa2, eps, H, mt, psi, dpsi, abei, ab, dab):
    """free_energy_conjgrad_coef (cg.h:477-731): 17 coefficients, exact in
    alpha_psi, 4th-order Taylor in alpha_A, exact magnetic quadratic.
    Returns them in kernel order c00..c04, c10..c14, c20..c24, c30, c40.
    Only the scalar epsilon is read (no epsilon_spatial lookup in the kernel).
    """
    mm, mp, pm, pp = node_cells(g, mt)
    inside = mm | mp | pm | pp
    rm0, rp0, r0m, r0p, gw = _du_weights(mm, mp, pm, pp)
    P = g.psi2(psi); D = g.psi2(dpsi)
    p2 = P.real * P.real + P.imag * P.imag
    d2 = D.real * D.real + D.imag * D.imag
    tw = 2.0 * (P.real * D.real + P.imag * D.imag)
    eps = float(eps)
    Z = lambda: np.zeros((g.Ny, g.Nx))  # noqa: E731
    C = {k: Z() for k in ('00', '01', '02', '03', '04', '10', '11', '12', '13', '14',
                          '20', '21', '22', '23', '24', '30', '40')}
    C['00'] = np.where(inside, gw * (0.5 * p2 - eps) * p2, 0.0)
    C['10'] = np.where(inside, gw * tw * (p2 - eps), 0.0)
    C['20'] = np.where(inside, gw * (-eps * d2 + 0.5 * tw * tw + p2 * d2), 0.0)
    C['30'] = np.where(inside, gw * tw * d2, 0.0)
    C['40'] = np.where(inside, gw * 0.5 * d2 * d2, 0.0)
    tot = _ab_sum(g, abei, ab)
    pha, phb = _ph_links(g, abei, ab)
    dan, dbn = g.a_node(dab), g.b_node(dab)
    for cond, r, idd, d, ph, dph_src, Pn, Dn in (
            (pm | pp, rp0, g.idx2, g.dx, pha, dan, shift(P, 0, 1), shift(D, 0, 1)),
            (mp | pp, r0p, g.idy2, g.dy, phb, dbn, shift(P, 1, 0), shift(D, 1, 0))):
        m = inside & cond
        dph = d * dph_src
        dph2 = dph * dph
        w = r * idd
        C['00'] = C['00'] + np.where(m, w * g_grad(P, ph, Pn), 0.0)
        z = np.conj(grad_c(P, ph, Pn)) * grad_c(D, ph, Dn)
        C['10'] = C['10'] + np.where(m, w * 2.0 * z.real, 0.0)
        C['20'] = C['20'] + np.where(m, w * g_grad(D, ph, Dn), 0.0)
        eip = np.cos(ph) + 1j * np.sin(ph)  # U(-ph)
        for key, zz in (('0', P * eip * np.conj(Pn)),
                        ('1', P * eip * np.conj(Dn) + D * eip * np.conj(Pn)),
                        ('2', D * eip * np.conj(Dn))):
            C[key + '1'] = C[key + '1'] + np.where(m, w * 2.0 * zz.imag * dph, 0.0)
            C[key + '2'] = C[key + '2'] + np.where(m, w * zz.real * dph2, 0.0)
            C[key + '3'] = C[key + '3'] + np.where(m, -(w / 3.0) * zz.imag * dph2 * dph, 0.0)
            C[key + '4'] = C[key + '4'] + np.where(m, -(w / 12.0) * zz.real * dph2 * dph2, 0.0)
    if kappa2 > 0.0:
        BH = _curl(g, tot) - H
        dB = _curl(g, dab)
        sl = (slice(0, g.Nyc), slice(0, g.Nxc))
        C['00'][sl] = C['00'][sl] + kappa2 * BH * BH
        C['01'][sl] = C['01'][sl] + kappa2 * 2.0 * BH * dB
        C['02'][sl] = C['02'][sl] + kappa2 * dB * dB
    dxdy = g.dx * g.dy
    order = ('00', '01', '02', '03', '04', '10', '11', '12', '13', '14',
             '20', '21', '22', '23', '24', '30', '40')
    return np.array([math.fsum(C[k].ravel()) * dxdy for k in order])


def coef17_to_matrix(c17):
    """cg.py:369-373: pack the 17 sums into the 5x5 polyval2d matrix."""
    c = np.zeros((5, 5))
    c[0, :] = c17[0:5]
    c[1, :] = c17[5:10]
    c[2, :] = c17[10:15]
    c[3, 0] = c17[15]
    c[4, 0] = c17[16]
    return c


def pr_beta(jac, jac_prev):
    """cg.py:422-449 + utils.h:12-70,139-146: max(sum g(g-gp) / sum gp^2, 0)
    with CUDA fmax semantics (a NaN quotient yields 0)."""
    if np.iscomplexobj(jac):
        num = math.fsum((jac.real * (jac.real - jac_prev.real) + jac.imag * (jac.imag - jac_prev.imag)).ravel())
        den = math.fsum((jac_prev.real * jac_prev.real + jac_prev.imag * jac_prev.imag).ravel())
    else:
        num = math.fsum((jac * (jac - jac_prev)).ravel())
        den = math.fsum((jac_prev * jac_prev).ravel())
    with np.errstate(divide='ignore', invalid='ignore'):
        q = np.float64(num) / np.float64(den)
    return float(np.fmax(q, 0.0))


def alpha_psi_min(c):
    """cg.py:227-235: smallest real root >= 0 of c1 + 2c2 a + 3c3 a^2 + 4c4 a^3."""
    am = np.polynomial.polynomial.polyroots([c[1], 2.0 * c[2], 3.0 * c[3], 4.0 * c[4]])
    am = am[np.isclose(am.imag, 0)].real
    am = am[am >= 0]
    return np.min(am)


def alpha_min(c, alpha0=(0.0, 0.0), tol=1e-8):
    """cg.py:378-419: scipy BFGS on polyval2d(alpha_psi, alpha_A, c)."""
    import scipy.optimize
    cj0 = np.polynomial.polynomial.polyder(c, axis=0)
    cj1 = np.polynomial.polynomial.polyder(c, axis=1)

    def f(al):
        return np.polynomial.polynomial.polyval2d(al[0], al[1], c)

    def jf(al):
        return np.array([np.polynomial.polynomial.polyval2d(al[0], al[1], cj0),
                         np.polynomial.polynomial.polyval2d(al[0], al[1], cj1)])

    r = scipy.optimize.minimize(f, x0=np.array(alpha0), jac=jf, method='BFGS', tol=tol)
    return r.x


# ---------------------------------------------------------------------------
# state initialisation (svirl/vars, svirl/mesh)
# ---------------------------------------------------------------------------

def random_psi(N, level=1.0, seed=None):
    """vars.py:93-109 (legacy MT19937 stream, |psi| draw first, phase second)."""
    if seed is not None:
        np.random.seed(seed)
    return (1.0 - level * np.random.rand(N)) * np.exp(level * 1.0j * np.pi * (2.0 * np.random.rand(N) - 1.0))


def xy_a_grid(g):
    """grid.py:49-60 (meshgrid 'ij', shape (Nxa, Nya))."""
    x = np.linspace(0.5 * g.dx, g.Lx - 0.5 * g.dx, num=g.Nxa, endpoint=True)
    y = np.linspace(0.0, g.Ly, num=g.Nya, endpoint=True)
    return np.meshgrid(x, y, indexing='ij')


def xy_b_grid(g):
    """grid.py:63-74 (meshgrid 'ij', shape (Nxb, Nyb))."""
    x = np.linspace(0.0, g.Lx, num=g.Nxb, endpoint=True)
    y = np.linspace(0.5 * g.dy, g.Ly - 0.5 * g.dy, num=g.Nyb, endpoint=True)
    return np.meshgrid(x, y, indexing='ij')


def symmetric_gauge(g, H, Lx=None, Ly=None):
    """params.py:135-170 with reset: a = -g (y - Ly/2) H, b = (1-g)(x - Lx/2) H,
    g = 1/2; returned in the flat device layout."""
    Lx = g.Lx if Lx is None else Lx
    Ly = g.Ly if Ly is None else Ly
    _, yg = xy_a_grid(g)
    xg, _ = xy_b_grid(g)
    a = np.zeros((g.Nxa, g.Nya)); b = np.zeros((g.Nxb, g.Nyb))
    a -= 0.5 * (yg - 0.5 * Ly) * H
    b += (1.0 - 0.5) * (xg - 0.5 * Lx) * H
    return np.concatenate([a.T.ravel(), b.T.ravel()])


def mask_at_nodes(g, mt):
    """grid.py:160-171: a node is in material if any adjacent cell is."""
    mm, mp, pm, pp = node_cells(g, mt)
    return (mm | mp | pm | pp).ravel()


# ---------------------------------------------------------------------------
# host loops
# ---------------------------------------------------------------------------

class OracleState:
    """CPU restatement of one GLSolver's state and its solver loops.

    Fields mirror the reference objects: psi (vars._psi), ab (vars._vp),
    abei (params._vpei; all zeros by default, params.py:36,287-303),
    eps / eps_map (params._epsilon), kappa (params._kappa), sigma, H, mt,
    vpi (fixed_vortices._vpi, the irregular vector potential of pinned
    vortices, or None) and lock_ns (fixed_vortices._phase_lock_ns or None).
    """

    def __init__(self, g, psi, ab, kappa=np.inf, sigma=1.0, H=0.0, eps=1.0,
                 eps_map=None, mt=None, abei=None, stop_psi=1e-6, stop_A=1e-6,
                 langevin_psi=0.0, langevin_A=0.0, random_t=1, convergence_rtol=1e-6,
                 vpi=None, lock_ns=None):
        self.g = g
        self.psi = np.array(psi, dtype=np.complex128)
        self.ab = np.array(ab, dtype=np.float64)
        self.abei = None if abei is None else np.array(abei, dtype=np.float64)
        self.kappa = float(kappa)
        self.sigma = float(sigma)
        self.rho = 1.0 / self.sigma
        self.H = float(H)
        self.eps = float(eps)
        self.eps_map = None if eps_map is None else np.array(eps_map, dtype=np.float64).ravel()
        self.mt = None if mt is None else np.array(mt, dtype=bool).ravel()
        # td.py:51-66 clamps (FP64)
        self.stop_psi = max(float(stop_psi), 1e-12)
        self.stop_A = max(float(stop_A), 1e-12)
        self.langevin_psi = float(langevin_psi)
        self.langevin_A = float(langevin_A)
        self.random_t = int(random_t) & 0xffffffff
        self.convergence_rtol = convergence_rtol
        self.vpi = None if vpi is None else np.array(vpi, dtype=np.float64)
        self.lock_ns = None if lock_ns is None else np.array(lock_ns, dtype=np.int64)
        # TD buffers: the reference's __gab_next persists between steps
        self._ab_next = np.zeros(g.Nab)
        self.sweeps = []  # (S_psi, S_A) per step
        # CG persistent state (cg.py:54-71)
        self._cg_ready = False

    @property
    def solveA(self):
        return not np.isposinf(self.kappa)

    @property
    def kappa2(self):
        """params.py:101-105: kappa^2, or -1 when kappa = inf."""
        return self.kappa ** 2 if self.solveA else -1.0

    # -- TDGL ---------------------------------------------------------------
    def ab_preprocess(self):
        """td.py:118-140: vp += vpi over the first N entries only (xpy_r
        with N = number of nodes, utils.h:117-124), before a psi solve."""
        if self.vpi is not None:
            N = self.g.N
            self.ab[:N] = self.ab[:N] + self.vpi[:N]

    def ab_postprocess(self):
        """td.py:143-154: vp -= vpi over the first N entries (xmy_r)."""
        if self.vpi is not None:
            N = self.g.N
            self.ab[:N] = self.ab[:N] - self.vpi[:N]

    def td_psi_solve(self, dt):
        """td.py:157-218: the Jacobi loop, then the phase lock of the
        listed nodes (td.py:205-216, td.h:295-307)."""
        g = self.g
        psi_rhs = self.psi.copy()
        cur = self.psi
        nsw = 0
        for j in range(1024):
            nxt, psi_rhs, r = psi_jacobi_sweep(
                g, dt, self.eps_scalar, self.eps_map, self.mt, self.ab, psi_rhs, cur,
                self.langevin_psi, j, self.random_t, self.stop_psi)
            cur = nxt
            nsw = j + 1
            if converged(r):
                break
        self.random_t = (self.random_t + 1) & 0xffffffff
        self.psi = cur
        if self.lock_ns is not None:
            self.psi = phase_lock(self.psi, self.lock_ns)
        return nsw

    def td_A_solve(self, dt):
        """td.py:252-325 with the aliasing of td.py:264-303: the supercurrent
        source is the buffer object that held A^n when the solve started,
        which the ping-pong swap overwrites on every odd sweep."""
        g = self.g
        X = self.ab            # buffer object bound as abi_ab_rhs
        Y = self._ab_next
        src = X
        if self.vpi is not None:
            # td.py:256-266: vpi += vp over the first N entries; vpi is then
            # the source of every sweep
            N = g.N
            self.vpi[:N] = self.vpi[:N] + X[:N]
            src = self.vpi
        rhs = X.copy()
        cur, nxt = X, Y
        nsw = 0
        for j in range(1024):
            src_vals = src.copy()  # each thread reads its own src before writing
            vals, rhs, r = a_jacobi_sweep(
                g, dt, self.kappa2, self.rho, self.H, self.mt, self.psi,
                src_vals, rhs, cur, self.langevin_A, j, self.random_t, self.stop_A)
            nxt[:] = vals
            cur, nxt = nxt, cur
            nsw = j + 1
            if converged(r):
                break
        self.random_t = (self.random_t + 1) & 0xffffffff
        self.ab, self._ab_next = cur, nxt
        if self.vpi is not None:
            # td.py:318-325: vpi -= vp (now A^{n+1}) over the first N entries
            N = g.N
            self.vpi[:N] = self.vpi[:N] - self.ab[:N]
        return nsw

    def td(self, dt=0.1, Nt=1000):
        """td.py:342-366: per step, the psi solve between the vpi pre- and
        post-processing, then (if kappa < inf) the A solve."""
        for _ in range(int(Nt)):
            self.ab_preprocess()
            s_psi = self.td_psi_solve(dt)
            self.ab_postprocess()
            s_A = self.td_A_solve(dt) if self.solveA else 0
            self.sweeps.append((s_psi, s_A))

    def td_order_parameter(self, dt=0.1, Nt=1000):
        """td.py:221-231 (eqn='order_parameter'): the vpi preprocessing
        once before the loop, the postprocessing after every step."""
        self.ab_preprocess()
        for _ in range(int(Nt)):
            self.sweeps.append((self.td_psi_solve(dt), 0))
            self.ab_postprocess()

    def td_vector_potential(self, dt=0.1, Nt=1000):
        """td.py:328-339 (eqn='vector_potential'): A solves only."""
        if not self.solveA:
            return
        for _ in range(int(Nt)):
            self.sweeps.append((0, self.td_A_solve(dt)))

    # -- observables ----------------------------------------------------------
    @property
    def eps_scalar(self):
        """params.py:79-83: the scalar handed to every kernel; 0.0 when an
        epsilon map is set (the map then replaces it where a kernel reads
        epsilon_spatial, but the CG coefficient kernels never do)."""
        return 0.0 if self.eps_map is not None else self.eps

    def free_energy(self):
        return free_energy(self.g, self.kappa2, self.eps_scalar, self.eps_map, self.H,
                           self.mt, self.psi, self.abei, self.ab)

    def current_density(self):
        """observables.py:86-110 (J_s when kappa = inf)."""
        if not self.solveA:
            return self.supercurrent_density()
        return current_density(self.g, self.kappa2, self.H, self.abei, self.ab)

    def supercurrent_density(self):
        return supercurrent_density(self.g, self.mt, self.psi, self.abei, self.ab)

    def magnetic_field(self):
        return magnetic_field(self.g, self.abei, self.ab)

    # -- CG -------------------------------------------------------------------
    def _cg_init(self):
        """cg.py:54-71: buffers and betas persist across cg() calls."""
        g = self.g
        if not self._cg_ready:
            self.dir_psi = np.zeros(g.N, dtype=np.complex128)
            self.jac_psi_prev = np.zeros(g.N, dtype=np.complex128)
            self.beta_psi = 0.0
            self.dir_A = np.zeros(g.Nab)
            self.jac_A_prev = np.zeros(g.Nab)
            self.beta_A = 0.0
            self._cg_ready = True

    def _jac_psi(self):
        return jacobian_psi(self.g, self.kappa2, self.eps_scalar, self.eps_map, self.H,
                            self.mt, self.psi, self.abei, self.ab)

    def _jac_A(self):
        return jacobian_A(self.g, self.kappa2, self.H, self.mt, self.psi, self.abei, self.ab)

    def cg(self, n_iter=1000):
        """cg.py:238-322 (kappa = inf) and cg.py:452-551 (kappa < inf)."""
        self._cg_init()
        self.cg_energies = []
        g = self.g
        if not self.solveA:
            self.dir_psi[:] = 0.0  # cg.py:264
        for i in range(int(n_iter)):
            jac = self._jac_psi()
            jacA = self._jac_A() if self.solveA else None
            if i > 0:
                self.beta_psi = pr_beta(jac, self.jac_psi_prev)
                if self.solveA:
                    self.beta_A = pr_beta(jacA, self.jac_A_prev)
            self.dir_psi = self.beta_psi * self.dir_psi - jac
            if self.solveA:
                self.dir_A = self.beta_A * self.dir_A - jacA
                c17 = coef(g, self.kappa2, self.eps_scalar, self.H, self.mt, self.psi,
                           self.dir_psi, self.abei, self.ab, self.dir_A)
                a_psi, a_A = alpha_min(coef17_to_matrix(c17))
                self.psi = a_psi * self.dir_psi + self.psi
                self.ab = a_A * self.dir_A + self.ab
                self.jac_A_prev = jacA.copy()
            else:
                c5 = coef_psi(g, self.kappa2, self.eps_scalar, self.H, self.mt, self.psi,
                              self.dir_psi, self.abei, self.ab)
                a_psi = alpha_psi_min(c5)
                self.psi = a_psi * self.dir_psi + self.psi
            self.jac_psi_prev = jac.copy()
            E = self.free_energy()
            self.cg_energies.append(E)
            if i > 0 and np.abs(self.cg_energies[i] / self.cg_energies[i - 1] - 1.0) < self.convergence_rtol:
                break
        return len(self.cg_energies)

    # -- vortex detector --------------------------------------------------------
    def vortices(self):
        return vortex_detector(self.g, self.psi, self.ab, self.H)


def winding_numbers(g, psi, ab, H):
    """The per-cell vorticity of vortex_detector.py:55-72, all cells at once;
    np.angle / np.mod as in the reference. Returns v[j, i] (Nyc, Nxc)."""
    t = np.angle(g.psi2(psi))
    a, b = g.a2(ab), g.b2(ab)
    t00 = t[:-1, :-1]; tp0 = t[:-1, 1:]; tpp = t[1:, 1:]; t0p = t[1:, :-1]
    v = -(0.5 / np.pi) * (
        np.mod(tp0 - t00 - g.dx * a[:-1, :] + np.pi, 2.0 * np.pi)
        + np.mod(tpp - tp0 - g.dy * b[:, 1:] + np.pi, 2.0 * np.pi)
        + np.mod(t0p - tpp + g.dx * a[1:, :] + np.pi, 2.0 * np.pi)
        + np.mod(t00 - t0p + g.dy * b[:, :-1] + np.pi, 2.0 * np.pi)
        - 4.0 * np.pi
        + g.dx * g.dy * H)
    return v


def vortex_cells(g, psi, ab, H):
    """Flagged cells in ascending n = i + Nxc*j and their vorticity
    (vortex_detector.py:55-72)."""
    v = winding_numbers(g, psi, ab, H).ravel()
    flag = (np.abs(v) > 0.5) & (np.abs(v - np.round(v)) < 0.1)
    n = np.nonzero(flag)[0]
    return n, v[n]


def triangulate(g, cells, vs, psi, a_ai, b_bi):
    """vortex_detector.py:75-147: sub-cell vortex positions from the zero
    lines of gauge-rotated Re/Im psi in each flagged cell.
    a_ai, b_bi are 2-D [i, j] arrays (host layout)."""
    def find_zero(x1, y1, f1, x2, y2, f2):
        return (f2 * x1 - x2 * f1) / (f2 - f1), (f2 * y1 - y2 * f1) / (f2 - f1)

    def find_zero_line(x1, y1, f1, x2, y2, f2, x3, y3, f3, x4, y4, f4):
        xy = []
        if f2 * f1 < -1e-10: xy.append(find_zero(x2, y2, f2, x1, y1, f1))
        if f3 * f2 < -1e-10: xy.append(find_zero(x3, y3, f3, x2, y2, f2))
        if f4 * f3 < -1e-10: xy.append(find_zero(x4, y4, f4, x3, y3, f3))
        if f1 * f4 < -1e-10: xy.append(find_zero(x1, y1, f1, x4, y4, f4))
        return xy

    def find_intersection(l1x1, l1y1, l1x2, l1y2, l2x1, l2y1, l2x2, l2y2):
        D = (l1x1 - l1x2) * (l2y1 - l2y2) - (l1y1 - l1y2) * (l2x1 - l2x2)
        ph = np.arctan2(D, (l1y1 - l1y2) * (l2y1 - l2y2) - (l1x1 - l1x2) * (l2x1 - l2x2))
        ph = np.mod(np.abs(ph), 0.5 * np.pi)
        if np.abs(ph) > 1e-10:
            ix = ((l1x1 * l1y2 - l1y1 * l1x2) * (l2x1 - l2x2) - (l1x1 - l1x2) * (l2x1 * l2y2 - l2y1 * l2x2)) / D
            iy = ((l1x1 * l1y2 - l1y1 * l1x2) * (l2y1 - l2y2) - (l1y1 - l1y2) * (l2x1 * l2y2 - l2y1 * l2x2)) / D
            return ix, iy, ph
        return np.nan, np.nan, ph

    P = g.psi2(psi).T  # [i, j]
    out = []
    for n, v in zip(cells, vs):
        i, j = int(n) % g.Nxc, int(n) // g.Nxc
        ip, jp = i + 1, j + 1
        x, y = g.dx * i, g.dy * j
        inta_00 = g.dx * a_ai[i, j]
        inta_0p = g.dx * a_ai[i, jp]
        intb_00 = g.dy * b_bi[i, j]
        intb_p0 = g.dy * b_bi[ip, j]
        tpsi_00 = P[i, j]
        tpsi_p0 = P[ip, j] * np.exp(-1j * (0.75 * inta_00 + 0.25 * (intb_00 + inta_0p - intb_p0)))
        tpsi_pp = P[ip, jp] * np.exp(-1j * (0.5 * (inta_00 + intb_p0) + 0.5 * (intb_00 + inta_0p)))
        tpsi_0p = P[i, jp] * np.exp(-1j * (0.75 * intb_00 + 0.25 * (inta_00 + intb_p0 - inta_0p)))
        re_xy = find_zero_line(x, y, np.real(tpsi_00), x + g.dx, y, np.real(tpsi_p0),
                               x + g.dx, y + g.dy, np.real(tpsi_pp), x, y + g.dy, np.real(tpsi_0p))
        im_xy = find_zero_line(x, y, np.imag(tpsi_00), x + g.dx, y, np.imag(tpsi_p0),
                               x + g.dx, y + g.dy, np.imag(tpsi_pp), x, y + g.dy, np.imag(tpsi_0p))
        if len(re_xy) == 2 and len(im_xy) == 2:
            ix, iy, ph = find_intersection(re_xy[0][0], re_xy[0][1], re_xy[1][0], re_xy[1][1],
                                           im_xy[0][0], im_xy[0][1], im_xy[1][0], im_xy[1][1])
            if x - g.dx < ix < x + 2.0 * g.dx and y - g.dy < iy < y + 2.0 * g.dy:
                out.append([ix, iy, np.round(v)])
    if out:
        return np.array(out, dtype=np.float64)
    return np.zeros((0, 3))


def fixed_vortices_phase(g, vpi):
    """svirl/vars/fixed_vortices.py:130-137 (_ab_phase of the irregular
    vector potential, :155-158): the phase at node (i, j) is the line
    integral of vpi from node (0, 0), first up the column i = 0 along the
    b links (dy times the running sum), then along row j over the a links
    (dx times the running sum), the two added. Flat node layout."""
    a = g.a2(vpi)                                                   # a[j, i]
    b = g.b2(vpi)                                                   # b[j, i]
    ca = g.dx * np.concatenate([np.zeros((g.Nya, 1)), a], axis=1).cumsum(axis=1)
    cb = g.dy * np.concatenate([np.zeros((1, g.Nx)), b], axis=0).cumsum(axis=0)
    return (np.repeat(cb[:, :1], g.Nx, axis=1) + ca).ravel()


def release_fixed_vortices(g, psi, vpi):
    """svirl/vars/fixed_vortices.py:63-77 (fixed_vortices_release): psi times
    exp(-i phase) with the phase of the irregular vector potential; the
    caller then drops vpi and the phase-locked nodes."""
    return psi * np.exp(-1.0j * fixed_vortices_phase(g, vpi))


def vortex_detector(g, psi, ab, H, abi=None):
    """vortex_detector.py:31-156: winding uses vars._vp only; the
    triangulation uses vp + vpi (abi = irregular vector potential)."""
    cells, vs = vortex_cells(g, psi, ab, H)
    tot = np.array(ab, dtype=np.float64) if abi is None else np.asarray(ab) + np.asarray(abi)
    a_ai = g.a2(tot).T
    b_bi = g.b2(tot).T
    det = triangulate(g, cells, vs, psi, a_ai, b_bi)
    return det[:, 0], det[:, 1], det[:, 2]
