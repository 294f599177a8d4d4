"""Per-thread loop interpreter of svirl's kernels, for tiny grids.

TEST INFRASTRUCTURE ONLY. Each function walks n over the reference's 1-D
launch (n = i + Nx*j) and evaluates the kernel body for one thread with
Python scalars, in the statement order of svirl/cuda/*.h. It exists to pin
the index handling of the vectorised restatement in ``gl_oracle``.
"""
import math


def _U(ph):
    return complex(math.cos(ph), -math.sin(ph))


def _cells(g, mt, i, j):
    """td.h:48-57."""
    Nx, Ny, Nxc = g.Nx, g.Ny, g.Nxc
    im, ip, jm, jp = i - 1, i + 1, j - 1, j + 1
    mm = (i > 0) and (j > 0)
    mp = (i > 0) and (jp < Ny)
    pm = (ip < Nx) and (j > 0)
    pp = (ip < Nx) and (jp < Ny)
    if mt is not None:
        if mm: mm = bool(mt[im + Nxc * jm])
        if mp: mp = bool(mt[im + Nxc * j])
        if pm: pm = bool(mt[i + Nxc * jm])
        if pp: pp = bool(mt[i + Nxc * j])
    return mm, mp, pm, pp


def _js(psi0, ph, psi1):
    x0, y0, x1, y1 = psi0.real, psi0.imag, psi1.real, psi1.imag
    return (x0 * y1 - y0 * x1) * math.cos(ph) - (x0 * x1 + y0 * y1) * math.sin(ph)


def _g_grad(psi0, ph, psi1):
    x1, y1 = psi1.real, psi1.imag
    c, s = math.cos(ph), math.sin(ph)
    re = x1 * c + y1 * s - psi0.real
    im = y1 * c - x1 * s - psi0.imag
    return re * re + im * im


def _grad_c(psi0, ph, psi1):
    x1, y1 = psi1.real, psi1.imag
    c, s = math.cos(ph), math.sin(ph)
    return complex(x1 * c + y1 * s, y1 * c - x1 * s) - psi0


def _residual(r, stop_eps):
    r = 1.0e4 * r / stop_eps
    if r > 1.0e8:
        r = 1.0e8
    return int(r)


def psi_sweep(g, dt, eps, eps_map, mt, ab, psi_rhs, psi, stop_eps):
    """td.h:4-133, Langevin off."""
    Nx, Ny, Nxa, Nxb, Na = g.Nx, g.Ny, g.Nxa, g.Nxb, g.Na
    out = [0j] * g.N
    r2 = 0
    for n in range(Nx * Ny):
        i, j = n % Nx, n // Nx
        im, ip, jm, jp = i - 1, i + 1, j - 1, j + 1
        mm, mp, pm, pp = _cells(g, mt, i, j)
        psi00 = complex(psi[n])
        if mm or mp or pm or pp:
            prhs = complex(psi_rhs[n])
            r_m0 = 1.0 if (mm or mp) else 0.0
            r_p0 = 1.0 if (pm or pp) else 0.0
            r_0m = 1.0 if (mm or pm) else 0.0
            r_0p = 1.0 if (mp or pp) else 0.0
            ig = 1.0
            e = float(eps_map[n]) if eps_map is not None else eps
            t_m0 = t_p0 = t_0m = t_0p = 0j
            if mm or mp:
                t_m0 = r_m0 * _U(-g.dx * ab[im + Nxa * j]) * complex(psi[im + Nx * j])
            if pm or pp:
                t_p0 = r_p0 * _U(g.dx * ab[i + Nxa * j]) * complex(psi[ip + Nx * j])
            if mm or pm:
                t_0m = r_0m * _U(-g.dy * ab[Na + i + Nxb * jm]) * complex(psi[i + Nx * jm])
            if mp or pp:
                t_0p = r_0p * _U(g.dy * ab[Na + i + Nxb * j]) * complex(psi[i + Nx * jp])
            diag = 1.0 + dt * (prhs.real * prhs.real + prhs.imag * prhs.imag - e
                               + ig * (g.idx2 * (r_m0 + r_p0) + g.idy2 * (r_0m + r_0p)))
            num = prhs + dt * ig * (g.idx2 * (t_m0 + t_p0) + g.idy2 * (t_0m + t_0p))
            nxt = complex(num.real / diag, num.imag / diag)
        else:
            nxt = 0j
        out[n] = nxt
        r = max(abs(nxt.real - psi00.real), abs(nxt.imag - psi00.imag))
        r2 = max(r2, _residual(r, stop_eps))
    return out, r2


def a_sweep(g, dt, kappa2, rho, H, mt, psi, src, ab_rhs, ab, stop_eps):
    """td.h:310-463, Langevin off. ``src`` is abi_ab_rhs."""
    Nx, Ny, Nxa, Nya, Nxb, Nyb, Na = g.Nx, g.Ny, g.Nxa, g.Nya, g.Nxb, g.Nyb, g.Na
    dt_rho = dt * rho
    dt_rho_kappa2 = dt_rho * kappa2
    out = [0.0] * g.Nab
    r2 = 0
    for n in range(Nx * Ny):
        i, j = n % Nx, n // Nx
        im, ip, jm, jp = i - 1, i + 1, j - 1, j + 1
        psi0 = complex(psi[i + Nx * j])
        rmax = 0.0
        _, mp, pm, pp = _cells(g, mt, i, j)
        if i < Nxa and j < Nya:
            m = i + Nxa * j
            rh, dd = 0.0, 1.0
            if j == 0:
                rh, dd = 2.0 * kappa2 * H * g.idy, 2.0
            elif jp == Nya:
                rh, dd = -2.0 * kappa2 * H * g.idy, 2.0
            jl = 0.0
            if pm or pp:
                jl = g.idx * _js(psi0, g.dx * src[m], complex(psi[ip + Nx * j]))
            t1 = (g.idy2 * ab[i + Nxa * jm] - g.idxy * ab[Na + i + Nxb * jm]
                  + g.idxy * ab[Na + ip + Nxb * jm]) if j > 0 else 0.0
            t2 = (g.idy2 * ab[i + Nxa * jp] + g.idxy * ab[Na + i + Nxb * j]
                  - g.idxy * ab[Na + ip + Nxb * j]) if jp < Nya else 0.0
            v = (ab_rhs[m] + dt_rho * (jl + rh) + dt_rho_kappa2 * dd * (t1 + t2)) / (1.0 + 2.0 * dt_rho_kappa2 * g.idy2)
            out[m] = v
            rmax = max(rmax, abs(v - ab[m]))
        if i < Nxb and j < Nyb:
            m = Na + i + Nxb * j
            rh, dd = 0.0, 1.0
            if i == 0:
                rh, dd = -2.0 * kappa2 * H * g.idx, 2.0
            elif ip == Nxb:
                rh, dd = 2.0 * kappa2 * H * g.idx, 2.0
            jl = 0.0
            if mp or pp:
                jl = g.idy * _js(psi0, g.dy * src[m], complex(psi[i + Nx * jp]))
            t1 = (g.idx2 * ab[Na + im + Nxb * j] - g.idxy * ab[im + Nxa * j]
                  + g.idxy * ab[im + Nxa * jp]) if i > 0 else 0.0
            t2 = (g.idx2 * ab[Na + ip + Nxb * j] + g.idxy * ab[i + Nxa * j]
                  - g.idxy * ab[i + Nxa * jp]) if ip < Nxb else 0.0
            v = (ab_rhs[m] + dt_rho * (jl + rh) + dt_rho_kappa2 * dd * (t1 + t2)) / (1.0 + 2.0 * dt_rho_kappa2 * g.idx2)
            out[m] = v
            rmax = max(rmax, abs(v - ab[m]))
        r2 = max(r2, _residual(rmax, stop_eps))
    return out, r2


def _p(abei, ab, k):
    p = 0.0
    if abei is not None: p += abei[k]
    if ab is not None: p += ab[k]
    return p


def magnetic_field(g, abei, ab):
    """observables.h:4-37."""
    out = [0.0] * g.Nc
    for n in range(g.Nc):
        i, j = n % g.Nxc, n // g.Nxc
        B = 0.0
        for v in (abei, ab):
            if v is not None:
                B += (g.idx * (v[g.Na + i + 1 + g.Nxb * j] - v[g.Na + i + g.Nxb * j])
                      - g.idy * (v[i + g.Nxa * (j + 1)] - v[i + g.Nxa * j]))
        out[n] = B
    return out


def _curlcurl(g, H, abei, ab, i, j, which):
    """The accumulator of current_density / free_energy_jacobian_A for one
    link (observables.h:66-155, cg.h:179-290), before the kappa2 factor."""
    Nxa, Nya, Nxb, Na = g.Nxa, g.Nya, g.Nxb, g.Na
    im, ip, jm, jp = i - 1, i + 1, j - 1, j + 1
    c, dd = 0.0, 1.0
    if which == 'a':
        n = i + Nxa * j
        if j == 0: c -= 2.0 * H * g.idy; dd = 2.0
        elif jp == Nya: c += 2.0 * H * g.idy; dd = 2.0
        if abei is not None: c += 2.0 / dd * g.idy2 * abei[n]
        if ab is not None: c += 2.0 * g.idy2 * ab[n]
        if j > 0:
            for v, s in ((abei, None), (ab, dd)):
                if v is None: continue
                t = -g.idy2 * v[i + Nxa * jm] + g.idxy * v[Na + i + Nxb * jm] - g.idxy * v[Na + ip + Nxb * jm]
                c += (dd / dd * t) if s is None else s * t
        if jp < Nya:
            for v, s in ((abei, None), (ab, dd)):
                if v is None: continue
                t = -g.idy2 * v[i + Nxa * jp] - g.idxy * v[Na + i + Nxb * j] + g.idxy * v[Na + ip + Nxb * j]
                c += (dd / dd * t) if s is None else s * t
    else:
        n = Na + i + Nxb * j
        if i == 0: c += 2.0 * H * g.idx; dd = 2.0
        elif ip == Nxb: c -= 2.0 * H * g.idx; dd = 2.0
        if abei is not None: c += 2.0 / dd * g.idx2 * abei[n]
        if ab is not None: c += 2.0 * g.idx2 * ab[n]
        if i > 0:
            for v, s in ((abei, None), (ab, dd)):
                if v is None: continue
                t = -g.idx2 * v[Na + im + Nxb * j] + g.idxy * v[im + Nxa * j] - g.idxy * v[im + Nxa * jp]
                c += (dd / dd * t) if s is None else s * t
        if ip < Nxb:
            for v, s in ((abei, None), (ab, dd)):
                if v is None: continue
                t = -g.idx2 * v[Na + ip + Nxb * j] - g.idxy * v[i + Nxa * j] + g.idxy * v[i + Nxa * jp]
                c += (dd / dd * t) if s is None else s * t
    return c


def current_density(g, kappa2, H, abei, ab):
    """observables.h:40-156."""
    out = [0.0] * g.Nab
    for n in range(g.N):
        i, j = n % g.Nx, n // g.Nx
        if i < g.Nxa and j < g.Nya:
            out[i + g.Nxa * j] = kappa2 * _curlcurl(g, H, abei, ab, i, j, 'a')
        if i < g.Nxb and j < g.Nyb:
            out[g.Na + i + g.Nxb * j] = kappa2 * _curlcurl(g, H, abei, ab, i, j, 'b')
    return out


def supercurrent_density(g, mt, psi, abei, ab):
    """observables.h:159-235 (EDGES_DU)."""
    out = [0.0] * g.Nab
    for n in range(g.N):
        i, j = n % g.Nx, n // g.Nx
        psi0 = complex(psi[i + g.Nx * j])
        _, mp, pm, pp = _cells(g, mt, i, j)
        if i < g.Nxa and j < g.Nya:
            m = i + g.Nxa * j
            jl = 0.0
            if pm or pp:
                r = 0.5 * ((1.0 if pm else 0.0) + (1.0 if pp else 0.0))
                jl = r * g.idx * _js(psi0, g.dx * _p(abei, ab, m), complex(psi[i + 1 + g.Nx * j]))
            out[m] = jl
        if i < g.Nxb and j < g.Nyb:
            m = g.Na + i + g.Nxb * j
            jl = 0.0
            if mp or pp:
                r = 0.5 * ((1.0 if mp else 0.0) + (1.0 if pp else 0.0))
                jl = r * g.idy * _js(psi0, g.dy * _p(abei, ab, m), complex(psi[i + g.Nx * (j + 1)]))
            out[m] = jl
    return out


def _du(mm, mp, pm, pp):
    f = lambda b: 1.0 if b else 0.0  # noqa: E731
    r_m0 = 0.5 * (f(mm) + f(mp)); r_p0 = 0.5 * (f(pm) + f(pp))
    r_0m = 0.5 * (f(mm) + f(pm)); r_0p = 0.5 * (f(mp) + f(pp))
    return r_m0, r_p0, r_0m, r_0p, 0.25 * (r_m0 + r_p0 + r_0m + r_0p)


def free_energy_terms(g, kappa2, eps, eps_map, H, mt, psi, abei, ab):
    """observables.h:250-357: the per-thread e values (before dx*dy)."""
    es = []
    Nx, Ny, Nxa, Nxb, Na = g.Nx, g.Ny, g.Nxa, g.Nxb, g.Na
    for n in range(g.N):
        i, j = n % Nx, n // Nx
        ip, jp = i + 1, j + 1
        e = 0.0
        mm, mp, pm, pp = _cells(g, mt, i, j)
        if mm or mp or pm or pp:
            _, r_p0, _, r_0p, gw = _du(mm, mp, pm, pp)
            psi00 = complex(psi[n])
            p2 = psi00.real * psi00.real + psi00.imag * psi00.imag
            ee = float(eps_map[n]) if eps_map is not None else eps
            e += gw * (0.5 * p2 - ee) * p2
            if pm or pp:
                e += r_p0 * g.idx2 * _g_grad(psi00, g.dx * _p(abei, ab, i + Nxa * j), complex(psi[ip + Nx * j]))
            if mp or pp:
                e += r_0p * g.idy2 * _g_grad(psi00, g.dy * _p(abei, ab, Na + i + Nxb * j), complex(psi[i + Nx * jp]))
        if kappa2 > 0.0 and i < g.Nxc and j < g.Nyc:
            dB = -H
            for v in (abei, ab):
                if v is not None:
                    dB += (g.idx * (v[Na + ip + Nxb * j] - v[Na + i + Nxb * j])
                           - g.idy * (v[i + Nxa * jp] - v[i + Nxa * j]))
            e += kappa2 * dB * dB
        es.append(e)
    return es


def jacobian_psi(g, kappa2, eps, eps_map, H, mt, psi, abei, ab):
    """cg.h:15-121."""
    Nx, Nxa, Nxb, Na = g.Nx, g.Nxa, g.Nxb, g.Na
    out = [0j] * g.N
    for n in range(g.N):
        i, j = n % Nx, n // Nx
        im, ip, jm, jp = i - 1, i + 1, j - 1, j + 1
        gj = 0j
        mm, mp, pm, pp = _cells(g, mt, i, j)
        if mm or mp or pm or pp:
            r_m0, r_p0, r_0m, r_0p, gw = _du(mm, mp, pm, pp)
            psi0 = complex(psi[n])
            ee = float(eps_map[n]) if eps_map is not None else eps
            p = psi0.real * psi0.real + psi0.imag * psi0.imag - ee
            gj += 2.0 * gw * p * psi0

            def ggj(ph, psi1):
                x1, y1 = psi1.real, psi1.imag
                c, s = math.cos(ph), math.sin(ph)
                return 2.0 * (psi0 - complex(x1 * c + y1 * s, y1 * c - x1 * s))
            if mm or mp:
                gj += r_m0 * g.idx2 * ggj(-g.dx * _p(abei, ab, im + Nxa * j), complex(psi[im + Nx * j]))
            if pm or pp:
                gj += r_p0 * g.idx2 * ggj(g.dx * _p(abei, ab, i + Nxa * j), complex(psi[ip + Nx * j]))
            if mm or pm:
                gj += r_0m * g.idy2 * ggj(-g.dy * _p(abei, ab, Na + i + Nxb * jm), complex(psi[i + Nx * jm]))
            if mp or pp:
                gj += r_0p * g.idy2 * ggj(g.dy * _p(abei, ab, Na + i + Nxb * j), complex(psi[i + Nx * jp]))
        out[n] = g.dx * g.dy * gj
    return out


def jacobian_A(g, kappa2, H, mt, psi, abei, ab):
    """cg.h:124-301."""
    out = [0.0] * g.Nab
    for n in range(g.N):
        i, j = n % g.Nx, n // g.Nx
        psi0 = complex(psi[n])
        _, mp, pm, pp = _cells(g, mt, i, j)
        if i < g.Nxa and j < g.Nya:
            m = i + g.Nxa * j
            gj = _curlcurl(g, H, abei, ab, i, j, 'a') * kappa2
            if pm or pp:
                r = 0.5 * ((1.0 if pm else 0.0) + (1.0 if pp else 0.0))
                gj += -r * g.idx * _js(psi0, g.dx * _p(abei, ab, m), complex(psi[i + 1 + g.Nx * j]))
            out[m] = 2.0 * g.dx * g.dy * gj
        if i < g.Nxb and j < g.Nyb:
            m = g.Na + i + g.Nxb * j
            gj = _curlcurl(g, H, abei, ab, i, j, 'b') * kappa2
            if mp or pp:
                r = 0.5 * ((1.0 if mp else 0.0) + (1.0 if pp else 0.0))
                gj += -r * g.idy * _js(psi0, g.dy * _p(abei, ab, m), complex(psi[i + g.Nx * (j + 1)]))
            out[m] = 2.0 * g.dx * g.dy * gj
    return out


def coef_terms(g, kappa2, eps, H, mt, psi, dpsi, abei, ab, dab=None):
    """cg.h:314-474 (dab None: 5 coefficients) or cg.h:477-731 (17)."""
    Nx, Nxa, Nxb, Na = g.Nx, g.Nxa, g.Nxb, g.Na
    L = 5 if dab is None else 17
    tot = [0.0] * L
    for n in range(g.N):
        i, j = n % Nx, n // Nx
        ip, jp = i + 1, j + 1
        C = [0.0] * L
        mm, mp, pm, pp = _cells(g, mt, i, j)
        if mm or mp or pm or pp:
            _, r_p0, _, r_0p, gw = _du(mm, mp, pm, pp)
            p0 = complex(psi[n]); d0 = complex(dpsi[n])
            p2 = p0.real * p0.real + p0.imag * p0.imag
            d2 = d0.real * d0.real + d0.imag * d0.imag
            tw = 2.0 * (p0.real * d0.real + p0.imag * d0.imag)
            kin = [gw * (0.5 * p2 - eps) * p2, gw * tw * (p2 - eps),
                   gw * (-eps * d2 + 0.5 * tw * tw + p2 * d2), gw * tw * d2, gw * 0.5 * d2 * d2]
            if L == 5:
                for k in range(5): C[k] += kin[k]
            else:
                C[0] += kin[0]; C[5] += kin[1]; C[10] += kin[2]; C[15] += kin[3]; C[16] += kin[4]
            links = []
            if pm or pp:
                links.append((r_p0, g.idx2, g.dx, i + Nxa * j, ip + Nx * j))
            if mp or pp:
                links.append((r_0p, g.idy2, g.dy, Na + i + Nxb * j, i + Nx * jp))
            for r, idd, d, k, nb in links:
                ph = 0.0
                if abei is not None: ph += d * abei[k]
                if ab is not None: ph += d * ab[k]
                p1 = complex(psi[nb]); d1 = complex(dpsi[nb])
                z = _grad_c(p0, ph, p1).conjugate() * _grad_c(d0, ph, d1)
                if L == 5:
                    C[2] += r * idd * _g_grad(d0, ph, d1)
                    C[1] += r * idd * 2.0 * z.real
                    C[0] += r * idd * _g_grad(p0, ph, p1)
                else:
                    dph = d * dab[k]; dph2 = dph * dph
                    C[0] += r * idd * _g_grad(p0, ph, p1)
                    C[5] += r * idd * 2.0 * z.real
                    C[10] += r * idd * _g_grad(d0, ph, d1)
                    Um = _U(-ph)
                    for base, zz in ((0, p0 * Um * p1.conjugate()),
                                     (5, p0 * Um * d1.conjugate() + d0 * Um * p1.conjugate()),
                                     (10, d0 * Um * d1.conjugate())):
                        C[base + 1] += r * idd * 2.0 * zz.imag * dph
                        C[base + 2] += r * idd * zz.real * dph2
                        C[base + 3] += -r * idd / 3.0 * zz.imag * dph2 * dph
                        C[base + 4] += -r * idd / 12.0 * zz.real * dph2 * dph2
        if kappa2 > 0.0 and i < g.Nxc and j < g.Nyc:
            if L == 5:
                dB = (g.idx * (ab[Na + ip + Nxb * j] - ab[Na + i + Nxb * j])
                      - g.idy * (ab[i + Nxa * jp] - ab[i + Nxa * j]) - H)
                C[0] += kappa2 * dB * dB
            else:
                a00 = _p(abei, ab, i + Nxa * j); a0p = _p(abei, ab, i + Nxa * jp)
                b00 = _p(abei, ab, Na + i + Nxb * j); bp0 = _p(abei, ab, Na + ip + Nxb * j)
                BH = g.idx * (bp0 - b00) - g.idy * (a0p - a00) - H
                dB = (g.idx * (dab[Na + ip + Nxb * j] - dab[Na + i + Nxb * j])
                      - g.idy * (dab[i + Nxa * jp] - dab[i + Nxa * j]))
                C[0] += kappa2 * BH * BH
                C[1] += kappa2 * 2.0 * BH * dB
                C[2] += kappa2 * dB * dB
        for k in range(L):
            tot[k] += C[k]
    return [t * (g.dx * g.dy) for t in tot]
