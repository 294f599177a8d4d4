// Row-marching forms of the two passes of a CG iteration over one grid:
// the gradient pass (free energy, both Jacobians, Polak-Ribiere sums, the
// state update optionally fused in) and the coefficient pass with the
// direction update. Same layout of the work as the TD sweeps: strips of 62
// columns, lanes 0 and 63 of a wave carry the columns either side, a wave
// marches down a range of rows keeping a window of its column in
// registers, and the values of the columns to the left and right come from
// the neighbouring lanes. Every field value is loaded once per pass (the
// node-parallel forms load psi five times and each link value up to four
// times, through the caches), and each link's sin/cos is evaluated once:
// the left and down link factors of a node are its neighbours' right and up
// link factors with the orientation reversed, (sin, cos)(-x) = (-sin x,
// cos x), which sincos_r returns bitwise. The node arithmetic is grad_eval /
// coef_eval's, so the results equal the node-parallel forms to the bit;
// only the order of the FP64 partial sums differs.
#pragma once

#include "glx_fields.hpp"
#include "glx_td.hpp"

namespace glx {

// Waves of a row-marching CG pass: about 16 rows per wave, at most 4096
// waves (1024 workgroups, the partials' capacity). The passes hand out the
// rows row-major (for_rows_rm): waves of consecutive index march the same
// rows of adjacent strips, so the chip reads whole grid rows together.
inline long cg_row_waves(int Nx, int Ny) {
    const long T = (long)nstrips_of<1>(Nx) * Ny;
    long W = (T + 15) / 16;
    if (W > 4 * RED_MAX_BLOCKS) W = 4 * RED_MAX_BLOCKS;
    return W < 1 ? 1 : W;
}

// The link values a lane holds for one link family at one row: abei (E)
// and ab (B), the latter with the state update applied (UPD).
template <typename R> struct LinkPair {
    R e, b;
};

// Rows per batch of the double-buffered march (two batches in flight).
template <typename R> struct CGRows { static constexpr int B = 1; };
template <> struct CGRows<float> { static constexpr int B = 2; };

// Every load of a marching step is issued unconditionally, from a valid
// address: a buffer that is absent is replaced by one of at least its size
// (psi holds 2 N reals, more than the Na + Nb links, the N nodes of an
// epsilon map or the bytes of the cell flags) and its values are dropped by
// a select where they are used. Straight-line loads let the compiler keep a
// whole batch in flight while it computes the previous one.

template <typename R, bool withA, bool UPD>
__global__ void __launch_bounds__(256) k_cg_grad_rows(Geo g, R kappa2, R eps, const R *eps_map, R H, const uint8_t *mt,
                                                      const cplx<R> *psi, const R *abei, const R *ab, cplx<R> *jac,
                                                      const cplx<R> *jac_prev, R *jacA, const R *jacA_prev,
                                                      double *partials, StateUpdate<R> S, const double *alpha, long W) {
    constexpr int B = CGRows<R>::B;
    const GeoR<R> q(g);
    if constexpr (UPD) {
        S.ap = R(alpha[0]);
        S.aa = R(alpha[1]);
    }
    const R dxdy = q.dx * q.dy, s2 = R(2) * q.dx * q.dy;
    const bool he = abei != nullptr, hb = ab != nullptr, has_mt = mt != nullptr, has_em = eps_map != nullptr;
    const bool upd_a = UPD && S.dab != nullptr;
    // an absent buffer is replaced by one read at the same index anyway, so
    // its loads hit lines already fetched (the links by A or abei, the cell
    // flags by the first byte of A's link, epsilon by the real part of the
    // previous gradient at the node)
    const R *lk = hb ? ab : (he ? abei : (const R *)psi);
    const R *abei_s = he ? abei : lk, *ab_s = hb ? ab : lk, *dab_s = upd_a ? S.dab : lk;
    const uint8_t *mt_s = has_mt ? mt : (const uint8_t *)lk;
    const int mt_st = has_mt ? 1 : (int)sizeof(R);
    const R *em_s = has_em ? eps_map : (const R *)jac_prev;
    const int em_st = has_em ? 1 : 2;
    const cplx<R> *dpsi_s = UPD ? S.dpsi : psi;
    const R *jAp_s = withA ? jacA_prev : lk;
    double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    // FP32 (LOCK): a workgroup takes four adjacent strips over one band of
    // rows; its waves march the same rows, kept in step by a barrier per
    // trip of the row loop, so the cache lines the strips share at their
    // edges are read by both while they are in the L2 (a CG iteration on a
    // 4097^2 grid: 0.676 against 0.698 ms). A wave past the last strip
    // repeats the last strip's loads and stores and sums nothing; every
    // wave of a workgroup has the same band, hence the same trip count.
    // FP64 keeps the wave-by-wave row-major partition, whose order of the
    // FP64 partial sums the parity tests are pinned to.
    constexpr bool LOCK = std::is_same<R, float>::value;
    auto body = [&](int s, int ja, int jb, bool live) {
        ColCtx X = col_ctx<1>(g.Nx, s);
        X.own = X.own && live;
        const int c = X.c;
        const uint32_t cc = X.cc, ca = X.ca;
        // Raw loads of marching step j: psi (and its update direction) at
        // row j+1 and the a links there; the b links, cell flag, epsilon and
        // previous gradients of row j. Indices are clamped into the grid.
        struct Step {
            cplx<R> p, dp, gp;
            R ae, ab, ad, be, bb, bd, ee, vpa, vpb;
            uint8_t mc;
        };
        auto kpsi = [&](int j) { return (long)g.Nx * min(max(j, 0), g.Ny - 1) + cc; };
        auto ka_of = [&](int j) { return ca + (long)g.Nxa * min(max(j, 0), g.Ny - 1); };
        auto kb_of = [&](int j) { return g.Na + cc + (long)g.Nxb * min(max(j, 0), g.Ny - 2); };
        auto kc_of = [&](int j) { return ca + (long)g.Nxc * min(max(j, 0), g.Ny - 2); };
        auto fetch = [&](int j) {
            Step t;
            const long kp = kpsi(j + 1), ka = ka_of(j + 1), kb = kb_of(j), n = kpsi(j);
            t.p = ld(psi, kp);
            t.dp = ld(dpsi_s, kp);
            t.ae = abei_s[ka];
            t.ab = ab_s[ka];
            t.ad = dab_s[ka];
            t.be = abei_s[kb];
            t.bb = ab_s[kb];
            t.bd = dab_s[kb];
            t.mc = mt_s[kc_of(j) * mt_st];
            t.ee = em_s[n * em_st];
            t.gp = ld(jac_prev, n);
            t.vpa = jAp_s[ka_of(j)];
            t.vpb = jAp_s[kb];
            return t;
        };
        auto upd_psi = [&](cplx<R> p, const cplx<R> &d) {
            if constexpr (UPD) {
                p.re = S.ap * d.re + p.re;
                p.im = S.ap * d.im + p.im;
            }
            return p;
        };
        auto upd_link = [&](R e, R b, R d) {
            LinkPair<R> v{he ? e : R(0), hb ? b : R(0)};
            if (upd_a) v.b = S.aa * d + v.b;
            return v;
        };
        auto pv = [&](const LinkPair<R> &v) {  // pv(abei, ab, k) of loaded values
            R p = R(0);
            if (he) p += v.e;
            if (hb) p += v.b;
            return p;
        };
        // window at the first row: psi at rows ja-1, ja; a at rows ja-1,
        // ja; b and the cell at row ja-1 and the up-link factors of row
        // ja-1 (the down link of row ja)
        cplx<R> Pd, Pc;
        LinkPair<R> Ad, Ac, Bd;
        bool md;
        {
            const long k0 = kpsi(ja - 1), k1 = kpsi(ja), a0 = ka_of(ja - 1), a1 = ka_of(ja), b0 = kb_of(ja - 1);
            Pd = upd_psi(ld(psi, k0), ld(dpsi_s, k0));
            Pc = upd_psi(ld(psi, k1), ld(dpsi_s, k1));
            Ad = upd_link(abei_s[a0], ab_s[a0], dab_s[a0]);
            Ac = upd_link(abei_s[a1], ab_s[a1], dab_s[a1]);
            Bd = upd_link(abei_s[b0], ab_s[b0], dab_s[b0]);
            md = has_mt ? mt[kc_of(ja - 1)] != 0 : true;
        }
        R sud = R(0), cud = R(1);
        if (ja > 0) {
            const bool mdl = shfl_up1((int)md) != 0;
            const bool lu_d = (c > 0 && mdl) || (c + 1 < g.Nx && md);  // cells (c-1, ja-1), (c, ja-1)
            if (X.col && lu_d) sincos_r(q.dy * pv(Bd), &sud, &cud);
        }
        auto consume = [&](const Step &t, int j) {
            const long n = (long)g.Nx * j + cc;
            const cplx<R> Pu = upd_psi(t.p, t.dp);
            const LinkPair<R> Au = upd_link(t.ae, t.ab, t.ad), Bc = upd_link(t.be, t.bb, t.bd);
            const bool mc = has_mt ? t.mc != 0 : true;
            GradIn<R> in;
            in.ee = has_em ? t.ee : eps;
            const cplx<R> gp = t.gp;
            const long ka = ca + (long)g.Nxa * j, kb = g.Na + cc + (long)g.Nxb * min(j, g.Ny - 2);
            const R vpa = withA ? t.vpa : R(0), vpb = withA ? t.vpb : R(0);
            // node values from the window and the neighbouring lanes
            in.p0 = Pc;
            in.pd = Pd;
            in.pu = Pu;
            in.pl.re = shfl_up1(Pc.re);
            in.pl.im = shfl_up1(Pc.im);
            in.pr.re = shfl_dn1(Pc.re);
            in.pr.im = shfl_dn1(Pc.im);
            in.c = cells_from(g, has_mt, c, j, shfl_up1((int)md) != 0, shfl_up1((int)mc) != 0, md, mc);
            const Cells cl = in.c;
            auto links = [&](bool e_part) {
                auto v = [&](const LinkPair<R> &x) { return e_part ? x.e : x.b; };
                Links10<R> L;
                L.a00 = v(Ac);
                L.am0 = shfl_up1(v(Ac));
                L.a0m = v(Ad);
                L.a0p = v(Au);
                L.amp = shfl_up1(v(Au));
                L.b00 = v(Bc);
                L.b0m = v(Bd);
                L.bm0 = shfl_up1(v(Bc));
                L.bp0 = shfl_dn1(v(Bc));
                L.bpm = shfl_dn1(v(Bd));
                return L;
            };
            in.E = he ? links(true) : Links10<R>{};
            in.B = hb ? links(false) : Links10<R>{};
            // link factors: the right and up links here, the left one from
            // the lane to the left, the down one from the row below
            in.sr = in.su = R(0);
            in.cr = in.cu = R(1);
            if (X.col && (cl.pm || cl.pp)) sincos_r(q.dx * pv(Ac), &in.sr, &in.cr);
            if (X.col && (cl.mp || cl.pp)) sincos_r(q.dy * pv(Bc), &in.su, &in.cu);
            in.sl = -shfl_up1(in.sr);
            in.cl = shfl_up1(in.cr);
            in.sd = -sud;
            in.cd = cud;
            if (X.own) {
                if constexpr (UPD) {
                    st(S.psi_out, n, Pc);
                    if (upd_a) {
                        if (c < g.Nxa) S.ab_out[ka] = Ac.b;
                        if (j < g.Nyb) S.ab_out[kb] = Bc.b;
                    }
                }
                R e, ga, gb;
                cplx<R> gj;
                bool ha, hbb;
                grad_eval<R, withA>(g, q, kappa2, H, he, hb, in, c, j, e, gj, ga, gb, ha, hbb, R(1));
                acc[0] += double(e);
                const cplx<R> gv = {dxdy * gj.re, dxdy * gj.im};
                st(jac, n, gv);
                acc[1] += double(gv.re * (gv.re - gp.re));
                acc[1] += double(gv.im * (gv.im - gp.im));
                acc[2] += double(gp.re * gp.re);
                acc[2] += double(gp.im * gp.im);
                if (withA) {
                    if (ha) {
                        const R v = s2 * ga;
                        jacA[ka] = v;
                        acc[3] += double(v * (v - vpa));
                        acc[4] += double(vpa * vpa);
                    }
                    if (hbb) {
                        const R v = s2 * gb;
                        jacA[kb] = v;
                        acc[3] += double(v * (v - vpb));
                        acc[4] += double(vpb * vpb);
                    }
                }
            }
            // roll the window
            Pd = Pc;
            Pc = Pu;
            Ad = Ac;
            Ac = Au;
            Bd = Bc;
            md = mc;
            sud = in.su;
            cud = in.cu;
        };
        // two batches of B steps in flight: the loads of one batch are
        // issued before the other is computed
        Step S0[B], S1[B];
#pragma unroll
        for (int k = 0; k < B; ++k) S0[k] = fetch(min(ja + k, jb - 1));
        for (int j = ja; j < jb; j += 2 * B) {
#pragma unroll
            for (int k = 0; k < B; ++k) S1[k] = fetch(min(j + B + k, jb - 1));
#pragma unroll
            for (int k = 0; k < B; ++k)
                if (j + k < jb) consume(S0[k], j + k);
#pragma unroll
            for (int k = 0; k < B; ++k) S0[k] = fetch(min(j + 2 * B + k, jb - 1));
#pragma unroll
            for (int k = 0; k < B; ++k)
                if (j + B + k < jb) consume(S1[k], j + B + k);
            if constexpr (LOCK) __syncthreads();
        }
    };
    if constexpr (LOCK) {
        const long nst = nstrips_of<1>(g.Nx), ng = (nst + 3) / 4, nwg = gridDim.x;
        const long nband = nwg >= ng ? nwg / ng : 1, U = nwg < nband * ng ? nwg : nband * ng;
        for (long item = xcd_block(); item < nband * ng; item += U) {  // workgroup-uniform
            const long band = item / ng;
            const int ja = (int)((band * g.Ny) / nband), jb = (int)(((band + 1) * g.Ny) / nband);
            const int s_raw = (int)((item % ng) * 4 + wave_id());
            if (ja < jb) body(s_raw < nst ? s_raw : (int)nst - 1, ja, jb, s_raw < nst);
        }
    } else {
        const long w = (long)xcd_block() * (blockDim.x / WAVE) + wave_id();
        if (w < W) for_rows_rm<1>(g.Nx, g.Ny, w, W, [&](int s, int ja, int jb) { body(s, ja, jb, true); });
    }
    block_store_partials<5>(acc, partials);
}

// The line-search coefficients of the new direction d = beta d_old - g,
// which the pass forms once per node where it reads it and writes to
// U.dout / U.dAout (glx_cg_direction_step; k_coef_dir's results, node by
// node). Partials as k_coef_dir leaves them: the L sums per workgroup, then
// (L = 17) each workgroup's largest link phase step.
template <typename R> struct LinkTriple {
    R e, b, d;  // abei, ab, the new A direction
};

template <typename R, int L>
__global__ void __launch_bounds__(256) k_coef_dir_rows(Geo g, R kappa2, R eps, const R *eps_map, R H, const uint8_t *mt,
                                                       const cplx<R> *psi, const R *abei, const R *ab, DirUpdate<R> U,
                                                       const double *beta, double *partials, long W) {
    constexpr int B = CGRows<R>::B;
    const GeoR<R> q(g);
    U.bp = R(beta[0]);
    U.ba = R(beta[1]);
    const bool he = abei != nullptr, hb = ab != nullptr, has_mt = mt != nullptr, has_em = eps_map != nullptr;
    // absent buffers replaced as in k_cg_grad_rows (epsilon by the real part
    // of the gradient at the node)
    const R *lk = hb ? ab : (he ? abei : (const R *)psi);
    const R *abei_s = he ? abei : lk, *ab_s = hb ? ab : lk;
    const R *dAold_s = L == 17 ? U.dAold : lk, *jacA_s = L == 17 ? U.jacA : lk;
    const uint8_t *mt_s = has_mt ? mt : (const uint8_t *)lk;
    const int mt_st = has_mt ? 1 : (int)sizeof(R);
    const R *em_s = has_em ? eps_map : (const R *)U.jac;
    const int em_st = has_em ? 1 : 2;
    double acc[L];
#pragma unroll
    for (int k = 0; k < L; ++k) acc[k] = 0.0;
    // the largest link phase step in the working precision: a maximum is
    // exact in either, and widening it at the end gives the FP64 value
    R th = R(0);
    const long w = (long)xcd_block() * (blockDim.x / WAVE) + wave_id();
    if (w < W) for_rows_rm<1>(g.Nx, g.Ny, w, W, [&](int s, int ja, int jb) {
        const ColCtx X = col_ctx<1>(g.Nx, s);
        const int c = X.c;
        const uint32_t cc = X.cc, ca = X.ca;
        // Raw loads of marching step j (psi, the old direction and the
        // gradient at row j+1, the a links there, the b links, cell flag and
        // epsilon of row j); the new direction beta d_old - g is formed where
        // they are consumed.
        struct Lk {
            R e, b, o, g;  // abei, ab, old A direction, A gradient
        };
        struct Step {
            cplx<R> p, o, gr;
            Lk a, b;
            R eps;
            uint8_t mc;
        };
        auto row = [&](int j) { return (long)g.Nx * min(max(j, 0), g.Ny - 1) + cc; };
        auto ka_of = [&](int j) { return ca + (long)g.Nxa * min(max(j, 0), g.Ny - 1); };
        auto kb_of = [&](int j) { return g.Na + cc + (long)g.Nxb * min(max(j, 0), g.Ny - 2); };
        auto kc_of = [&](int j) { return ca + (long)g.Nxc * min(max(j, 0), g.Ny - 2); };
        auto link = [&](long k) { return Lk{abei_s[k], ab_s[k], dAold_s[k], jacA_s[k]}; };
        auto dlink = [&](const Lk &v) {
            return LinkTriple<R>{he ? v.e : R(0), hb ? v.b : R(0), L == 17 ? U.ba * v.o - v.g : R(0)};
        };
        auto dnode = [&](const cplx<R> &o, const cplx<R> &gk) {
            return cplx<R>{U.bp * o.re - gk.re, U.bp * o.im - gk.im};
        };
        // each load of a step at a row pointer (uniform) plus the lane's
        // byte offset (loop-invariant, 32-bit: grids hold at most 2^30 nodes)
        const uint32_t o_n = cc * (uint32_t)sizeof(cplx<R>), o_a = ca * (uint32_t)sizeof(R),
                       o_b = cc * (uint32_t)sizeof(R), o_m = ca * (uint32_t)mt_st,
                       o_e = cc * (uint32_t)(em_st * sizeof(R));
        auto link_at = [&](size_t rk, uint32_t o) {
            return Lk{at(abei_s + rk, o), at(ab_s + rk, o), at(dAold_s + rk, o), at(jacA_s + rk, o)};
        };
        auto fetch = [&](int j) {
            Step t;
            const int jp = min(max(j + 1, 0), g.Ny - 1), jr = min(max(j, 0), g.Ny - 1), jq = min(max(j, 0), g.Ny - 2);
            const size_t rn = (size_t)g.Nx * jp;
            t.p = ld_at(psi + rn, o_n);
            t.o = ld_at(U.dold + rn, o_n);
            t.gr = ld_at(U.jac + rn, o_n);
            t.a = link_at((size_t)g.Nxa * jp, o_a);
            t.b = link_at((size_t)g.Na + (size_t)g.Nxb * jq, o_b);
            t.mc = at(mt_s + (size_t)g.Nxc * jq * mt_st, o_m);
            t.eps = at(em_s + (size_t)g.Nx * jr * em_st, o_e);
            return t;
        };
        // window: psi, direction and the a links at row ja; the cell below
        cplx<R> Pc = ld(psi, row(ja)), Dc = dnode(ld(U.dold, row(ja)), ld(U.jac, row(ja)));
        LinkTriple<R> Ac = dlink(link(ka_of(ja)));
        bool md = has_mt ? mt[kc_of(ja - 1)] != 0 : true;
        // FP32: a lane's coefficients of the 2B rows of one trip are summed
        // in FP32 (at most four terms each) and the sums added to the FP64
        // accumulators once per trip; FP64 adds each node's terms directly
        constexpr bool PART = std::is_same<R, float>::value;
        R part[L];
#pragma unroll
        for (int k = 0; k < L; ++k) part[k] = R(0);
        auto consume = [&](const Step &t, int j) {
            const long n = (long)g.Nx * j + cc;
            const cplx<R> Pu = t.p, Du = dnode(t.o, t.gr);
            const LinkTriple<R> Au = dlink(t.a), Bc = dlink(t.b);
            const bool mc = has_mt ? t.mc != 0 : true;
            CoefIn<R> in;
            in.eps = has_em ? t.eps : eps;
            in.p0 = Pc;
            in.p_u = Pu;
            in.p_r.re = shfl_dn1(Pc.re);
            in.p_r.im = shfl_dn1(Pc.im);
            in.d0 = Dc;
            in.d_u = Du;
            in.d_r.re = shfl_dn1(Dc.re);
            in.d_r.im = shfl_dn1(Dc.im);
            in.c = cells_from(g, has_mt, c, j, shfl_up1((int)md) != 0, shfl_up1((int)mc) != 0, md, mc);
            in.E = Links4<R>{Ac.e, Au.e, Bc.e, shfl_dn1(Bc.e)};
            in.B = Links4<R>{Ac.b, Au.b, Bc.b, shfl_dn1(Bc.b)};
            in.D = Links4<R>{};
            if (L == 17) in.D = Links4<R>{Ac.d, Au.d, Bc.d, shfl_dn1(Bc.d)};
            if (X.own) {
                st(U.dout, n, Dc);
                if (L == 17) {
                    if (c < g.Nxa) U.dAout[ca + (long)g.Nxa * j] = Ac.d;
                    if (j < g.Nyb) U.dAout[g.Na + cc + (long)g.Nxb * j] = Bc.d;
                    th = fmax(th, fmax(fabs(q.dx * Ac.d), fabs(q.dy * Bc.d)));
                }
                R C[L];
#pragma unroll
                for (int k = 0; k < L; ++k) C[k] = R(0);
                coef_eval<R, L>(q, kappa2, H, he, hb, in, c, j, c < g.Nxc && j < g.Nyc, C);
#pragma unroll
                for (int k = 0; k < L; ++k) {
                    if constexpr (PART) part[k] += C[k];
                    else acc[k] += double(C[k]);
                }
            }
            // values that outlive their row buffer get registers of their
            // own (see own_copy), so the next batch can load into it
            Pc = own_copy(Pu);
            Dc = Du;
            Ac = LinkTriple<R>{own_copy(Au.e), own_copy(Au.b), Au.d};
            md = mc;
        };
        Step S0[B], S1[B];
#pragma unroll
        for (int k = 0; k < B; ++k) S0[k] = fetch(min(ja + k, jb - 1));
        for (int j = ja; j < jb; j += 2 * B) {
#pragma unroll
            for (int k = 0; k < B; ++k) S1[k] = fetch(min(j + B + k, jb - 1));
#pragma unroll
            for (int k = 0; k < B; ++k)
                if (j + k < jb) consume(S0[k], j + k);
#pragma unroll
            for (int k = 0; k < B; ++k) S0[k] = fetch(min(j + 2 * B + k, jb - 1));
#pragma unroll
            for (int k = 0; k < B; ++k)
                if (j + B + k < jb) consume(S1[k], j + B + k);
            if constexpr (PART) {
#pragma unroll
                for (int k = 0; k < L; ++k) {
                    acc[k] += double(part[k]);
                    part[k] = R(0);
                }
            }
        }
    });
    block_store_partials<L>(acc, partials);
    if (L == 17) block_store_phase_max(double(th), partials, L);
}

}  // namespace glx
