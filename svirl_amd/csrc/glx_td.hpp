// TDGL Jacobi sweeps for psi and A on gfx950.
//
// Layout of the work: the grid is cut into vertical strips of 62 columns
// (58 or 60 for passes that fuse three or two sweeps); one wave marches down
// a range of rows of one strip. Lanes 0 and 63 of a wave hold the columns
// left and right of the strip (halos), lanes 1..62 own one column each.
// Each lane keeps a three-row window of its column in registers, so every
// field value is read from HBM once per pass; horizontal neighbours come
// from adjacent lanes by cross-lane shuffles. Each lane evaluates sin/cos
// once per link it carries (2 per node per pass) and hands the values to
// its right neighbour.
//
// Arithmetic follows svirl/cuda/td.h term by term (same operand order, no
// reassociation); FP64 instantiations are compiled without FMA contraction.
#pragma once

#include <type_traits>

#include "glx_device.hpp"

namespace glx {

template <typename R> struct PsiSweepArgs {
    int Nx, Ny;
    R dx, dy, idx2, idy2;
    R dt, eps;
    const R *eps_map;        // per node or NULL
    const uint8_t *mt;       // per cell or NULL
    const R *ab;             // vector potential (a block then b block)
    const cplx<R> *psi;      // psi^j
    cplx<R> *psi_next;       // psi^{j+1}
    const cplx<R> *rhs;      // right-hand side read
    cplx<R> *rhs_out;        // Langevin: noisy rhs written here at jstep 0 (else NULL)
    R langevin_c;
    uint32_t jstep, rand_t;
    double stop_eps;
    // gauge-shifted storage (FP32-mixed run): the buffers hold A - A_g with
    // A_g the symmetric gauge of field gH; link phases add A_g back in double
    double gH, gdx, gdy, gx0, gy0;
};

template <typename R> struct ASweepArgs {
    int Nx, Ny;
    R dx, dy, idx, idy, idx2, idy2, idxy;
    R dt_rho, dt_rho_kappa2;
    R rh_a, rh_b;            // 2 kappa^2 H idy, 2 kappa^2 H idx
    R den_a, den_b;          // 1 + 2 dt rho kappa^2 idy2 (idx2)
    R iden_a, iden_b;        // 1/den_a, 1/den_b (FP32 multiplies by them instead of dividing)
    const uint8_t *mt;
    const cplx<R> *psi;
    const R *src;            // supercurrent source (may alias ab_next)
    const R *rhs;            // right-hand side read
    R *rhs_out;              // Langevin: noisy rhs written here at jstep 0 (else NULL)
    const R *ab;             // A^j
    R *ab_next;              // A^{j+1} (a fused pass: its last sweep's result)
    R *ab_mid;               // three-sweep pass: A^{j+2} (the next sweep's source)
    R langevin_c;
    uint32_t jstep, rand_t;
    double stop_eps;
    double gH, gdx, gdy, gx0, gy0;  // as in PsiSweepArgs
};

// Symmetric gauge of field H (the one svirl builds, params.py:159-167) on
// link a(c, j) and b(c, j): a_g = -H/2 (y_j - Ly/2), b_g = H/2 (x_c - Lx/2).
__device__ __forceinline__ double gauge_a(double H, double dy, double y0, int j) { return -0.5 * H * (double(j) * dy - y0); }
__device__ __forceinline__ double gauge_b(double H, double dx, double x0, int c) { return 0.5 * H * (double(c) * dx - x0); }

// Link phase d*A of a stored value v: plain, or (gauge-shifted storage) with
// the gauge's phase dg = d*A_g added back, dg being evaluated in double once
// per row or column and rounded to the working precision.
template <typename R, bool GAUGE>
__device__ __forceinline__ R link_phase(R d, R v, R dg) {
    if (GAUGE) return fma(d, v, dg);
    return d * v;
}
template <typename R, typename Args> __device__ __forceinline__ R gauge_phase_a(const Args &A, int j) {
    return R(double(A.dx) * gauge_a(A.gH, A.gdy, A.gy0, j));
}
template <typename R, typename Args> __device__ __forceinline__ R gauge_phase_b(const Args &A, int c) {
    return R(double(A.dy) * gauge_b(A.gH, A.gdx, A.gx0, c));
}

// Rows are loaded in batches of Pipe<R>::ROWS (single sweeps), ROWS2 (fused
// psi passes) or AROWS2 / AROWS3 (fused A passes of two / three sweeps): the
// loads of batch k+1 are issued before batch k is computed, and the two
// batches live in separately named registers (the loop body is unrolled
// over the two), so no loaded value is copied before it is used and 1 to 2
// batches of loads per wave stay in flight during the arithmetic. The
// persistent solves run two waves per SIMD, which leaves registers for
// two-row batches (measured at 4097^2 on two-sweep psi passes: FP32 93 ->
// 85 us per sweep, FP64 162 -> 155; deeper batches and FP32 A passes gain
// nothing). Loads beyond the wave's last row are clamped to that row (cache
// hits, never used). The batched kernel runs its solves with SH (shallow)
// batches, which keep its register count within two waves per SIMD.
template <typename R, bool SH = false> struct Pipe { static constexpr int ROWS = 2, ROWS2 = 2, AROWS2 = 1, AROWS3 = 1; };
template <> struct Pipe<double, false> { static constexpr int ROWS = 2, ROWS2 = 2, AROWS2 = 2, AROWS3 = 1; };
template <> struct Pipe<float, true> { static constexpr int ROWS = 2, ROWS2 = 1, AROWS2 = 1, AROWS3 = 1; };
template <> struct Pipe<double, true> { static constexpr int ROWS = 1, ROWS2 = 1, AROWS2 = 1, AROWS3 = 1; };

// Sweeps per fused psi pass in the persistent single-instance solve (FP32:
// four, psi_rowsK<4> with each row's loads issued three rows ahead, 227
// VGPRs, two workgroups per CU; FP64 keeps three, psi_rows3, whose
// register footprint under psi_rowsK<3> would leave one wave per SIMD).
template <typename R> struct Fuse { static constexpr int PSI = 3; };
template <> struct Fuse<float> { static constexpr int PSI = 4; };

// Per-lane column context of a strip: c may lie outside the grid (halo lanes
// at the edges); cc and ca are c clamped into the node and link ranges, so
// that every load of a row is issued unconditionally from a valid address
// (straight-line loads let the compiler count outstanding loads exactly and
// keep the next batch in flight); values that do not exist are masked at
// their use instead.
struct ColCtx {
    int c;
    uint32_t cc, ca;
    bool col, own, hasa;
};

// Strips of WAVE - 2*HALO columns; lanes [HALO, WAVE - HALO) own a column.
template <int HALO = 1>
__device__ __forceinline__ ColCtx col_ctx(int Nx, int s) {
    const int lane = threadIdx.x & (WAVE - 1);
    ColCtx X;
    X.c = s * (WAVE - 2 * HALO) - HALO + lane;
    X.col = (X.c >= 0) && (X.c < Nx);
    X.own = X.col && lane >= HALO && lane < WAVE - HALO;
    X.hasa = X.col && X.c < Nx - 1;
    X.cc = (uint32_t)min(max(X.c, 0), Nx - 1);
    X.ca = min(X.cc, (uint32_t)(Nx - 2));
    return X;
}

// Raw loads of one row for the psi sweep: psi at row j+1, a and b at row j,
// the mask byte of cell (c, j), the right-hand side and epsilon at row j.
template <typename R> struct PsiRow {
    cplx<R> pp, rhs;
    R av, bv, e;
    int cell;
};

template <typename R>
__device__ __forceinline__ PsiRow<R> psi_load_row(const PsiSweepArgs<R> &A, const ColCtx &X, const uint8_t *mtp,
                                                  const R *emp, int j) {
    const uint32_t Nx = A.Nx, Ny = A.Ny, Nxa = Nx - 1;
    const uint32_t jr = min((uint32_t)j, Ny - 1), jp = min((uint32_t)j + 1, Ny - 1), jq = min((uint32_t)j, Ny - 2);
    const uint32_t oc = X.cc * (uint32_t)sizeof(R), oa = X.ca * (uint32_t)sizeof(R);
    PsiRow<R> L;
    L.pp = ld_at(A.psi + (size_t)Nx * jp, 2 * oc);
    L.av = at(A.ab + (size_t)Nxa * jr, oa);
    L.bv = at(A.ab + (size_t)Nxa * Ny + (size_t)Nx * jq, oc);
    L.cell = at(mtp + (size_t)Nxa * jq, X.ca);
    L.rhs = ld_at(A.rhs + (size_t)Nx * jr, 2 * oc);
    L.e = at(emp + (size_t)Nx * jr, oc);
    return L;
}

// Backward-Euler Jacobi update of one psi node (svirl/cuda/td.h:4-133),
// in two parts. psi_den: the diagonal 1 + dt (|rhs|^2 - eps + sum of the
// link weights / h^2) from the four adjacent cells (mm = (c-1, j-1),
// mp = (c-1, j), pmc = (c, j-1), ppc = (c, j)); FP32 returns its reciprocal
// (hardware estimate + one Newton step, within 1 ulp, one reciprocal for
// both components instead of two IEEE divisions), FP64 the diagonal itself.
// It depends on the sweep only through rhs, so the sweeps of a fused pass
// share it.
template <typename R>
__device__ __forceinline__ R psi_den(const PsiSweepArgs<R> &A, const cplx<R> &rhs, R e, bool mm, bool mp, bool pmc,
                                     bool ppc) {
    const R rm0 = (mm || mp) ? R(1) : R(0);
    const R rp0 = (pmc || ppc) ? R(1) : R(0);
    const R r0m = (mm || pmc) ? R(1) : R(0);
    const R r0p = (mp || ppc) ? R(1) : R(0);
    R q = rhs.re * rhs.re + rhs.im * rhs.im - e;
    q = q + (A.idx2 * (rm0 + rp0) + A.idy2 * (r0m + r0p));
    const R diag = R(1) + A.dt * q;
    if constexpr (std::is_same<R, float>::value) {
        float rd = __builtin_amdgcn_rcpf(diag);
        return fmaf(rd, fmaf(-diag, rd, 1.0f), rd);
    } else {
        return diag;
    }
}

// psi_apply: the off-diagonal part and the division, from the neighbours pl,
// pr, pm, pp (left, right, down, up) and the link factors (c, s) = (cos, sin)
// of dx a_left, dx a, dy b_down, dy b. A link that touches no material cell
// carries the factors (0, 0) (mask_link), which makes its term 0 as the
// reference's link weight 0 does.
template <typename R>
__device__ __forceinline__ cplx<R> psi_apply(const PsiSweepArgs<R> &A, const cplx<R> &rhs, R den, R pl_re, R pl_im,
                                             R pr_re, R pr_im, const cplx<R> &pm, const cplx<R> &pp, R cal, R sal,
                                             R ca, R sa, R cbm, R sbm, R cb, R sb) {
    // U(-dx a_left) psi_left, U(dx a) psi_right, U(-dy b_down) psi_down, U(dy b) psi_up
    const R m0r = cal * pl_re - sal * pl_im, m0i = cal * pl_im + sal * pl_re;
    const R p0r = ca * pr_re + sa * pr_im, p0i = ca * pr_im - sa * pr_re;
    const R q0mr = cbm * pm.re - sbm * pm.im, q0mi = cbm * pm.im + sbm * pm.re;
    const R q0pr = cb * pp.re + sb * pp.im, q0pi = cb * pp.im - sb * pp.re;
    const R nre = rhs.re + A.dt * (A.idx2 * (m0r + p0r) + A.idy2 * (q0mr + q0pr));
    const R nim = rhs.im + A.dt * (A.idx2 * (m0i + p0i) + A.idy2 * (q0mi + q0pi));
    if constexpr (std::is_same<R, float>::value) return {nre * den, nim * den};
    else return {nre / den, nim / den};
}

// psi_apply for FP32 passes with two products folded in per row: the link
// factors arrive scaled by dt/dx^2 (a links) and dt/dy^2 (b links), and rd
// is rhs * den, so each component is one sum of the four link terms and
// one fma (the same quantity as psi_apply's, rounded differently).
template <typename R>
__device__ __forceinline__ cplx<R> psi_apply_folded(const cplx<R> &rd, R den, R pl_re, R pl_im, R pr_re, R pr_im,
                                                    const cplx<R> &pm, const cplx<R> &pp, R cal, R sal, R ca, R sa,
                                                    R cbm, R sbm, R cb, R sb) {
    const R m0r = cal * pl_re - sal * pl_im, m0i = cal * pl_im + sal * pl_re;
    const R p0r = ca * pr_re + sa * pr_im, p0i = ca * pr_im - sa * pr_re;
    const R q0mr = cbm * pm.re - sbm * pm.im, q0mi = cbm * pm.im + sbm * pm.re;
    const R q0pr = cb * pp.re + sb * pp.im, q0pi = cb * pp.im - sb * pp.re;
    const R sre = (m0r + p0r) + (q0mr + q0pr), sim = (m0i + p0i) + (q0mi + q0pi);
    return {fma(sre, den, rd.re), fma(sim, den, rd.im)};
}

// Link factors of a link that touches no material cell become (0, 0).
template <typename R> __device__ __forceinline__ void mask_link(bool on, R &c, R &s) {
    c = on ? c : R(0);
    s = on ? s : R(0);
}

// The reference's Langevin term added to the right-hand side at node n.
template <typename R>
__device__ __forceinline__ void psi_noise(const PsiSweepArgs<R> &A, uint32_t n, cplx<R> &rhs) {
    const R n1 = R(double(noise_1<R>(n, A.rand_t)) - 0.5);
    const R n2 = R(double(noise_2<R>(n, A.rand_t)) - 0.5);
    rhs.re = rhs.re + A.langevin_c * n1;
    rhs.im = rhs.im + A.langevin_c * n2;
}

// One psi sweep over rows [ja, jb) of strip s. Returns the lane's largest
// |delta| (max of real and imaginary parts); the stop rule's scaled integer
// residual is monotone in it, so it is applied once per wave by the caller.
template <typename R, bool GAUGE, bool SH = false>
__device__ __forceinline__ R psi_rows(const PsiSweepArgs<R> &A, int s, int ja, int jb) {
    constexpr int B = Pipe<R, SH>::ROWS;
    const ColCtx X = col_ctx<1>(A.Nx, s);
    const int c = X.c;
    const bool own = X.own;
    const int Nx = A.Nx, Ny = A.Ny, Nxa = Nx - 1;
    const R *b = A.ab + (long)Nxa * Ny;
    const bool noisy = A.rhs_out != nullptr;
    const bool has_mt = A.mt != nullptr;
    const uint8_t *mtp = has_mt ? A.mt : reinterpret_cast<const uint8_t *>(A.ab);     // any readable bytes
    const R *emp = A.eps_map ? A.eps_map : reinterpret_cast<const R *>(A.rhs);         // when absent
    const R gb = GAUGE ? gauge_phase_b<R>(A, c) : R(0);
    R rr = R(0);

    // window: psi at rows j-1, j of this column; U of b(c, j-1); cell (c, j-1)
    cplx<R> pm = {R(0), R(0)}, pc = {R(0), R(0)};
    R cbm = R(1), sbm = R(0);
    int cellm = 0;
    if (X.col) {
        pc = ld(A.psi, c + (long)Nx * ja);
        if (ja > 0) {
            pm = ld(A.psi, c + (long)Nx * (ja - 1));
            sincos_r(link_phase<R, GAUGE>(A.dy, b[c + (long)Nx * (ja - 1)], gb), &sbm, &cbm);
            if (c < Nxa) cellm = has_mt ? int(A.mt[c + (long)Nxa * (ja - 1)]) : 1;
        }
    }
    mask_link(shfl_up1(cellm) != 0 || cellm != 0, cbm, sbm);  // b(c, ja-1): cells (c-1, ja-1), (c, ja-1)

    auto row = [&](const PsiRow<R> &cu, int j) {
        const long rowoff = (long)Nx * j;
        const cplx<R> pp = (X.col && j + 1 < Ny) ? cu.pp : cplx<R>{R(0), R(0)};
        const R av = X.hasa ? cu.av : R(0);
        const R bv = (X.col && j < Ny - 1) ? cu.bv : R(0);
        const int cellc = (X.hasa && j < Ny - 1) ? (has_mt ? cu.cell : 1) : 0;
        cplx<R> rhs = cu.rhs;
        const R e = A.eps_map ? cu.e : A.eps;
        R sa, ca, sb, cb;
        sincos_r(link_phase<R, GAUGE>(A.dx, av, GAUGE ? gauge_phase_a<R>(A, j) : R(0)), &sa, &ca);
        sincos_r(link_phase<R, GAUGE>(A.dy, bv, gb), &sb, &cb);
        const bool cl_m = shfl_up1(cellm) != 0;  // cell (c-1, j-1)
        const bool cl_c = shfl_up1(cellc) != 0;  // cell (c-1, j)
        mask_link(cellm != 0 || cellc != 0, ca, sa);  // a(c, j): cells (c, j-1), (c, j)
        mask_link(cl_c || cellc != 0, cb, sb);        // b(c, j): cells (c-1, j), (c, j)

        // neighbours from adjacent lanes
        const R pl_re = shfl_up1(pc.re), pl_im = shfl_up1(pc.im);
        const R pr_re = shfl_dn1(pc.re), pr_im = shfl_dn1(pc.im);
        const R cal = shfl_up1(ca), sal = shfl_up1(sa);

        if (own) {
            const bool mm = cl_m, mp = cl_c, pmc = cellm != 0, ppc = cellc != 0;
            const bool inside = mm || mp || pmc || ppc;
            cplx<R> nxt = {R(0), R(0)};
            if (inside) {
                if (noisy && A.jstep == 0 && A.langevin_c > R(1.0e-32)) psi_noise(A, (uint32_t)(c + rowoff), rhs);
                nxt = psi_apply(A, rhs, psi_den(A, rhs, e, mm, mp, pmc, ppc), pl_re, pl_im, pr_re, pr_im, pm, pp, cal,
                                sal, ca, sa, cbm, sbm, cb, sb);
            }
            if (noisy && A.jstep == 0) st(A.rhs_out, c + rowoff, rhs);
            st_at(A.psi_next + rowoff, 2 * X.cc * (uint32_t)sizeof(R), nxt);
            rr = fmax(rr, fmax(fabs(nxt.re - pc.re), fabs(nxt.im - pc.im)));
        }
        // roll the window
        pm = pc;
        pc = pp;
        cbm = cb;
        sbm = sb;
        cellm = cellc;
    };

    PsiRow<R> P0[B], P1[B];
#pragma unroll
    for (int k = 0; k < B; ++k) P0[k] = psi_load_row(A, X, mtp, emp, min(ja + k, jb - 1));
    for (int j = ja; j < jb; j += 2 * B) {
#pragma unroll
        for (int k = 0; k < B; ++k) P1[k] = psi_load_row(A, X, mtp, emp, min(j + B + k, jb - 1));
#pragma unroll
        for (int k = 0; k < B; ++k)
            if (j + k < jb) row(P0[k], j + k);
#pragma unroll
        for (int k = 0; k < B; ++k) P0[k] = psi_load_row(A, X, mtp, emp, min(j + 2 * B + k, jb - 1));
#pragma unroll
        for (int k = 0; k < B; ++k)
            if (j + B + k < jb) row(P1[k], j + B + k);
    }
    return rr;
}

// Two consecutive psi sweeps j and j+1 in one pass over rows [ja, jb) of
// strip s (strips of WAVE - 4 columns, two halo lanes a side): the first
// sweep runs one row ahead of the second and on lanes 1..62, so the second
// finds its three rows and both lateral neighbours of psi^{j+1} in
// registers. Link factors and cells are computed once for both sweeps;
// psi^{j+1} never leaves registers. Writes psi^{j+2} (and, at jstep 0, the
// Langevin right-hand side); returns the lane's largest |delta| of each sweep
// in r1 and r2. The arithmetic of each node update is that of psi_rows.
template <typename R, bool GAUGE, bool SH = false>
__device__ __forceinline__ void psi_rows2(const PsiSweepArgs<R> &A, int s, int ja, int jb, R &r1, R &r2) {
    constexpr int B = Pipe<R, SH>::ROWS2;
    const ColCtx X = col_ctx<2>(A.Nx, s);
    const int lane = threadIdx.x & (WAVE - 1);
    const int c = X.c;
    const bool own = X.own;                                  // lanes 2..61: outputs
    const bool mid = X.col && lane >= 1 && lane < WAVE - 1;  // lanes 1..62: first sweep
    const int Nx = A.Nx, Ny = A.Ny, Nxa = Nx - 1;
    const R *b = A.ab + (long)Nxa * Ny;
    const bool noisy = A.rhs_out != nullptr;
    const bool noise_on = noisy && A.jstep == 0 && A.langevin_c > R(1.0e-32);
    const bool has_mt = A.mt != nullptr;
    const uint8_t *mtp = has_mt ? A.mt : reinterpret_cast<const uint8_t *>(A.ab);
    const R *emp = A.eps_map ? A.eps_map : reinterpret_cast<const R *>(A.rhs);
    const R gb = GAUGE ? gauge_phase_b<R>(A, c) : R(0);
    const cplx<R> zero = {R(0), R(0)};
    const int t0 = ja > 0 ? ja - 1 : 0;

    // first sweep window: psi^j at rows t-1, t; U of b(t-1); cell (c, t-1)
    cplx<R> pm = zero, pc = zero;
    R cb1 = R(1), sb1 = R(0);
    int cell1 = 0;
    if (X.col) {
        pc = ld(A.psi, c + (long)Nx * t0);
        if (t0 > 0) {
            pm = ld(A.psi, c + (long)Nx * (t0 - 1));
            sincos_r(link_phase<R, GAUGE>(A.dy, b[c + (long)Nx * (t0 - 1)], gb), &sb1, &cb1);
            if (c < Nxa) cell1 = has_mt ? int(A.mt[c + (long)Nxa * (t0 - 1)]) : 1;
        }
    }
    mask_link(shfl_up1(cell1) != 0 || cell1 != 0, cb1, sb1);  // b(c, t0-1)
    // second sweep window (row u = t - 1): psi^{j+1} at u-1, u; U of a(u),
    // b(u-1); rhs, diagonal and material test of node (c, u)
    cplx<R> qm = zero, qc = zero, rhs_u = zero;
    R ca2 = R(0), sa2 = R(0), cal2 = R(0), sal2 = R(0), cb2 = R(0), sb2 = R(0), den_u = R(1);
    bool in_u = false;

    auto row = [&](const PsiRow<R> &cu, int t) {
        const long rowoff = (long)Nx * t;
        const bool real = t < Ny;  // t == Ny: past the last row (second sweep only)
        const cplx<R> pp = (X.col && t + 1 < Ny) ? cu.pp : zero;
        const R av = X.hasa ? cu.av : R(0);
        const R bv = (X.col && t < Ny - 1) ? cu.bv : R(0);
        const int cellc = (X.hasa && t < Ny - 1) ? (has_mt ? cu.cell : 1) : 0;
        cplx<R> rhs = cu.rhs;
        const R e = A.eps_map ? cu.e : A.eps;
        R sa, ca, sb, cb;
        sincos_r(link_phase<R, GAUGE>(A.dx, av, GAUGE ? gauge_phase_a<R>(A, t) : R(0)), &sa, &ca);
        sincos_r(link_phase<R, GAUGE>(A.dy, bv, gb), &sb, &cb);
        const bool mm = shfl_up1(cell1) != 0;  // cell (c-1, t-1)
        const bool mp = shfl_up1(cellc) != 0;  // cell (c-1, t)
        const bool pmc = cell1 != 0, ppc = cellc != 0;
        mask_link(pmc || ppc, ca, sa);  // a(c, t): cells (c, t-1), (c, t)
        mask_link(mp || ppc, cb, sb);   // b(c, t): cells (c-1, t), (c, t)
        const R cal = shfl_up1(ca), sal = shfl_up1(sa);
        const bool inside = mm || mp || pmc || ppc;

        // first sweep at row t -> psi^{j+1}(t)
        cplx<R> q = zero;
        R den = R(1);
        {
            const R pl_re = shfl_up1(pc.re), pl_im = shfl_up1(pc.im);
            const R pr_re = shfl_dn1(pc.re), pr_im = shfl_dn1(pc.im);
            if (mid && real && inside) {
                if (noise_on) psi_noise(A, (uint32_t)(c + rowoff), rhs);
                den = psi_den(A, rhs, e, mm, mp, pmc, ppc);
                q = psi_apply(A, rhs, den, pl_re, pl_im, pr_re, pr_im, pm, pp, cal, sal, ca, sa, cb1, sb1, cb, sb);
            }
            if (own && real && t >= ja && t < jb) {
                if (noisy && A.jstep == 0) st(A.rhs_out, c + rowoff, rhs);
                r1 = fmax(r1, fmax(fabs(q.re - pc.re), fabs(q.im - pc.im)));
            }
        }
        // second sweep at row u = t - 1 -> psi^{j+2}(u)
        const int u = t - 1;
        if (u >= ja) {
            const R pl_re = shfl_up1(qc.re), pl_im = shfl_up1(qc.im);
            const R pr_re = shfl_dn1(qc.re), pr_im = shfl_dn1(qc.im);
            if (own) {
                cplx<R> nxt = zero;
                if (in_u)
                    nxt = psi_apply(A, rhs_u, den_u, pl_re, pl_im, pr_re, pr_im, qm, q, cal2, sal2, ca2, sa2, cb2, sb2,
                                    cb1, sb1);
                st_at(A.psi_next + (size_t)Nx * u, 2 * X.cc * (uint32_t)sizeof(R), nxt);
                r2 = fmax(r2, fmax(fabs(nxt.re - qc.re), fabs(nxt.im - qc.im)));
            }
        }
        // roll both windows
        qm = qc;
        qc = q;
        cb2 = cb1;
        sb2 = sb1;
        ca2 = ca;
        sa2 = sa;
        cal2 = cal;
        sal2 = sal;
        in_u = inside;
        den_u = den;
        rhs_u = rhs;
        pm = pc;
        pc = pp;
        cb1 = cb;
        sb1 = sb;
        cell1 = cellc;
    };

    const int t1 = jb + 1;  // rows t in [t0, t1)
    PsiRow<R> P0[B], P1[B];
#pragma unroll
    for (int k = 0; k < B; ++k) P0[k] = psi_load_row(A, X, mtp, emp, min(t0 + k, t1 - 1));
    for (int t = t0; t < t1; t += 2 * B) {
#pragma unroll
        for (int k = 0; k < B; ++k) P1[k] = psi_load_row(A, X, mtp, emp, min(t + B + k, t1 - 1));
#pragma unroll
        for (int k = 0; k < B; ++k)
            if (t + k < t1) row(P0[k], t + k);
#pragma unroll
        for (int k = 0; k < B; ++k) P0[k] = psi_load_row(A, X, mtp, emp, min(t + 2 * B + k, t1 - 1));
#pragma unroll
        for (int k = 0; k < B; ++k)
            if (t + B + k < t1) row(P1[k], t + B + k);
    }
}

// Three consecutive psi sweeps j, j+1, j+2 in one pass over rows [ja, jb) of
// strip s (strips of WAVE - 6 columns, three halo lanes a side), built like
// psi_rows2 with one more stage: sweep j runs at row t on lanes 1..62,
// sweep j+1 at row t-1 on lanes 2..61 and sweep j+2 at row t-2 on lanes
// 3..60, each from the previous stage's rows kept in registers. Link
// factors, cells and node diagonals are computed once for the three sweeps.
// Writes psi^{j+3} (and, at jstep 0, the Langevin right-hand side); r1, r2
// and r3 get the lane's largest |delta| of each sweep.
template <typename R, bool GAUGE, bool SH = false>
__device__ __forceinline__ void psi_rows3(const PsiSweepArgs<R> &A, int s, int ja, int jb, R &r1, R &r2, R &r3) {
    constexpr int B = Pipe<R, SH>::ROWS2;
    const ColCtx X = col_ctx<3>(A.Nx, s);
    const int lane = threadIdx.x & (WAVE - 1);
    const int c = X.c;
    const bool own = X.own;                                   // lanes 3..60: outputs
    const bool mid1 = X.col && lane >= 1 && lane < WAVE - 1;  // lanes 1..62: sweep j
    const bool mid2 = X.col && lane >= 2 && lane < WAVE - 2;  // lanes 2..61: sweep j+1
    const int Nx = A.Nx, Ny = A.Ny, Nxa = Nx - 1;
    const R *b = A.ab + (long)Nxa * Ny;
    const bool noisy = A.rhs_out != nullptr;
    const bool noise_on = noisy && A.jstep == 0 && A.langevin_c > R(1.0e-32);
    const bool has_mt = A.mt != nullptr;
    const uint8_t *mtp = has_mt ? A.mt : reinterpret_cast<const uint8_t *>(A.ab);
    const R *emp = A.eps_map ? A.eps_map : reinterpret_cast<const R *>(A.rhs);
    const R gb = GAUGE ? gauge_phase_b<R>(A, c) : R(0);
    const cplx<R> zero = {R(0), R(0)};
    const int t0 = ja > 1 ? ja - 2 : 0;

    // sweep j window: psi^j at rows t-1, t; U of b(t-1); cell (c, t-1)
    cplx<R> pm = zero, pc = zero;
    R cb1 = R(1), sb1 = R(0);
    int cell1 = 0;
    if (X.col) {
        pc = ld(A.psi, c + (long)Nx * t0);
        if (t0 > 0) {
            pm = ld(A.psi, c + (long)Nx * (t0 - 1));
            sincos_r(link_phase<R, GAUGE>(A.dy, b[c + (long)Nx * (t0 - 1)], gb), &sb1, &cb1);
            if (c < Nxa) cell1 = has_mt ? int(A.mt[c + (long)Nxa * (t0 - 1)]) : 1;
        }
    }
    mask_link(shfl_up1(cell1) != 0 || cell1 != 0, cb1, sb1);  // b(c, t0-1)
    // sweep j+1 window (row u = t - 1): psi^{j+1} at u-1, u; U of a(u),
    // b(u-1); rhs, diagonal and material test of node (c, u)
    cplx<R> qm = zero, qc = zero, rhs_u = zero;
    R ca2 = R(0), sa2 = R(0), cal2 = R(0), sal2 = R(0), cb2 = R(0), sb2 = R(0), den_u = R(1);
    bool in_u = false;
    // sweep j+2 window (row v = t - 2): psi^{j+2} at v-1, v; U of a(v),
    // b(v-1); rhs, diagonal and material test of node (c, v)
    cplx<R> wm = zero, wc = zero, rhs_v = zero;
    R ca3 = R(0), sa3 = R(0), cal3 = R(0), sal3 = R(0), cb3 = R(0), sb3 = R(0), den_v = R(1);
    bool in_v = false;

    auto row = [&](const PsiRow<R> &cu, int t) {
        const long rowoff = (long)Nx * t;
        const bool real = t < Ny;  // t >= Ny: past the last row (later sweeps only)
        const cplx<R> pp = (X.col && t + 1 < Ny) ? cu.pp : zero;
        const R av = X.hasa ? cu.av : R(0);
        const R bv = (X.col && t < Ny - 1) ? cu.bv : R(0);
        const int cellc = (X.hasa && t < Ny - 1) ? (has_mt ? cu.cell : 1) : 0;
        cplx<R> rhs = cu.rhs;
        const R e = A.eps_map ? cu.e : A.eps;
        R sa, ca, sb, cb;
        sincos_r(link_phase<R, GAUGE>(A.dx, av, GAUGE ? gauge_phase_a<R>(A, t) : R(0)), &sa, &ca);
        sincos_r(link_phase<R, GAUGE>(A.dy, bv, gb), &sb, &cb);
        const bool mm = shfl_up1(cell1) != 0;  // cell (c-1, t-1)
        const bool mp = shfl_up1(cellc) != 0;  // cell (c-1, t)
        const bool pmc = cell1 != 0, ppc = cellc != 0;
        mask_link(pmc || ppc, ca, sa);  // a(c, t): cells (c, t-1), (c, t)
        mask_link(mp || ppc, cb, sb);   // b(c, t): cells (c-1, t), (c, t)
        const R cal = shfl_up1(ca), sal = shfl_up1(sa);
        const bool inside = mm || mp || pmc || ppc;

        // sweep j at row t -> psi^{j+1}(t)
        cplx<R> q = zero;
        R den = R(1);
        {
            const R pl_re = shfl_up1(pc.re), pl_im = shfl_up1(pc.im);
            const R pr_re = shfl_dn1(pc.re), pr_im = shfl_dn1(pc.im);
            if (mid1 && real && inside) {
                if (noise_on) psi_noise(A, (uint32_t)(c + rowoff), rhs);
                den = psi_den(A, rhs, e, mm, mp, pmc, ppc);
                q = psi_apply(A, rhs, den, pl_re, pl_im, pr_re, pr_im, pm, pp, cal, sal, ca, sa, cb1, sb1, cb, sb);
            }
            if (own && real && t >= ja && t < jb) {
                if (noisy && A.jstep == 0) st(A.rhs_out, c + rowoff, rhs);
                r1 = fmax(r1, fmax(fabs(q.re - pc.re), fabs(q.im - pc.im)));
            }
        }
        // sweep j+1 at row u = t - 1 -> psi^{j+2}(u)
        const int u = t - 1;
        cplx<R> w = zero;
        if (u >= 0 && u < Ny) {
            const R pl_re = shfl_up1(qc.re), pl_im = shfl_up1(qc.im);
            const R pr_re = shfl_dn1(qc.re), pr_im = shfl_dn1(qc.im);
            if (mid2 && in_u)
                w = psi_apply(A, rhs_u, den_u, pl_re, pl_im, pr_re, pr_im, qm, q, cal2, sal2, ca2, sa2, cb2, sb2, cb1,
                              sb1);
            if (own && u >= ja && u < jb) r2 = fmax(r2, fmax(fabs(w.re - qc.re), fabs(w.im - qc.im)));
        }
        // sweep j+2 at row v = t - 2 -> psi^{j+3}(v)
        const int v = t - 2;
        if (v >= ja) {
            const R pl_re = shfl_up1(wc.re), pl_im = shfl_up1(wc.im);
            const R pr_re = shfl_dn1(wc.re), pr_im = shfl_dn1(wc.im);
            if (own) {
                cplx<R> nxt = zero;
                if (in_v)
                    nxt = psi_apply(A, rhs_v, den_v, pl_re, pl_im, pr_re, pr_im, wm, w, cal3, sal3, ca3, sa3, cb3, sb3,
                                    cb2, sb2);
                st_at(A.psi_next + (size_t)Nx * v, 2 * X.cc * (uint32_t)sizeof(R), nxt);
                r3 = fmax(r3, fmax(fabs(nxt.re - wc.re), fabs(nxt.im - wc.im)));
            }
        }
        // roll the three windows
        wm = wc;
        wc = w;
        cb3 = cb2;
        sb3 = sb2;
        ca3 = ca2;
        sa3 = sa2;
        cal3 = cal2;
        sal3 = sal2;
        in_v = in_u;
        den_v = den_u;
        rhs_v = rhs_u;
        qm = qc;
        qc = q;
        cb2 = cb1;
        sb2 = sb1;
        ca2 = ca;
        sa2 = sa;
        cal2 = cal;
        sal2 = sal;
        in_u = inside;
        den_u = den;
        rhs_u = rhs;
        pm = pc;
        pc = pp;
        cb1 = cb;
        sb1 = sb;
        cell1 = cellc;
    };

    const int t1 = jb + 2;  // rows t in [t0, t1)
    PsiRow<R> P0[B], P1[B];
#pragma unroll
    for (int k = 0; k < B; ++k) P0[k] = psi_load_row(A, X, mtp, emp, min(t0 + k, t1 - 1));
    for (int t = t0; t < t1; t += 2 * B) {
#pragma unroll
        for (int k = 0; k < B; ++k) P1[k] = psi_load_row(A, X, mtp, emp, min(t + B + k, t1 - 1));
#pragma unroll
        for (int k = 0; k < B; ++k)
            if (t + k < t1) row(P0[k], t + k);
#pragma unroll
        for (int k = 0; k < B; ++k) P0[k] = psi_load_row(A, X, mtp, emp, min(t + 2 * B + k, t1 - 1));
#pragma unroll
        for (int k = 0; k < B; ++k)
            if (t + B + k < t1) row(P1[k], t + B + k);
    }
}

// A copy of v in registers of its own. The record ring keeps a row's
// right-hand side for K - 1 more rows; without the copy the register
// allocator keeps it in the load's destination, so the next load of that
// row buffer needs other registers and the loop's back edge then copies
// loads still in flight, which waits for all of them.
template <typename R> __device__ __forceinline__ R own_copy(R v) {
    if constexpr (std::is_same<R, float>::value) {
        R o;
        asm("v_mov_b32 %0, %1" : "=v"(o) : "v"(v));
        return o;
    } else {
        return v;
    }
}
template <typename R> __device__ __forceinline__ cplx<R> own_copy(cplx<R> v) {
    return {own_copy(v.re), own_copy(v.im)};
}

// What stage 0 hands down the stages from its row r: the link factors of
// a(c, r), of the link to its left and of b(c, r-1), and the node's
// right-hand side (FP32: rhs * den, psi_apply_folded), diagonal and
// material test.
template <typename R> struct PsiStage {
    R ca, sa, cal, sal, cbd, sbd;
    cplx<R> rhs;
    R den;
    bool in;
};

// K >= 2 fused psi sweeps in one pass, psi_rows3's pipeline with K stages:
// at row t, stage s runs sweep j+s at row t-s on lanes s+1 .. 62-s from the
// rows stage s-1 left in registers (stage 0 from the loaded rows), so the
// strips are WAVE - 2K columns wide and each stage's lateral neighbours come
// from the lanes either side. Link factors, cells and node diagonals are
// computed once per row by stage 0 and handed down the stages. Writes
// psi^{j+K} (and, at jstep 0, the Langevin right-hand side); r[s] gets the
// lane's largest |delta| of sweep j+s.
//
// The loop body takes Q = 4 rows, and everything a row leaves for later
// rows (stage 0's record, each stage's result, psi^j, the b link factors
// and the cell flag) sits in a ring of Q entries indexed by the row's place
// p in the body: row p writes entry p and reads entries p-1 .. p-K+1, all
// compile-time indices. Each value is written once where it is computed
// and read where it is needed, so no window is copied from row to row (a
// rolled window costs a register move per value per row).
template <typename R, int K, bool GAUGE, bool SH = false>
__device__ __forceinline__ void psi_rowsK(const PsiSweepArgs<R> &A, int s_, int ja, int jb, R (&r)[K]) {
    static_assert(K >= 2, "two or more fused sweeps");
    constexpr int Q = 4;
    static_assert(K <= Q, "a ring entry is read at most Q - 1 rows after it is written");
    const ColCtx X = col_ctx<K>(A.Nx, s_);
    const int lane = threadIdx.x & (WAVE - 1);
    const int c = X.c;
    const bool own = X.own;  // lanes K .. 63-K: the last stage's outputs
    const int Nx = A.Nx, Ny = A.Ny, Nxa = Nx - 1;
    const R *b = A.ab + (long)Nxa * Ny;
    const bool noisy = A.rhs_out != nullptr;
    const bool noise_on = noisy && A.jstep == 0 && A.langevin_c > R(1.0e-32);
    const bool has_mt = A.mt != nullptr;
    const uint8_t *mtp = has_mt ? A.mt : reinterpret_cast<const uint8_t *>(A.ab);
    const R *emp = A.eps_map ? A.eps_map : reinterpret_cast<const R *>(A.rhs);
    const R gb = GAUGE ? gauge_phase_b<R>(A, c) : R(0);
    const cplx<R> zero = {R(0), R(0)};
    const int t0 = ja > K - 1 ? ja - (K - 1) : 0;
    // FP32 folds dt/h^2 into the link factors and rhs * den into one value
    // per row (psi_apply_folded); FP64 keeps the reference's operation order
    constexpr bool FOLD = std::is_same<R, float>::value;
    const R kxa = FOLD ? A.dt * A.idx2 : R(1), kyb = FOLD ? A.dt * A.idy2 : R(1);

    // rings, entry (p mod Q) written by the row at place p of the body:
    // pj[p]: psi^j at that row (written one row earlier, as its pp);
    // cbq/sbq/cq: its b link factors and cell flag; rec: stage 0's record;
    // o[s]: stage s's result. The entries before the first row hold what
    // the pipeline starts from: psi^j at rows t0 - 1 and t0, b(c, t0-1),
    // cell (c, t0-1); empty stages.
    cplx<R> pj[Q], o[K - 1][Q];
    R cbq[Q], sbq[Q];
    int cq[Q];
    PsiStage<R> rec[Q];
#pragma unroll
    for (int k = 0; k < Q; ++k) {
        pj[k] = zero;
        cbq[k] = R(0);
        sbq[k] = R(0);
        cq[k] = 0;
        rec[k] = PsiStage<R>{R(0), R(0), R(0), R(0), R(0), R(0), zero, R(1), false};
#pragma unroll
        for (int st_ = 0; st_ < K - 1; ++st_) o[st_][k] = zero;
    }
    {
        cplx<R> pm = zero, pc = zero;
        R cb1 = R(1), sb1 = R(0);
        int cell1 = 0;
        if (X.col) {
            pc = ld(A.psi, c + (long)Nx * t0);
            if (t0 > 0) {
                pm = ld(A.psi, c + (long)Nx * (t0 - 1));
                sincos_r(link_phase<R, GAUGE>(A.dy, b[c + (long)Nx * (t0 - 1)], gb), &sb1, &cb1);
                if (c < Nxa) cell1 = has_mt ? int(A.mt[c + (long)Nxa * (t0 - 1)]) : 1;
            }
        }
        mask_link(shfl_up1(cell1) != 0 || cell1 != 0, cb1, sb1);  // b(c, t0-1)
        if constexpr (FOLD) {
            cb1 *= kyb;
            sb1 *= kyb;
        }
        pj[Q - 1] = pm;
        pj[0] = pc;
        cbq[Q - 1] = cb1;
        sbq[Q - 1] = sb1;
        cq[Q - 1] = cell1;
    }

    // Every stage is computed on all lanes and its result selected (zero
    // where the stage has no node), so a row is one basic block the
    // scheduler can interleave across stages; only the last stage's store
    // stays under a lane condition. The Langevin variant (NZ: the first
    // sweep of a step with a right-hand side to record) is a second copy of
    // the loop, so the common one carries no branch for it.
    auto row = [&](const PsiRow<R> &cu, int t, auto place, auto nz) {
        constexpr int p = decltype(place)::value;
        constexpr bool NZ = decltype(nz)::value;
        constexpr int p1 = (p + Q - 1) % Q, pn = (p + 1) % Q;
        const long rowoff = (long)Nx * t;
        const bool real = t < Ny;  // t >= Ny: past the last row (later stages only)
        const cplx<R> pc = pj[p], pm = pj[p1];
        const R cb1 = cbq[p1], sb1 = sbq[p1];
        const int cell1 = cq[p1];
        const cplx<R> pp = (X.col && t + 1 < Ny) ? cu.pp : zero;
        const R av = X.hasa ? cu.av : R(0);
        const R bv = (X.col && t < Ny - 1) ? cu.bv : R(0);
        const int cellc = (X.hasa && t < Ny - 1) ? (has_mt ? cu.cell : 1) : 0;
        cplx<R> rhs = cu.rhs;
        const R e = A.eps_map ? cu.e : A.eps;
        R sa, ca, sb, cb;
        sincos_r(link_phase<R, GAUGE>(A.dx, av, GAUGE ? gauge_phase_a<R>(A, t) : R(0)), &sa, &ca);
        sincos_r(link_phase<R, GAUGE>(A.dy, bv, gb), &sb, &cb);
        const bool mm = shfl_up1(cell1) != 0;  // cell (c-1, t-1)
        const bool mp = shfl_up1(cellc) != 0;  // cell (c-1, t)
        const bool pmc = cell1 != 0, ppc = cellc != 0;
        mask_link(pmc || ppc, ca, sa);  // a(c, t): cells (c, t-1), (c, t)
        mask_link(mp || ppc, cb, sb);   // b(c, t): cells (c-1, t), (c, t)
        if constexpr (FOLD) {
            ca *= kxa;
            sa *= kxa;
            cb *= kyb;
            sb *= kyb;
        }
        const R cal = shfl_up1(ca), sal = shfl_up1(sa);
        const bool inside = mm || mp || pmc || ppc;

        cplx<R> out[K];  // out[s]: sweep j+s's result at row t-s
        // stage 0: sweep j at row t
        R den;
        cplx<R> rd;  // what the later stages take of the right-hand side
        {
            const bool act = X.col && lane >= 1 && lane < WAVE - 1 && real && inside;
            if constexpr (NZ) {
                if (noise_on) {
                    cplx<R> nr = rhs;
                    psi_noise(A, (uint32_t)(c + rowoff), nr);
                    rhs = act ? nr : rhs;
                }
            }
            const R d = psi_den(A, rhs, e, mm, mp, pmc, ppc);
            den = act ? d : R(1);
            cplx<R> qa;
            if constexpr (FOLD) {
                rd = {rhs.re * den, rhs.im * den};
                qa = psi_apply_folded(rd, den, shfl_up1(pc.re), shfl_up1(pc.im), shfl_dn1(pc.re), shfl_dn1(pc.im), pm,
                                      pp, cal, sal, ca, sa, cb1, sb1, cb, sb);
            } else {
                rd = own_copy(rhs);
                qa = psi_apply(A, rhs, d, shfl_up1(pc.re), shfl_up1(pc.im), shfl_dn1(pc.re), shfl_dn1(pc.im), pm, pp,
                               cal, sal, ca, sa, cb1, sb1, cb, sb);
            }
            const cplx<R> q = act ? qa : zero;
            const bool res = own && real && t >= ja && t < jb;
            if constexpr (NZ) {
                if (res) st(A.rhs_out, c + rowoff, rhs);
            }
            r[0] = res ? fmax(r[0], fmax(fabs(q.re - pc.re), fabs(q.im - pc.im))) : r[0];
            out[0] = q;
        }
        // stage 0's record of row t, for stage s at row t + s
        rec[p] = PsiStage<R>{ca, sa, cal, sal, cb1, sb1, rd, den, inside};
        // stages 1 .. K-1: sweep j+s at row t-s, from stage s-1's results of
        // rows t-s-1 and t-s (computed at rows t-2 and t-1) and of row t-s+1
        // (this row); its down link is row t-s's b(c, t-s-1), its up link
        // row t-s+1's
#pragma unroll
        for (int st_ = 1; st_ < K; ++st_) {
            const PsiStage<R> &W = rec[(p + Q - st_) % Q];
            const R ucb = rec[(p + Q - st_ + 1) % Q].cbd, usb = rec[(p + Q - st_ + 1) % Q].sbd;
            const cplx<R> wc = o[st_ - 1][p1], wm = o[st_ - 1][(p + Q - 2) % Q];
            const int rs = t - st_;
            cplx<R> wa;
            if constexpr (FOLD)
                wa = psi_apply_folded(W.rhs, W.den, shfl_up1(wc.re), shfl_up1(wc.im), shfl_dn1(wc.re), shfl_dn1(wc.im),
                                      wm, out[st_ - 1], W.cal, W.sal, W.ca, W.sa, W.cbd, W.sbd, ucb, usb);
            else
                wa = psi_apply(A, W.rhs, W.den, shfl_up1(wc.re), shfl_up1(wc.im), shfl_dn1(wc.re), shfl_dn1(wc.im),
                               wm, out[st_ - 1], W.cal, W.sal, W.ca, W.sa, W.cbd, W.sbd, ucb, usb);
            const bool res = own && rs >= ja && rs < jb;  // rows of [ja, jb) lie in [0, Ny)
            cplx<R> w;
            if (st_ < K - 1) {
                const bool act = rs >= 0 && rs < Ny && X.col && lane >= st_ + 1 && lane < WAVE - 1 - st_ && W.in;
                w = act ? wa : zero;
            } else {
                w = W.in ? wa : zero;
                if (res) st_at(A.psi_next + (size_t)Nx * rs, 2 * X.cc * (uint32_t)sizeof(R), w);
            }
            r[st_] = res ? fmax(r[st_], fmax(fabs(w.re - wc.re), fabs(w.im - wc.im))) : r[st_];
            out[st_] = w;
        }
#pragma unroll
        for (int st_ = 0; st_ < K - 1; ++st_) o[st_][p] = out[st_];
        pj[pn] = pp;
        cbq[p] = cb;
        sbq[p] = sb;
        cq[p] = cellc;
    };

    // Q row buffers: before row t + p is evaluated, the loads of row
    // t + p + Q - 1 go into the buffer row t + p - 1 has finished with, so
    // each row's loads are in flight while the Q - 1 rows before it run.
    // The body evaluates all Q rows unconditionally (up to Q - 1 rows past
    // t1 - 1 in the last trip: their loads are clamped to row t1 - 1, and
    // they store nothing and change no residual since every store and
    // residual is limited to rows below jb), which keeps the loads and
    // stores straight-line, so the memory counter waits stay exact across
    // the loop's back edge instead of draining it.
    using P0_ = std::integral_constant<int, 0>;
    using P1_ = std::integral_constant<int, 1>;
    using P2_ = std::integral_constant<int, 2>;
    using P3_ = std::integral_constant<int, 3>;
    static_assert(Q == 4, "the body below is written out for four rows");
    const int t1 = jb + K - 1;  // rows t in [t0, t1)
    auto run = [&](auto nz) {
        PsiRow<R> P[Q];
#pragma unroll
        for (int k = 0; k < Q - 1; ++k) {
            P[k] = psi_load_row(A, X, mtp, emp, min(t0 + k, t1 - 1));
            __builtin_amdgcn_sched_barrier(0);  // issue in row order, as the loop does
        }
        for (int t = t0; t < t1; t += Q) {
            P[3] = psi_load_row(A, X, mtp, emp, min(t + 3, t1 - 1));
            row(P[0], t, P0_{}, nz);
            P[0] = psi_load_row(A, X, mtp, emp, min(t + 4, t1 - 1));
            row(P[1], t + 1, P1_{}, nz);
            P[1] = psi_load_row(A, X, mtp, emp, min(t + 5, t1 - 1));
            row(P[2], t + 2, P2_{}, nz);
            P[2] = psi_load_row(A, X, mtp, emp, min(t + 6, t1 - 1));
            row(P[3], t + 3, P3_{}, nz);
        }
    };
    if (noisy && A.jstep == 0)
        run(std::true_type{});
    else
        run(std::false_type{});
}

// Raw loads of one row for the A sweep: a(c, j+1), b(c, j), psi(c, j+1), the
// mask byte of cell (c, j), and the owned links' source and right-hand side
// at row j.
template <typename R> struct ARow {
    R ap, bc, sa, sb, ra, rb;
    cplx<R> pp;
    int cell;
};

template <typename R>
__device__ __forceinline__ ARow<R> a_load_row(const ASweepArgs<R> &A, const ColCtx &X, const uint8_t *mtp, int j) {
    const uint32_t Nx = A.Nx, Ny = A.Ny, Nxa = Nx - 1, Na = Nxa * Ny;
    // rows past the last one (the later sweeps of a fused pass load up to
    // two beyond) are clamped like the others: their values are not used
    const uint32_t jr = min((uint32_t)j, Ny - 1), jp = min((uint32_t)j + 1, Ny - 1), jq = min((uint32_t)j, Ny - 2);
    const uint32_t oc = X.cc * (uint32_t)sizeof(R), oa = X.ca * (uint32_t)sizeof(R);
    const size_t ra = (size_t)Nxa * jr, rb = (size_t)Na + (size_t)Nx * jq;
    ARow<R> L;
    L.ap = at(A.ab + (size_t)Nxa * jp, oa);
    L.bc = at(A.ab + rb, oc);
    L.pp = ld_at(A.psi + (size_t)Nx * jp, 2 * oc);
    L.cell = at(mtp + (size_t)Nxa * jq, X.ca);
    L.sa = at(A.src + ra, oa);
    L.ra = at(A.rhs + ra, oa);
    L.sb = at(A.src + rb, oc);
    L.rb = at(A.rhs + rb, oc);
    return L;
}

// The A sweep arguments for a row kernel. In the FP32 persistent solves the
// constants of the link updates are held in vector registers (see in_vgpr);
// FP64 keeps them in scalar registers (two vector registers each would push
// its kernels past two waves per SIMD), and so do the batched kernels.
template <typename R, bool SH>
__device__ __forceinline__ const ASweepArgs<R> &a_row_args(const ASweepArgs<R> &A0, ASweepArgs<R> &tmp) {
    if constexpr (std::is_same<R, float>::value && !SH) {
        tmp = A0;
        tmp.dx = in_vgpr(tmp.dx);
        tmp.dy = in_vgpr(tmp.dy);
        tmp.idx = in_vgpr(tmp.idx);
        tmp.idy = in_vgpr(tmp.idy);
        tmp.idx2 = in_vgpr(tmp.idx2);
        tmp.idy2 = in_vgpr(tmp.idy2);
        tmp.idxy = in_vgpr(tmp.idxy);
        tmp.dt_rho = in_vgpr(tmp.dt_rho);
        tmp.dt_rho_kappa2 = in_vgpr(tmp.dt_rho_kappa2);
        tmp.rh_a = in_vgpr(tmp.rh_a);
        tmp.rh_b = in_vgpr(tmp.rh_b);
        tmp.iden_a = in_vgpr(tmp.iden_a);
        tmp.iden_b = in_vgpr(tmp.iden_b);
        return tmp;
    } else {
        return A0;
    }
}

// Backward-Euler Jacobi update of the a link (c, j) (td.h:366-410): its
// supercurrent jl, the right-hand side ra, and the neighbouring links am =
// a(c, j-1), bm = b(c, j-1), brm = b(c+1, j-1), ap = a(c, j+1), bc = b(c, j),
// brc = b(c+1, j). FP32 multiplies by the rounded reciprocal of the
// constant denominator (within 1 ulp of the division).
template <typename R>
__device__ __forceinline__ R a_link_update(const ASweepArgs<R> &A, int j, R ra, R jl, R am, R bm, R brm, R ap, R bc,
                                           R brc) {
    const int Ny = A.Ny;
    R rh = R(0), dd = R(1);
    if (j == 0) { rh = A.rh_a; dd = R(2); }
    else if (j + 1 == Ny) { rh = -A.rh_a; dd = R(2); }
    R t1 = R(0), t2 = R(0);
    if (j > 0) t1 = A.idy2 * am - A.idxy * bm + A.idxy * brm;
    if (j + 1 < Ny) t2 = A.idy2 * ap + A.idxy * bc - A.idxy * brc;
    const R num = ra + A.dt_rho * (jl + rh) + A.dt_rho_kappa2 * dd * (t1 + t2);
    if constexpr (std::is_same<R, float>::value) return num * A.iden_a;
    else return num / A.den_a;
}

// The b link (c, j) (td.h:412-456): blc = b(c-1, j), alc = a(c-1, j),
// alp = a(c-1, j+1), brc = b(c+1, j), ac = a(c, j), ap = a(c, j+1).
template <typename R>
__device__ __forceinline__ R b_link_update(const ASweepArgs<R> &A, int c, R rb, R jl, R blc, R alc, R alp, R brc, R ac,
                                           R ap) {
    const int Nx = A.Nx;
    R rh = R(0), dd = R(1);
    if (c == 0) { rh = -A.rh_b; dd = R(2); }
    else if (c + 1 == Nx) { rh = A.rh_b; dd = R(2); }
    R t1 = R(0), t2 = R(0);
    if (c > 0) t1 = A.idx2 * blc - A.idxy * alc + A.idxy * alp;
    if (c + 1 < Nx) t2 = A.idx2 * brc + A.idxy * ac - A.idxy * ap;
    const R num = rb + A.dt_rho * (jl + rh) + A.dt_rho_kappa2 * dd * (t1 + t2);
    if constexpr (std::is_same<R, float>::value) return num * A.iden_b;
    else return num / A.den_b;
}

// Supercurrent of a link, id * Im(conj psi0 U psi1) in the reference's form.
template <typename R>
__device__ __forceinline__ R link_current(R id, cplx<R> p0, R x1, R y1, R cs, R sn) {
    const R x0 = p0.re, y0 = p0.im;
    return id * ((x0 * y1 - y0 * x1) * cs - (x0 * x1 + y0 * y1) * sn);
}

// One A sweep over rows [ja, jb) of strip s (both link families of every
// owned node). Returns the lane's largest |delta|.
template <typename R, bool GAUGE, bool SH = false>
__device__ __forceinline__ R a_rows(const ASweepArgs<R> &A0, int s, int ja, int jb) {
    ASweepArgs<R> tmp;
    const ASweepArgs<R> &A = a_row_args<R, SH>(A0, tmp);
    constexpr int B = Pipe<R, SH>::ROWS;
    const ColCtx X = col_ctx<1>(A.Nx, s);
    const int c = X.c;
    const bool own = X.own, hasa = X.hasa;
    const int Nx = A.Nx, Ny = A.Ny, Nxa = Nx - 1;
    const long Na = (long)Nxa * Ny;
    const R *a = A.ab;
    const R *b = A.ab + Na;
    const bool noisy = A.rhs_out != nullptr;
    const bool has_mt = A.mt != nullptr;
    const uint8_t *mtp = has_mt ? A.mt : reinterpret_cast<const uint8_t *>(A.ab);  // any readable bytes when absent
    const R gb = GAUGE ? gauge_phase_b<R>(A, c) : R(0);
    R rr = R(0);

    // window of this column: a(c, j-1), a(c, j), b(c, j-1), psi(c, j), cell (c, j-1)
    R am = R(0), ac = R(0), bm = R(0);
    cplx<R> pc = {R(0), R(0)};
    int cellm = 0;
    if (X.col) {
        pc = ld(A.psi, c + (long)Nx * ja);
        if (hasa) ac = a[c + (long)Nxa * ja];
        if (ja > 0) {
            if (hasa) am = a[c + (long)Nxa * (ja - 1)];
            bm = b[c + (long)Nx * (ja - 1)];
            if (hasa) cellm = has_mt ? int(A.mt[c + (long)Nxa * (ja - 1)]) : 1;
        }
    }

    auto row = [&](const ARow<R> &cu, int j) {
        const R ap = (hasa && j + 1 < Ny) ? cu.ap : R(0);
        const R bc = (X.col && j < Ny - 1) ? cu.bc : R(0);
        const cplx<R> pp = (X.col && j < Ny - 1) ? cu.pp : cplx<R>{R(0), R(0)};
        const int cellv = (hasa && j < Ny - 1) ? (has_mt ? cu.cell : 1) : 0;
        const bool cellc = cellv != 0;
        // right column values
        const R brc = shfl_dn1(bc), brm = shfl_dn1(bm);
        const R pr_re = shfl_dn1(pc.re), pr_im = shfl_dn1(pc.im);
        // left column values
        const R alc = shfl_up1(ac), alp = shfl_up1(ap), blc = shfl_up1(bc);
        const bool cl_c = shfl_up1(cellv) != 0;  // cell (c-1, j)

        const long na = c + (long)Nxa * j;
        const long nb = Na + c + (long)Nx * j;
        const bool owna = own && c < Nxa;
        const bool ownb = own && j < Ny - 1;
        R ra = cu.ra, rb = cu.rb;
        if (noisy && A.jstep == 0 && A.langevin_c > R(1.0e-32)) {
            if (owna) ra = ra + A.langevin_c * R(double(noise_1<R>((uint32_t)na, A.rand_t)) - 0.5);
            if (ownb) rb = rb + A.langevin_c * R(double(noise_2<R>((uint32_t)nb, A.rand_t)) - 0.5);
        }
        R ssa, csa, ssb, csb;
        sincos_r(link_phase<R, GAUGE>(A.dx, cu.sa, GAUGE ? gauge_phase_a<R>(A, j) : R(0)), &ssa, &csa);
        sincos_r(link_phase<R, GAUGE>(A.dy, cu.sb, gb), &ssb, &csb);

        if (owna) {
            // cells (c, j-1), (c, j)
            const R jl = (cellm != 0 || cellc) ? link_current(A.idx, pc, pr_re, pr_im, csa, ssa) : R(0);
            const R v = a_link_update(A, j, ra, jl, am, bm, brm, ap, bc, brc);
            if (noisy && A.jstep == 0) A.rhs_out[na] = ra;
            st_at(A.ab_next + (size_t)Nxa * j, X.cc * (uint32_t)sizeof(R), v);
            rr = fmax(rr, fabs(v - ac));
        }
        if (ownb) {
            // cells (c-1, j), (c, j)
            const R jl = (cl_c || cellc) ? link_current(A.idy, pc, pp.re, pp.im, csb, ssb) : R(0);
            const R v = b_link_update(A, c, rb, jl, blc, alc, alp, brc, ac, ap);
            if (noisy && A.jstep == 0) A.rhs_out[nb] = rb;
            st_at(A.ab_next + (size_t)Na + (size_t)Nx * j, X.cc * (uint32_t)sizeof(R), v);
            rr = fmax(rr, fabs(v - bc));
        }
        // roll
        am = ac;
        ac = ap;
        bm = bc;
        pc = pp;
        cellm = cellv;
    };

    ARow<R> P0[B], P1[B];
#pragma unroll
    for (int k = 0; k < B; ++k) P0[k] = a_load_row(A, X, mtp, min(ja + k, jb - 1));
    for (int j = ja; j < jb; j += 2 * B) {
#pragma unroll
        for (int k = 0; k < B; ++k) P1[k] = a_load_row(A, X, mtp, min(j + B + k, jb - 1));
#pragma unroll
        for (int k = 0; k < B; ++k)
            if (j + k < jb) row(P0[k], j + k);
#pragma unroll
        for (int k = 0; k < B; ++k) P0[k] = a_load_row(A, X, mtp, min(j + 2 * B + k, jb - 1));
#pragma unroll
        for (int k = 0; k < B; ++k)
            if (j + B + k < jb) row(P1[k], j + B + k);
    }
    return rr;
}

// Two consecutive A sweeps j and j+1 (j even) in one pass over rows
// [ja, jb) of strip s, strips of WAVE - 4 columns as in psi_rows2. Over such
// a pair the supercurrent source is the pair's input for both sweeps
// (A^(2*floor(j/2)) = A^j), so each link's supercurrent is evaluated once;
// the first sweep runs one row ahead of the second on lanes 1..62 and its
// result A^{j+1} stays in registers. Writes A^{j+2} (and, at jstep 0, the
// Langevin right-hand side); r1 and r2 get the lane's largest |delta| of
// each sweep. Each link update is that of a_rows.
template <typename R, bool GAUGE, bool SH = false>
__device__ __forceinline__ void a_rows2(const ASweepArgs<R> &A0, int s, int ja, int jb, R &r1, R &r2) {
    ASweepArgs<R> tmp;
    const ASweepArgs<R> &A = a_row_args<R, SH>(A0, tmp);
    constexpr int B = Pipe<R, SH>::AROWS2;
    const ColCtx X = col_ctx<2>(A.Nx, s);
    const int lane = threadIdx.x & (WAVE - 1);
    const int c = X.c;
    const bool own = X.own, hasa = X.hasa;
    const bool mid = X.col && lane >= 1 && lane < WAVE - 1;
    const int Nx = A.Nx, Ny = A.Ny, Nxa = Nx - 1;
    const long Na = (long)Nxa * Ny;
    const R *a = A.ab;
    const R *b = A.ab + Na;
    const bool noisy = A.rhs_out != nullptr;
    const bool noise_on = noisy && A.jstep == 0 && A.langevin_c > R(1.0e-32);
    const bool has_mt = A.mt != nullptr;
    const uint8_t *mtp = has_mt ? A.mt : reinterpret_cast<const uint8_t *>(A.ab);
    const R gb = GAUGE ? gauge_phase_b<R>(A, c) : R(0);
    const cplx<R> zero = {R(0), R(0)};
    const int t0 = ja > 0 ? ja - 1 : 0;

    // first sweep window (row t): a(c, t-1), a(c, t), b(c, t-1), psi(c, t), cell (c, t-1)
    R am = R(0), ac = R(0), bm = R(0);
    cplx<R> pc = zero;
    int cellm = 0;
    if (X.col) {
        pc = ld(A.psi, c + (long)Nx * t0);
        if (hasa) ac = a[c + (long)Nxa * t0];
        if (t0 > 0) {
            if (hasa) am = a[c + (long)Nxa * (t0 - 1)];
            bm = b[c + (long)Nx * (t0 - 1)];
            if (hasa) cellm = has_mt ? int(A.mt[c + (long)Nxa * (t0 - 1)]) : 1;
        }
    }
    // second sweep window (row u = t - 1): A^{j+1} at a(c, u-1), a(c, u),
    // b(c, u-1), b(c, u); supercurrents and right-hand sides of row u
    R am1 = R(0), ac1 = R(0), bm1 = R(0), bc1 = R(0), jla1 = R(0), jlb1 = R(0), ra1 = R(0), rb1 = R(0);

    auto row = [&](const ARow<R> &cu, int t) {
        const bool real = t < Ny;
        const R ap = (hasa && t + 1 < Ny) ? cu.ap : R(0);
        const R bc = (X.col && t < Ny - 1) ? cu.bc : R(0);
        const cplx<R> pp = (X.col && t < Ny - 1) ? cu.pp : zero;
        const int cellv = (hasa && t < Ny - 1) ? (has_mt ? cu.cell : 1) : 0;
        const bool cellc = cellv != 0;
        const R brc = shfl_dn1(bc), brm = shfl_dn1(bm);
        const R pr_re = shfl_dn1(pc.re), pr_im = shfl_dn1(pc.im);
        const R alc = shfl_up1(ac), alp = shfl_up1(ap), blc = shfl_up1(bc);
        const bool cl_c = shfl_up1(cellv) != 0;  // cell (c-1, t)
        const long na = c + (long)Nxa * t;
        const long nb = Na + c + (long)Nx * t;
        const bool hasa_t = hasa && real;
        const bool hasb_t = X.col && t < Ny - 1;
        R ra = cu.ra, rb = cu.rb;
        if (noise_on) {
            if (mid && hasa_t) ra = ra + A.langevin_c * R(double(noise_1<R>((uint32_t)na, A.rand_t)) - 0.5);
            if (mid && hasb_t) rb = rb + A.langevin_c * R(double(noise_2<R>((uint32_t)nb, A.rand_t)) - 0.5);
        }
        R ssa, csa, ssb, csb;
        sincos_r(link_phase<R, GAUGE>(A.dx, cu.sa, GAUGE ? gauge_phase_a<R>(A, t) : R(0)), &ssa, &csa);
        sincos_r(link_phase<R, GAUGE>(A.dy, cu.sb, gb), &ssb, &csb);
        const R jla = (cellm != 0 || cellc) ? link_current(A.idx, pc, pr_re, pr_im, csa, ssa) : R(0);
        const R jlb = (cl_c || cellc) ? link_current(A.idy, pc, pp.re, pp.im, csb, ssb) : R(0);

        // first sweep at row t -> A^{j+1}(t)
        R va = R(0), vb = R(0);
        if (mid && hasa_t) va = a_link_update(A, t, ra, jla, am, bm, brm, ap, bc, brc);
        if (mid && hasb_t) vb = b_link_update(A, c, rb, jlb, blc, alc, alp, brc, ac, ap);
        if (own && t >= ja && t < jb) {
            if (hasa_t) {
                if (noisy && A.jstep == 0) A.rhs_out[na] = ra;
                r1 = fmax(r1, fabs(va - ac));
            }
            if (hasb_t) {
                if (noisy && A.jstep == 0) A.rhs_out[nb] = rb;
                r1 = fmax(r1, fabs(vb - bc));
            }
        }
        // second sweep at row u = t - 1 -> A^{j+2}(u)
        const int u = t - 1;
        {
            const R brc1 = shfl_dn1(bc1), brm1 = shfl_dn1(bm1);
            const R alc1 = shfl_up1(ac1), alp1 = shfl_up1(va), blc1 = shfl_up1(bc1);
            if (u >= ja && own) {
                if (hasa) {
                    const R v = a_link_update(A, u, ra1, jla1, am1, bm1, brm1, va, bc1, brc1);
                    st_at(A.ab_next + (size_t)Nxa * u, X.cc * (uint32_t)sizeof(R), v);
                    r2 = fmax(r2, fabs(v - ac1));
                }
                if (u < Ny - 1) {
                    const R v = b_link_update(A, c, rb1, jlb1, blc1, alc1, alp1, brc1, ac1, va);
                    st_at(A.ab_next + (size_t)Na + (size_t)Nx * u, X.cc * (uint32_t)sizeof(R), v);
                    r2 = fmax(r2, fabs(v - bc1));
                }
            }
        }
        // roll both windows
        am1 = ac1;
        ac1 = va;
        bm1 = bc1;
        bc1 = vb;
        jla1 = jla;
        jlb1 = jlb;
        ra1 = ra;
        rb1 = rb;
        am = ac;
        ac = ap;
        bm = bc;
        pc = pp;
        cellm = cellv;
    };

    const int t1 = jb + 1;  // rows t in [t0, t1)
    ARow<R> P0[B], P1[B];
#pragma unroll
    for (int k = 0; k < B; ++k) P0[k] = a_load_row(A, X, mtp, min(min(t0 + k, t1 - 1), Ny - 1));
    for (int t = t0; t < t1; t += 2 * B) {
#pragma unroll
        for (int k = 0; k < B; ++k) P1[k] = a_load_row(A, X, mtp, min(min(t + B + k, t1 - 1), Ny - 1));
#pragma unroll
        for (int k = 0; k < B; ++k)
            if (t + k < t1) row(P0[k], t + k);
#pragma unroll
        for (int k = 0; k < B; ++k) P0[k] = a_load_row(A, X, mtp, min(min(t + 2 * B + k, t1 - 1), Ny - 1));
#pragma unroll
        for (int k = 0; k < B; ++k)
            if (t + B + k < t1) row(P1[k], t + B + k);
    }
}

// Three consecutive A sweeps j, j+1, j+2 (j even) in one pass over rows
// [ja, jb) of strip s, strips of WAVE - 6 columns. Sweeps j and j+1 take
// their supercurrent from the pass's input A^j, sweep j+2 from A^{j+2}, the
// output of sweep j+1, which the pass holds in registers (the source
// schedule A^(2*floor(j/2))). Sweep j runs at row t on lanes 1..62, sweep j+1
// at row t-1 on lanes 2..61, sweep j+2 at row t-2 on lanes 3..60. Writes
// A^{j+3} to ab_next and A^{j+2} to ab_mid (the source of sweep j+3 should
// the solve continue); r1, r2, r3 get the lane's largest |delta| of each
// sweep. Each link update is that of a_rows.
template <typename R, bool GAUGE, bool SH = false>
__device__ __forceinline__ void a_rows3(const ASweepArgs<R> &A0, int s, int ja, int jb, R &r1, R &r2, R &r3) {
    ASweepArgs<R> tmp;
    const ASweepArgs<R> &A = a_row_args<R, SH>(A0, tmp);
    constexpr int B = Pipe<R, SH>::AROWS3;
    // FP64 persistent solve: the link updates and sweep j+2's sincos are
    // computed on all lanes and selected (stores and residuals stay under
    // their lane conditions), so a row's sweeps can be interleaved; the same
    // expressions, so the same values
    constexpr bool BF = std::is_same<R, double>::value && !SH;
    const ColCtx X = col_ctx<3>(A.Nx, s);
    const int lane = threadIdx.x & (WAVE - 1);
    const int c = X.c;
    const bool own = X.own, hasa = X.hasa;
    const bool mid1 = X.col && lane >= 1 && lane < WAVE - 1;
    const bool mid2 = X.col && lane >= 2 && lane < WAVE - 2;
    const int Nx = A.Nx, Ny = A.Ny, Nxa = Nx - 1;
    const long Na = (long)Nxa * Ny;
    const R *a = A.ab;
    const R *b = A.ab + Na;
    const bool noisy = A.rhs_out != nullptr;
    const bool noise_on = noisy && A.jstep == 0 && A.langevin_c > R(1.0e-32);
    const bool has_mt = A.mt != nullptr;
    const uint8_t *mtp = has_mt ? A.mt : reinterpret_cast<const uint8_t *>(A.ab);
    const R gb = GAUGE ? gauge_phase_b<R>(A, c) : R(0);
    const cplx<R> zero = {R(0), R(0)};
    const uint32_t oc = X.cc * (uint32_t)sizeof(R);
    const int t0 = ja > 1 ? ja - 2 : 0;

    // sweep j window (row t): a(c, t-1), a(c, t), b(c, t-1), psi(c, t), cell (c, t-1)
    R am = R(0), ac = R(0), bm = R(0);
    cplx<R> pc = zero;
    int cellm = 0;
    if (X.col) {
        pc = ld(A.psi, c + (long)Nx * t0);
        if (hasa) ac = a[c + (long)Nxa * t0];
        if (t0 > 0) {
            if (hasa) am = a[c + (long)Nxa * (t0 - 1)];
            bm = b[c + (long)Nx * (t0 - 1)];
            if (hasa) cellm = has_mt ? int(A.mt[c + (long)Nxa * (t0 - 1)]) : 1;
        }
    }
    // sweep j+1 window (row u = t - 1): A^{j+1} at a(c, u-1), a(c, u),
    // b(c, u-1), b(c, u); supercurrents and right-hand sides of row u
    R am1 = R(0), ac1 = R(0), bm1 = R(0), bc1 = R(0), jla1 = R(0), jlb1 = R(0), ra1 = R(0), rb1 = R(0);
    // sweep j+2 window (row v = t - 2): A^{j+2} at a(c, v-1), a(c, v),
    // b(c, v-1), b(c, v); psi(c, v), psi(c, v+1); right-hand sides and
    // supercurrent switches of row v
    R am2 = R(0), ac2 = R(0), bm2 = R(0), bc2 = R(0), ra2 = R(0), rb2 = R(0);
    cplx<R> pv = zero, pu = zero;
    bool ena1 = false, enb1 = false, ena2 = false, enb2 = false;

    auto row = [&](const ARow<R> &cu, int t) {
        const bool real = t < Ny;
        const R ap = (hasa && t + 1 < Ny) ? cu.ap : R(0);
        const R bc = (X.col && t < Ny - 1) ? cu.bc : R(0);
        const cplx<R> pp = (X.col && t < Ny - 1) ? cu.pp : zero;
        const int cellv = (hasa && t < Ny - 1) ? (has_mt ? cu.cell : 1) : 0;
        const bool cellc = cellv != 0;
        const R brc = shfl_dn1(bc), brm = shfl_dn1(bm);
        const R pr_re = shfl_dn1(pc.re), pr_im = shfl_dn1(pc.im);
        const R alc = shfl_up1(ac), alp = shfl_up1(ap), blc = shfl_up1(bc);
        const bool cl_c = shfl_up1(cellv) != 0;  // cell (c-1, t)
        const long na = c + (long)Nxa * t;
        const long nb = Na + c + (long)Nx * t;
        const bool hasa_t = hasa && real;
        const bool hasb_t = X.col && t < Ny - 1;
        const bool ena = cellm != 0 || cellc;  // a(c, t) touches the material
        const bool enb = cl_c || cellc;        // b(c, t) touches the material
        R ra = cu.ra, rb = cu.rb;
        if (noise_on) {
            if (mid1 && hasa_t) ra = ra + A.langevin_c * R(double(noise_1<R>((uint32_t)na, A.rand_t)) - 0.5);
            if (mid1 && hasb_t) rb = rb + A.langevin_c * R(double(noise_2<R>((uint32_t)nb, A.rand_t)) - 0.5);
        }
        R ssa, csa, ssb, csb;
        sincos_r(link_phase<R, GAUGE>(A.dx, cu.sa, GAUGE ? gauge_phase_a<R>(A, t) : R(0)), &ssa, &csa);
        sincos_r(link_phase<R, GAUGE>(A.dy, cu.sb, gb), &ssb, &csb);
        const R jla = ena ? link_current(A.idx, pc, pr_re, pr_im, csa, ssa) : R(0);
        const R jlb = enb ? link_current(A.idy, pc, pp.re, pp.im, csb, ssb) : R(0);

        // sweep j at row t -> A^{j+1}(t)
        R va = R(0), vb = R(0);
        if constexpr (BF) {
            const R xa = a_link_update(A, t, ra, jla, am, bm, brm, ap, bc, brc);
            const R xb = b_link_update(A, c, rb, jlb, blc, alc, alp, brc, ac, ap);
            va = (mid1 && hasa_t) ? xa : R(0);
            vb = (mid1 && hasb_t) ? xb : R(0);
        } else {
            if (mid1 && hasa_t) va = a_link_update(A, t, ra, jla, am, bm, brm, ap, bc, brc);
            if (mid1 && hasb_t) vb = b_link_update(A, c, rb, jlb, blc, alc, alp, brc, ac, ap);
        }
        if (own && t >= ja && t < jb) {
            if (hasa_t) {
                if (noisy && A.jstep == 0) A.rhs_out[na] = ra;
                r1 = fmax(r1, fabs(va - ac));
            }
            if (hasb_t) {
                if (noisy && A.jstep == 0) A.rhs_out[nb] = rb;
                r1 = fmax(r1, fabs(vb - bc));
            }
        }
        // sweep j+1 at row u = t - 1 -> A^{j+2}(u)
        const int u = t - 1;
        R wa = R(0), wb = R(0);
        {
            const R brc1 = shfl_dn1(bc1), brm1 = shfl_dn1(bm1);
            const R alc1 = shfl_up1(ac1), alp1 = shfl_up1(va), blc1 = shfl_up1(bc1);
            if constexpr (BF) {
                const R xa = a_link_update(A, u, ra1, jla1, am1, bm1, brm1, va, bc1, brc1);
                const R xb = b_link_update(A, c, rb1, jlb1, blc1, alc1, alp1, brc1, ac1, va);
                const bool m2 = mid2 && u >= 0 && u < Ny;
                wa = (m2 && hasa) ? xa : R(0);
                wb = (m2 && u < Ny - 1) ? xb : R(0);
            } else if (mid2 && u >= 0 && u < Ny) {
                if (hasa) wa = a_link_update(A, u, ra1, jla1, am1, bm1, brm1, va, bc1, brc1);
                if (u < Ny - 1) wb = b_link_update(A, c, rb1, jlb1, blc1, alc1, alp1, brc1, ac1, va);
            }
            if (own && u >= ja && u < jb) {
                if (hasa) {
                    st_at(A.ab_mid + (size_t)Nxa * u, oc, wa);
                    r2 = fmax(r2, fabs(wa - ac1));
                }
                if (u < Ny - 1) {
                    st_at(A.ab_mid + (size_t)Na + (size_t)Nx * u, oc, wb);
                    r2 = fmax(r2, fabs(wb - bc1));
                }
            }
        }
        // sweep j+2 at row v = t - 2 -> A^{j+3}(v), source A^{j+2}
        const int v = t - 2;
        {
            const R brc2 = shfl_dn1(bc2), brm2 = shfl_dn1(bm2);
            const R alc2 = shfl_up1(ac2), alp2 = shfl_up1(wa), blc2 = shfl_up1(bc2);
            const R pvr_re = shfl_dn1(pv.re), pvr_im = shfl_dn1(pv.im);
            if constexpr (BF) {
                R s2a, c2a, s2b, c2b;
                sincos_r(link_phase<R, GAUGE>(A.dx, ac2, GAUGE ? gauge_phase_a<R>(A, v) : R(0)), &s2a, &c2a);
                sincos_r(link_phase<R, GAUGE>(A.dy, bc2, gb), &s2b, &c2b);
                const R jla2 = ena2 ? link_current(A.idx, pv, pvr_re, pvr_im, c2a, s2a) : R(0);
                const R xa = a_link_update(A, v, ra2, jla2, am2, bm2, brm2, wa, bc2, brc2);
                const R jlb2 = enb2 ? link_current(A.idy, pv, pu.re, pu.im, c2b, s2b) : R(0);
                const R xb = b_link_update(A, c, rb2, jlb2, blc2, alc2, alp2, brc2, ac2, wa);
                if (v >= ja && own) {
                    if (hasa) {
                        st_at(A.ab_next + (size_t)Nxa * v, oc, xa);
                        r3 = fmax(r3, fabs(xa - ac2));
                    }
                    if (v < Ny - 1) {
                        st_at(A.ab_next + (size_t)Na + (size_t)Nx * v, oc, xb);
                        r3 = fmax(r3, fabs(xb - bc2));
                    }
                }
            } else if (v >= ja && own) {
                R s2a, c2a, s2b, c2b;
                sincos_r(link_phase<R, GAUGE>(A.dx, ac2, GAUGE ? gauge_phase_a<R>(A, v) : R(0)), &s2a, &c2a);
                sincos_r(link_phase<R, GAUGE>(A.dy, bc2, gb), &s2b, &c2b);
                if (hasa) {
                    const R jl = ena2 ? link_current(A.idx, pv, pvr_re, pvr_im, c2a, s2a) : R(0);
                    const R x = a_link_update(A, v, ra2, jl, am2, bm2, brm2, wa, bc2, brc2);
                    st_at(A.ab_next + (size_t)Nxa * v, oc, x);
                    r3 = fmax(r3, fabs(x - ac2));
                }
                if (v < Ny - 1) {
                    const R jl = enb2 ? link_current(A.idy, pv, pu.re, pu.im, c2b, s2b) : R(0);
                    const R x = b_link_update(A, c, rb2, jl, blc2, alc2, alp2, brc2, ac2, wa);
                    st_at(A.ab_next + (size_t)Na + (size_t)Nx * v, oc, x);
                    r3 = fmax(r3, fabs(x - bc2));
                }
            }
        }
        // roll the three windows
        am2 = ac2;
        ac2 = wa;
        bm2 = bc2;
        bc2 = wb;
        ra2 = ra1;
        rb2 = rb1;
        ena2 = ena1;
        enb2 = enb1;
        pv = pu;
        pu = pc;
        am1 = ac1;
        ac1 = va;
        bm1 = bc1;
        bc1 = vb;
        jla1 = jla;
        jlb1 = jlb;
        ra1 = ra;
        rb1 = rb;
        ena1 = ena;
        enb1 = enb;
        am = ac;
        ac = ap;
        bm = bc;
        pc = pp;
        cellm = cellv;
    };

    const int t1 = jb + 2;  // rows t in [t0, t1)
    ARow<R> P0[B], P1[B];
#pragma unroll
    for (int k = 0; k < B; ++k) P0[k] = a_load_row(A, X, mtp, min(min(t0 + k, t1 - 1), Ny - 1));
    for (int t = t0; t < t1; t += 2 * B) {
#pragma unroll
        for (int k = 0; k < B; ++k) P1[k] = a_load_row(A, X, mtp, min(min(t + B + k, t1 - 1), Ny - 1));
#pragma unroll
        for (int k = 0; k < B; ++k)
            if (t + k < t1) row(P0[k], t + k);
#pragma unroll
        for (int k = 0; k < B; ++k) P0[k] = a_load_row(A, X, mtp, min(min(t + 2 * B + k, t1 - 1), Ny - 1));
#pragma unroll
        for (int k = 0; k < B; ++k)
            if (t + B + k < t1) row(P1[k], t + B + k);
    }
}

// ---- work partition --------------------------------------------------------
// A grid of W waves shares T = nstrips*Ny wave-rows; wave w takes the
// contiguous global rows [w*T/W, (w+1)*T/W), which may span two strips.
template <int HALO = 1> __host__ __device__ inline int nstrips_of(int Nx) {
    return (Nx + WAVE - 2 * HALO - 1) / (WAVE - 2 * HALO);
}

template <typename F>
__device__ __forceinline__ void for_rows(int Nx, int Ny, long w, long W, F &&f) {
    const long T = (long)nstrips_of(Nx) * Ny;
    long g0 = (w * T) / W, g1 = ((w + 1) * T) / W;
    while (g0 < g1) {
        const int s = (int)(g0 / Ny);
        const int ja = (int)(g0 % Ny);
        const long left = g1 - g0;
        const int jb = (int)((long)ja + left < (long)Ny ? (long)ja + left : (long)Ny);
        f(s, ja, jb);
        g0 += jb - ja;
    }
}


// Row-major partition for the persistent solves: the rows are cut into
// S = max(1, W / nstrips) segments and the work items k = (segment, strip)
// are numbered strip-fastest; wave w takes items w, w + U, ... with
// U = min(W, S * nstrips). Waves of consecutive index then sweep the same rows
// of adjacent strips at the same time (whole grid rows are read together);
// spare waves (w >= U) idle.
template <int HALO, typename F>
__device__ __forceinline__ void for_rows_rm(int Nx, int Ny, long w, long W, F &&f) {
    const long nst = nstrips_of<HALO>(Nx);
    const long S = W >= nst ? W / nst : 1;
    const long U = W < S * nst ? W : S * nst;
    for (long k = w; k < S * nst; k += U) {
        const long seg = k / nst;
        const int ja = (int)((seg * Ny) / S), jb = (int)(((seg + 1) * Ny) / S);
        if (ja < jb) f((int)(k % nst), ja, jb);
    }
}

// ---- single-sweep kernels ---------------------------------------------------

template <typename R>
__global__ void __launch_bounds__(256) k_psi_sweep(PsiSweepArgs<R> A, int *r2max, long W) {
    __shared__ int lds[4];
    const long w = (long)blockIdx.x * (blockDim.x / WAVE) + wave_id();
    R rr = R(0);
    if (w < W) for_rows(A.Nx, A.Ny, w, W, [&](int s, int ja, int jb) { rr = fmax(rr, psi_rows<R, false, true>(A, s, ja, jb)); });
    const int rmax = block_max_int(scaled_residual<R>(rr, A.stop_eps), lds);
    if (threadIdx.x == 0) atomic_max_agent(r2max, rmax);
}

template <typename R>
__global__ void __launch_bounds__(256) k_a_sweep(ASweepArgs<R> A, int *r2max, long W) {
    __shared__ int lds[4];
    const long w = (long)blockIdx.x * (blockDim.x / WAVE) + wave_id();
    R rr = R(0);
    if (w < W) for_rows(A.Nx, A.Ny, w, W, [&](int s, int ja, int jb) { rr = fmax(rr, a_rows<R, false, true>(A, s, ja, jb)); });
    const int rmax = block_max_int(scaled_residual<R>(rr, A.stop_eps), lds);
    if (threadIdx.x == 0) atomic_max_agent(r2max, rmax);
}

// ---- persistent solve kernels ----------------------------------------------
//
// One launch runs one Jacobi solve (psi or A) to the reference's stop rule:
// every sweep ends at a grid barrier that also reduces the sweep's residual,
// so every workgroup sees the same value and takes the same branch.
// Per-step launches are enqueued back to back; the buffer holding the
// current state and the barrier epoch carry over between launches in device
// memory, so there is no host round trip per step.

enum { ST_PSI = 0, ST_AB = 1, ST_EPOCH = 2, ST_STEP = 4, ST_PRED_A = 5, ST_ERRSTEP = 6, ST_PRED_PSI = 7 };
// Pass counters of a single-instance run (glx_td_passes): memory passes of
// the psi and A solves, the redone ones among them, and the A passes that
// move one more A buffer (the three-sweep pass stores A^2; an odd sweep
// j >= 3, and every sweep with a fixed source, reads a separate source).
enum { ST_PASS_PSI = 8, ST_REDO_PSI = 9, ST_PASS_A = 10, ST_REDO_A = 11, ST_XBUF_A = 12, ST_WORDS = 13 };

struct TDScratch {
    GridSync gs;
    int *state;             // ST_* fields (64-bit epoch at ST_EPOCH)
    int *sweeps;            // 2 * nsteps
};

template <typename R> struct TDSolveArgs {
    PsiSweepArgs<R> psi;    // per-sweep pointers are filled in the kernel
    ASweepArgs<R> A;
    cplx<R> *P[3];
    R *Q[3];
    cplx<R> *rhs_psi;       // Langevin rhs scratch or NULL
    R *rhs_ab;
    int max_sweeps;
    int step;               // index into sweeps[]
    int solveA;             // batched runs: the instance has an A solve
    TDScratch sc;
    const R *a_src;         // fixed supercurrent source of every A sweep (k_td_a_solve<R, G, true>)
};

// P[k] for a run-time k without indexing the kernel-argument array (which
// would copy it to scratch memory).
template <typename T> __device__ __forceinline__ T *pick3(T *const (&P)[3], int k) {
    return k == 0 ? P[0] : (k == 1 ? P[1] : P[2]);
}

__device__ __forceinline__ unsigned long long *state64(int *state, int k) {
    return reinterpret_cast<unsigned long long *>(state + k);
}

// The scaled residuals of up to three sweeps of one pass go through the team
// barrier as 10-bit fields min(r, 16383) >> 4 (only the comparison with
// 10000 steers the solve, and r < 10000 iff the field is below 625 =
// 10000 / 16). sweep_sync: the workgroup's largest residuals, team barrier,
// team-wide fields in r_out; res_conv(r_out, k): sweep k of the pass met the
// stop rule everywhere. Four sweeps go as 8-bit fields: r >> 6 (at most
// 156) below 10000, at least 157 from 10000 on.
__device__ __forceinline__ int res_field(int r) { return min(r, 16383) >> 4; }
__device__ __forceinline__ bool res_conv(int packed, int k) { return ((packed >> (10 * k)) & 0x3ff) < 625; }
__device__ __forceinline__ int res_field8(int r) { return r < 10000 ? (r >> 6) : max(157, min(r, 16383) >> 6); }
__device__ __forceinline__ bool res_conv8(int packed, int k) { return ((packed >> (8 * k)) & 0xff) < 157; }

__device__ __forceinline__ bool sweep_sync4(const GridSync &gs, const Team &tm, const int (&r)[4],
                                            unsigned long long &epoch, int *lds, int &r_out) {
    int packed = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) packed |= block_max_int(res_field8(r[k]), lds) << (8 * k);
    return team_sync_max<4>(gs, tm, epoch, packed, r_out);
}

template <int K = 3>
__device__ __forceinline__ bool sweep_sync(const GridSync &gs, const Team &tm, int r1, int r2, int r3,
                                           unsigned long long &epoch, int *lds, int &r_out) {
    r1 = K >= 1 ? block_max_int(res_field(r1), lds) : 0;  // K: fields carried (passes of at most K sweeps)
    r2 = K >= 2 ? block_max_int(res_field(r2), lds) : 0;
    r3 = K >= 3 ? block_max_int(res_field(r3), lds) : 0;
    return team_sync_max(gs, tm, epoch, (r3 << 20) | (r2 << 10) | r1, r_out);
}

// Result of one solve, identical in every workgroup of the team.
struct SolveOut {
    int cur;    // buffer index of the final iterate
    int nsw;    // sweeps
    bool ok;    // false: a barrier timed out
    int cnt;    // pass counters, packed (one register): PASS_ONE per memory
                // pass (a fused pass counts once), REDO_ONE per pass redone
                // up to an earlier sweep that met the stop rule, XBUF_ONE per
                // A pass that moves one more A buffer
};
// a solve has at most 1024 sweeps, so each field fits its bits
constexpr int PASS_ONE = 1, REDO_ONE = 1 << 11, XBUF_ONE = 1 << 13;
__host__ __device__ inline int cnt_passes(int c) { return c & 0x7ff; }
__host__ __device__ inline int cnt_redo(int c) { return (c >> 11) & 0x3; }
__host__ __device__ inline int cnt_xbuf(int c) { return (c >> 13) & 0x7ff; }

// The psi solve of one TD step on the team's waves (w of W), in passes of
// KMAX fused sweeps (psi_rows3 or psi_rows2: several sweeps per pass over
// memory). A pass reports each sweep's residual; when an earlier sweep of the
// pass already met the stop rule, the solve is redone up to that sweep from
// the pass's input (psi_rows or psi_rows2 into the same output), so the
// result and the sweep count are those of sweep-by-sweep iteration. The
// first pass is shortened to pred mod KMAX sweeps (pred: the previous step's
// count) so that the full passes end where the solve is likely to end.
template <typename R, bool GAUGE, bool SH = false, int KB = 0>
__device__ __forceinline__ SolveOut psi_solve(const TDSolveArgs<R> &T, const GridSync &gs, const Team &tm, long w,
                                              long W, int pc, int qc, int pred, unsigned long long &epoch, int *lds) {
    // KB > 0: passes of at most KB sweeps (the batched kernel's tile depth,
    // picked by its autotuner; strips are 64 - 2 KB columns wide)
    constexpr int KMAX = KB > 0 ? KB : (SH ? 3 : Fuse<R>::PSI);
    static_assert(KMAX >= 1 && KMAX <= 4, "one to four sweeps per pass");
    const int o1 = (pc + 1) % 3, o2 = (pc + 2) % 3;
    PsiSweepArgs<R> A = T.psi;
    A.ab = pick3(T.Q, qc);
    A.rhs_out = T.rhs_psi;
    int in = pc, out = o1, j = 0, nsw = 0, cnt = 0;
    bool done = false;
    auto pass = [&](int k, R (&r)[4]) {
        cnt += PASS_ONE;
        if constexpr (KMAX == 4) {
            if (k == 4) {
                for_rows_rm<4>(A.Nx, A.Ny, w, W, [&](int s, int ja, int jb) {
                    R q[4] = {r[0], r[1], r[2], r[3]};
                    psi_rowsK<R, 4, GAUGE, SH>(A, s, ja, jb, q);
                    r[0] = q[0]; r[1] = q[1]; r[2] = q[2]; r[3] = q[3];
                });
                return;
            }
        }
        if constexpr (KMAX >= 3) {
            if (k == 3) {
                for_rows_rm<3>(A.Nx, A.Ny, w, W,
                               [&](int s, int ja, int jb) { psi_rows3<R, GAUGE, SH>(A, s, ja, jb, r[0], r[1], r[2]); });
                return;
            }
        }
        if (k == 2)
            for_rows_rm<2>(A.Nx, A.Ny, w, W, [&](int s, int ja, int jb) { psi_rows2<R, GAUGE, SH>(A, s, ja, jb, r[0], r[1]); });
        else
            for_rows_rm<1>(A.Nx, A.Ny, w, W, [&](int s, int ja, int jb) { r[0] = fmax(r[0], psi_rows<R, GAUGE, SH>(A, s, ja, jb)); });
    };
    while (!done && j < T.max_sweeps) {
        A.psi = pick3(T.P, in);
        A.psi_next = pick3(T.P, out);
        A.rhs = (T.rhs_psi && j > 0) ? T.rhs_psi : pick3(T.P, pc);
        A.jstep = (uint32_t)j;
        int k = (j == 0 && pred % KMAX != 0) ? pred % KMAX : KMAX;
        k = min(k, T.max_sweeps - j);
        R rr[4] = {R(0), R(0), R(0), R(0)};
        pass(k, rr);
        int r = 0;
        bool ok;
        if constexpr (KMAX == 4) {
            const int sr[4] = {scaled_residual<R>(rr[0], A.stop_eps), scaled_residual<R>(rr[1], A.stop_eps),
                               scaled_residual<R>(rr[2], A.stop_eps), scaled_residual<R>(rr[3], A.stop_eps)};
            ok = sweep_sync4(gs, tm, sr, epoch, lds, r);
        } else {
            ok = sweep_sync<KMAX>(gs, tm, scaled_residual<R>(rr[0], A.stop_eps), scaled_residual<R>(rr[1], A.stop_eps),
                                  scaled_residual<R>(rr[2], A.stop_eps), epoch, lds, r);
        }
        if (!ok) return {in, nsw, false, cnt};
        auto conv = [&](int q) { return KMAX == 4 ? res_conv8(r, q) : res_conv(r, q); };
        // the first sweep of the pass that met the stop rule, or k
        int m = k;
        for (int q = k - 2; q >= 0; --q)
            if (conv(q)) m = q + 1;
        if (m < k) {
            // redo the pass up to sweep m into the same output
            R d[4] = {R(0), R(0), R(0), R(0)};
            pass(m, d);
            cnt += REDO_ONE;
            int rdummy = 0;
            if (!sweep_sync<0>(gs, tm, 0, 0, 0, epoch, lds, rdummy)) return {in, nsw, false, cnt};
            done = true;
        } else {
            done = conv(k - 1);
        }
        j += m;
        nsw = j;
        in = out;
        out = (out == o1) ? o2 : o1;
    }
    return {in, nsw, true, cnt};
}

// The A solve of one TD step on the team's waves, with the reference's
// source schedule A^(2*floor(j/2)) (td.py:264-303). Sweeps run in fused
// pairs (a_rows2) starting at even j, where both sweeps take their source
// from the pair's input. When the previous step needed an odd number pred >= 3
// of sweeps, the solve opens with a three-sweep pass (a_rows3, which also
// stores A^2, the source of sweep 3). Single sweeps (a_rows) are used at odd
// j, at the last sweep allowed, and at j = pred - 1. As in psi_solve, a pass
// whose earlier sweep met the stop rule is redone up to that sweep.
// FIXED: every sweep takes its source from T.a_src (svirl's fixed-vortex
// path, td.py:256-266, binds vpi + A^n there); pairs may then start at any
// j, and the three-sweep pass (which makes its own A^2 source) is not used.
template <typename R, bool GAUGE, bool SH = false, bool FIXED = false>
__device__ __forceinline__ SolveOut a_solve(const TDSolveArgs<R> &T, const GridSync &gs, const Team &tm, long w,
                                            long W, int pc, int qc, int pred, unsigned long long &epoch, int *lds) {
    constexpr bool TRIPLE = !SH && !FIXED;
    const int o1 = (qc + 1) % 3, o2 = (qc + 2) % 3;
    ASweepArgs<R> A = T.A;
    A.psi = pick3(T.P, pc);
    A.rhs_out = T.rhs_ab;
    int in = qc, out = o1, j = 0, nsw = 0, cnt = 0;
    bool done = false;
    auto pass = [&](int k, R &r1, R &r2, R &r3) {
        cnt += PASS_ONE;
        if constexpr (TRIPLE) {
            if (k == 3) {
                for_rows_rm<3>(A.Nx, A.Ny, w, W,
                               [&](int s, int ja, int jb) { a_rows3<R, GAUGE, SH>(A, s, ja, jb, r1, r2, r3); });
                return;
            }
        }
        if (k == 2)
            for_rows_rm<2>(A.Nx, A.Ny, w, W, [&](int s, int ja, int jb) { a_rows2<R, GAUGE, SH>(A, s, ja, jb, r1, r2); });
        else
            for_rows_rm<1>(A.Nx, A.Ny, w, W, [&](int s, int ja, int jb) { r1 = fmax(r1, a_rows<R, GAUGE, SH>(A, s, ja, jb)); });
    };
    while (!done && j < T.max_sweeps) {
        A.jstep = (uint32_t)j;
        A.ab = pick3(T.Q, in);
        A.ab_next = pick3(T.Q, out);
        if constexpr (TRIPLE) A.ab_mid = pick3(T.Q, out == o1 ? o2 : o1);
        // source A^(2*floor(j/2)): the input on even sweeps, A^n on sweep 1,
        // the output buffer's previous contents on odd j >= 3 (A^(j-1), left
        // there by the single sweep j - 1 or by the three-sweep pass)
        if constexpr (FIXED) A.src = T.a_src;
        else A.src = pick3(T.Q, (j % 2 == 0) ? in : (j == 1 ? qc : out));
        A.rhs = (T.rhs_ab && j > 0) ? T.rhs_ab : pick3(T.Q, qc);
        int k = 2;
        if (!FIXED && j % 2 == 1) k = 1;
        else if (TRIPLE && j == 0 && (pred & 1) && pred >= 3 && T.max_sweeps >= 3) k = 3;
        else if (j + 1 == pred || j + 2 > T.max_sweeps) k = 1;
        R r1 = R(0), r2 = R(0), r3 = R(0);
        pass(k, r1, r2, r3);
        cnt += (FIXED || k == 3 || (j % 2 == 1 && j >= 3)) ? XBUF_ONE : 0;
        int r = 0;
        if (!sweep_sync<TRIPLE ? 3 : 2>(gs, tm, scaled_residual<R>(r1, A.stop_eps), scaled_residual<R>(r2, A.stop_eps),
                                        scaled_residual<R>(r3, A.stop_eps), epoch, lds, r))
            return {in, nsw, false, cnt};
        const int m = (k > 1 && res_conv(r, 0)) ? 1 : ((k > 2 && res_conv(r, 1)) ? 2 : k);
        if (m < k) {
            R d1 = R(0), d2 = R(0), d3 = R(0);
            pass(TRIPLE && m == 2 ? 2 : 1, d1, d2, d3);  // m is 1 unless the pass had three sweeps
            cnt += REDO_ONE + (FIXED ? XBUF_ONE : 0);
            int rdummy = 0;
            if (!sweep_sync<0>(gs, tm, 0, 0, 0, epoch, lds, rdummy)) return {in, nsw, false, cnt};
            done = true;
        } else {
            done = res_conv(r, k - 1);
        }
        j += m;
        nsw = j;
        in = out;
        out = (out == o1) ? o2 : o1;
    }
    return {in, nsw, true, cnt};
}

// Single instance: one launch per solve over the whole grid; the state
// carries the buffer indices, the barrier epoch and the sweep-count
// prediction between launches.
template <typename R, bool GAUGE>
__global__ void __launch_bounds__(256) k_td_psi_solve(TDSolveArgs<R> T) {
    __shared__ int lds[16];
    const long W = (long)gridDim.x * (blockDim.x / WAVE);
    const long w = (long)xcd_block() * (blockDim.x / WAVE) + wave_id();
    int *st = T.sc.state;
    unsigned long long epoch = *state64(st, ST_EPOCH);
    const Team tm{0, (int)gridDim.x, NREL, T.sc.gs.rel};
    const SolveOut o = psi_solve<R, GAUGE>(T, T.sc.gs, tm, w, W, st[ST_PSI], st[ST_AB], st[ST_PRED_PSI], epoch, lds);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        if (!o.ok) {
            st[ST_ERRSTEP] = T.step + 1;
            return;
        }
        st[ST_PSI] = o.cur;
        st[ST_PRED_PSI] = o.nsw;
        st[ST_PASS_PSI] += cnt_passes(o.cnt);
        st[ST_REDO_PSI] += cnt_redo(o.cnt);
        *state64(st, ST_EPOCH) = epoch;
        T.sc.sweeps[2 * T.step] = o.nsw;
    }
}

template <typename R, bool GAUGE, bool FIXED = false>
__global__ void __launch_bounds__(256) k_td_a_solve(TDSolveArgs<R> T) {
    __shared__ int lds[16];
    const long W = (long)gridDim.x * (blockDim.x / WAVE);
    const long w = (long)xcd_block() * (blockDim.x / WAVE) + wave_id();
    int *st = T.sc.state;
    unsigned long long epoch = *state64(st, ST_EPOCH);
    const Team tm{0, (int)gridDim.x, NREL, T.sc.gs.rel};
    const SolveOut o =
        a_solve<R, GAUGE, false, FIXED>(T, T.sc.gs, tm, w, W, st[ST_PSI], st[ST_AB], st[ST_PRED_A], epoch, lds);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        if (!o.ok) {
            st[ST_ERRSTEP] = T.step + 1;
            return;
        }
        st[ST_AB] = o.cur;
        st[ST_PRED_A] = o.nsw;
        st[ST_PASS_A] += cnt_passes(o.cnt);
        st[ST_REDO_A] += cnt_redo(o.cnt);
        st[ST_XBUF_A] += cnt_xbuf(o.cnt);
        *state64(st, ST_EPOCH) = epoch;
        T.sc.sweeps[2 * T.step + 1] = o.nsw;
    }
}

// ---- batched TD ---------------------------------------------------------------
//
// Many independent instances advance one TD step per launch. Each instance
// owns a team of consecutive workgroups (sized to its grid by the host) and
// runs its psi solve and then its A solve with team-local barriers, so its
// sweep counts are its own. The per-instance state (buffer indices, barrier
// epoch, step counter, sweep-count prediction) lives in device memory, so
// every step launch is identical and a run can be replayed from a captured
// HIP graph.

template <typename R> struct BatchArgs {
    const TDSolveArgs<R> *args;  // per instance
    const int *team_of_block;    // per workgroup: its instance, or -1 (idle)
    const int2 *teams;           // per instance: first workgroup, workgroups
    unsigned long long *rel;     // per instance: one release word, REL_STRIDE apart
    int *state;                  // per instance: 8 ints (ST_*)
    int *sweeps;                 // [step][instance][2]
    GridSync gs;                 // arrival words (per workgroup) and error flag
    int n;
};

// KB: sweeps per psi pass (2, 3 or 4).
template <typename R, bool GAUGE, int KB = 3>
__global__ void __launch_bounds__(256) k_batch_td_step(BatchArgs<R> B) {
    __shared__ int lds[16];
    const int inst = B.team_of_block[blockIdx.x];
    if (inst < 0) return;
    const int2 tb = B.teams[inst];
    const Team tm{tb.x, tb.y, 1, B.rel + (long)inst * REL_STRIDE};
    const long W = (long)tb.y * (blockDim.x / WAVE);
    const long w = (long)(blockIdx.x - tb.x) * (blockDim.x / WAVE) + wave_id();
    const TDSolveArgs<R> &T = B.args[inst];
    int *st = B.state + 8 * inst;
    unsigned long long epoch = *state64(st, ST_EPOCH);
    const int step = st[ST_STEP];
    const int qc = st[ST_AB];
    const SolveOut o = psi_solve<R, GAUGE, true, KB>(T, B.gs, tm, w, W, st[ST_PSI], qc, st[ST_PRED_PSI], epoch, lds);
    SolveOut oa{qc, 0, o.ok, 0};
    if (o.ok && T.solveA) oa = a_solve<R, GAUGE, true>(T, B.gs, tm, w, W, o.cur, qc, st[ST_PRED_A], epoch, lds);
    if ((int)blockIdx.x == tb.x && threadIdx.x == 0) {
        if (!oa.ok) {
            st[ST_ERRSTEP] = step + 1;
            return;
        }
        st[ST_PSI] = o.cur;
        st[ST_AB] = oa.cur;
        st[ST_PRED_PSI] = o.nsw;
        st[ST_PRED_A] = oa.nsw;
        st[ST_STEP] = step + 1;
        *state64(st, ST_EPOCH) = epoch;
        B.sweeps[((long)step * B.n + inst) * 2] = o.nsw;
        B.sweeps[((long)step * B.n + inst) * 2 + 1] = oa.nsw;
    }
}

// Zero the arrival words of the first nb workgroups and n release words, and
// set the error flag to err0 (0, or 1 to make every team leave at its first
// barrier); a kernel, so that nothing here is a memset node.
static __global__ void k_batch_reset(unsigned long long *arrive, int nb, unsigned long long *rel, int n, int *err,
                                     int err0) {
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < nb; k += gridDim.x * blockDim.x) arrive[k] = 0ull;
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) rel[(long)k * REL_STRIDE] = 0ull;
    if (blockIdx.x == 0 && threadIdx.x == 0) *err = err0;
}

// Zero the barrier words and run state before the first solve of a run, set
// the buffer indices and the error flag (err0 as in k_batch_reset).
static __global__ void k_td_reset(TDScratch sc, int pc, int qc, int err0) {
    const int t = threadIdx.x;
    for (int k = t; k < MAX_GRID_BLOCKS; k += blockDim.x) sc.gs.arrive[k] = 0ull;
    for (int k = t; k < NREL * REL_STRIDE; k += blockDim.x) sc.gs.rel[k] = 0ull;
    if (t == 0) {
        *sc.gs.err = err0;
        sc.state[ST_PSI] = pc;
        sc.state[ST_AB] = qc;
        *state64(sc.state, ST_EPOCH) = 0ull;
        sc.state[ST_ERRSTEP] = 0;
        sc.state[ST_PRED_PSI] = 0;
        sc.state[ST_PRED_A] = 0;
        for (int k = ST_PASS_PSI; k < ST_WORDS; ++k) sc.state[k] = 0;
    }
}

// A <-> A - A_g on every link, in place (sign -1: remove, +1: add back).
template <typename R>
__global__ void k_gauge_shift(R *ab, int Nx, int Ny, double H, double dx, double dy, double x0, double y0, double sign) {
    const long Na = (long)(Nx - 1) * Ny, Nab = Na + (long)Nx * (Ny - 1);
    for (long k = blockIdx.x * (long)blockDim.x + threadIdx.x; k < Nab; k += (long)gridDim.x * blockDim.x) {
        double g;
        if (k < Na) g = gauge_a(H, dy, y0, (int)(k / (Nx - 1)));
        else g = gauge_b(H, dx, x0, (int)((k - Na) % Nx));
        ab[k] = R(double(ab[k]) + sign * g);
    }
}

// k_gauge_shift on the buffer holding the run's current vector potential.
template <typename R>
__global__ void k_gauge_shift_state(R *q0, R *q1, R *q2, const int *state, int Nx, int Ny, double H, double dx,
                                    double dy, double x0, double y0, double sign) {
    const int k = state[ST_AB];
    R *ab = k == 0 ? q0 : (k == 1 ? q1 : q2);
    const long Na = (long)(Nx - 1) * Ny, Nab = Na + (long)Nx * (Ny - 1);
    for (long n = blockIdx.x * (long)blockDim.x + threadIdx.x; n < Nab; n += (long)gridDim.x * blockDim.x) {
        double g;
        if (n < Na) g = gauge_a(H, dy, y0, (int)(n / (Nx - 1)));
        else g = gauge_b(H, dx, x0, (int)((n - Na) % Nx));
        ab[n] = R(double(ab[n]) + sign * g);
    }
}

}  // namespace glx
