// TDGL Jacobi sweeps for psi and A on gfx950.
//
// Layout of the work: the grid is cut into vertical strips of 63 columns;
// one wave marches down a range of rows of one strip. Lane 0 of a wave holds
// the column left of the strip (halo), lanes 1..63 own one column each.
// Each lane keeps a three-row window of its column in registers, so every
// field value is read from HBM once per sweep; horizontal neighbours come
// from adjacent lanes by cross-lane shuffles, and only lane 63 loads the
// column right of the strip. Each lane evaluates sin/cos once per link it
// carries (2 per node per sweep) and hands the values to its right
// neighbour.
//
// Arithmetic follows svirl/cuda/td.h term by term (same operand order, no
// reassociation); FP64 instantiations are compiled without FMA contraction.
#pragma once

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
    const uint8_t *mt;
    const cplx<R> *psi;
    const R *src;            // supercurrent source (may alias ab_next)
    const R *rhs;            // right-hand side read
    R *rhs_out;              // Langevin: noisy rhs written here at jstep 0 (else NULL)
    const R *ab;             // A^j
    R *ab_next;              // A^{j+1}
    R langevin_c;
    uint32_t jstep, rand_t;
    double stop_eps;
    double gH, gdx, gdy, gx0, gy0;  // as in PsiSweepArgs
};

// Symmetric gauge of field H (the one svirl builds, params.py:159-167) on
// link a(c, j) and b(c, j): a_g = -H/2 (y_j - Ly/2), b_g = H/2 (x_c - Lx/2).
__device__ __forceinline__ double gauge_a(double H, double dy, double y0, int j) { return -0.5 * H * (double(j) * dy - y0); }
__device__ __forceinline__ double gauge_b(double H, double dx, double x0, int c) { return 0.5 * H * (double(c) * dx - x0); }

// Link phase d*A of a stored value v: plain, or with the gauge added back.
template <typename R, bool GAUGE>
__device__ __forceinline__ R link_phase(R d, R v, double g) {
    if (GAUGE) return R(double(d) * (g + double(v)));
    return d * v;
}

template <typename R> __device__ __forceinline__ R shfl_up1(R v) { return __shfl_up(v, 1, WAVE); }
template <typename R> __device__ __forceinline__ R shfl_dn1(R v) { return __shfl_down(v, 1, WAVE); }

__device__ __forceinline__ bool cell_at(const uint8_t *mt, int Nxc, int Nyc, int c, int j) {
    if (c < 0 || c >= Nxc || j < 0 || j >= Nyc) return false;
    return mt ? (mt[c + (long)Nxc * j] != 0) : true;
}

// Loads of one row for the psi sweep: psi at row j+1, a and b at row j,
// the cell (c, j), the right-hand side and epsilon at row j, and (lane 63)
// psi right of the strip at row j. The sweep issues the loads of row j+1
// before it computes row j, so one row of loads per wave is in flight
// during the arithmetic.
template <typename R> struct PsiRow {
    cplx<R> pp, rhs, pr;
    R av, bv, e;
    bool cell;
};

template <typename R>
__device__ __forceinline__ PsiRow<R> psi_load_row(const PsiSweepArgs<R> &A, int c, int lane, bool col, bool own, int j) {
    const int Nx = A.Nx, Ny = A.Ny, Nxa = Nx - 1;
    const long rowoff = (long)Nx * j;
    PsiRow<R> L;
    L.pp = {R(0), R(0)};
    L.rhs = {R(0), R(0)};
    L.pr = {R(0), R(0)};
    L.av = R(0);
    L.bv = R(0);
    L.e = A.eps;
    if (col) {
        if (j + 1 < Ny) L.pp = ld(A.psi, c + rowoff + Nx);
        if (c < Nxa) L.av = A.ab[c + (long)Nxa * j];
        if (j < Ny - 1) L.bv = A.ab[(long)Nxa * Ny + c + rowoff];
    }
    L.cell = cell_at(A.mt, Nx - 1, Ny - 1, c, j);
    if (own) {
        L.rhs = ld(A.rhs, c + rowoff);
        if (A.eps_map) L.e = A.eps_map[c + rowoff];
    }
    if (lane == WAVE - 1 && c + 1 < Nx) L.pr = ld(A.psi, c + 1 + rowoff);
    return L;
}

// One psi sweep over rows [ja, jb) of strip s. Returns the lane's largest
// scaled residual.
template <typename R, bool GAUGE>
__device__ int psi_rows(const PsiSweepArgs<R> &A, int s, int ja, int jb) {
    const int lane = threadIdx.x & (WAVE - 1);
    const int Nx = A.Nx, Ny = A.Ny, Nxa = Nx - 1, Nxc = Nx - 1, Nyc = Ny - 1;
    const long Na = (long)Nxa * Ny;
    const int c = s * STRIP - 1 + lane;
    const bool col = (c >= 0) && (c < Nx);
    const bool own = col && lane >= 1;
    const R *b = A.ab + Na;
    const bool noisy = A.rhs_out != nullptr;
    int rmax = 0;

    // window: psi at rows j-1, j of this column; U of b(c, j-1); cell (c, j-1)
    cplx<R> pm = {R(0), R(0)}, pc = {R(0), R(0)};
    R cbm = R(1), sbm = R(0);
    if (col) {
        pc = ld(A.psi, c + (long)Nx * ja);
        if (ja > 0) {
            pm = ld(A.psi, c + (long)Nx * (ja - 1));
            sincos_r(link_phase<R, GAUGE>(A.dy, b[c + (long)Nx * (ja - 1)], gauge_b(A.gH, A.gdx, A.gx0, c)), &sbm, &cbm);
        }
    }
    bool cellm = cell_at(A.mt, Nxc, Nyc, c, ja - 1);
    PsiRow<R> nx = psi_load_row(A, c, lane, col, own, ja);

    for (int j = ja; j < jb; ++j) {
        const long rowoff = (long)Nx * j;
        const PsiRow<R> cu = nx;
        if (j + 1 < jb) nx = psi_load_row(A, c, lane, col, own, j + 1);
        const cplx<R> pp = cu.pp;
        const bool cellc = cu.cell;
        cplx<R> rhs = cu.rhs;
        const R e = cu.e;
        R sa, ca, sb, cb;
        sincos_r(link_phase<R, GAUGE>(A.dx, cu.av, gauge_a(A.gH, A.gdy, A.gy0, j)), &sa, &ca);
        sincos_r(link_phase<R, GAUGE>(A.dy, cu.bv, gauge_b(A.gH, A.gdx, A.gx0, c)), &sb, &cb);

        // neighbours from adjacent lanes
        const R pl_re = shfl_up1(pc.re), pl_im = shfl_up1(pc.im);
        R pr_re = shfl_dn1(pc.re), pr_im = shfl_dn1(pc.im);
        const R cal = shfl_up1(ca), sal = shfl_up1(sa);
        const bool cl_m = shfl_up1((int)cellm) != 0;  // cell (c-1, j-1)
        const bool cl_c = shfl_up1((int)cellc) != 0;  // cell (c-1, j)
        if (lane == WAVE - 1) {
            pr_re = cu.pr.re;
            pr_im = cu.pr.im;
        }

        if (own) {
            const bool mm = cl_m, mp = cl_c, pmc = cellm, ppc = cellc;
            const bool inside = mm || mp || pmc || ppc;
            cplx<R> nxt = {R(0), R(0)};
            if (inside) {
                if (noisy && A.jstep == 0 && A.langevin_c > R(1.0e-32)) {
                    const uint32_t n = (uint32_t)(c + rowoff);
                    const R n1 = R(double(noise_1<R>(n, A.rand_t)) - 0.5);
                    const R n2 = R(double(noise_2<R>(n, A.rand_t)) - 0.5);
                    rhs.re = rhs.re + A.langevin_c * n1;
                    rhs.im = rhs.im + A.langevin_c * n2;
                }
                const R rm0 = (mm || mp) ? R(1) : R(0);
                const R rp0 = (pmc || ppc) ? R(1) : R(0);
                const R r0m = (mm || pmc) ? R(1) : R(0);
                const R r0p = (mp || ppc) ? R(1) : R(0);
                // U(-dx a_left) psi_left, U(dx a) psi_right, U(-dy b_down) psi_down, U(dy b) psi_up
                R m0r = R(0), m0i = R(0), p0r = R(0), p0i = R(0);
                R q0mr = R(0), q0mi = R(0), q0pr = R(0), q0pi = R(0);
                if (mm || mp) { m0r = cal * pl_re - sal * pl_im; m0i = cal * pl_im + sal * pl_re; }
                if (pmc || ppc) { p0r = ca * pr_re + sa * pr_im; p0i = ca * pr_im - sa * pr_re; }
                if (mm || pmc) { q0mr = cbm * pm.re - sbm * pm.im; q0mi = cbm * pm.im + sbm * pm.re; }
                if (mp || ppc) { q0pr = cb * pp.re + sb * pp.im; q0pi = cb * pp.im - sb * pp.re; }
                R q = rhs.re * rhs.re + rhs.im * rhs.im - e;
                q = q + (A.idx2 * (rm0 + rp0) + A.idy2 * (r0m + r0p));
                const R diag = R(1) + A.dt * q;
                const R nre = rhs.re + A.dt * (A.idx2 * (m0r + p0r) + A.idy2 * (q0mr + q0pr));
                const R nim = rhs.im + A.dt * (A.idx2 * (m0i + p0i) + A.idy2 * (q0mi + q0pi));
                nxt.re = nre / diag;
                nxt.im = nim / diag;
            }
            if (noisy && A.jstep == 0) st(A.rhs_out, c + rowoff, rhs);
            st(A.psi_next, c + rowoff, nxt);
            const double r = fmax(fabs(double(nxt.re) - double(pc.re)), fabs(double(nxt.im) - double(pc.im)));
            rmax = max(rmax, scaled_residual(r, A.stop_eps));
        }
        // roll the window
        pm = pc;
        pc = pp;
        cbm = cb;
        sbm = sb;
        cellm = cellc;
    }
    return rmax;
}

// Loads of one row for the A sweep: a(c, j+1), b(c, j), psi(c, j+1), the
// cell (c, j), the owned links' source and right-hand side at row j, and
// (lane 63) b and psi right of the strip at row j.
template <typename R> struct ARow {
    R ap, bc, sa, sb, ra, rb, brc;
    cplx<R> pp, pr;
    bool cell;
};

template <typename R>
__device__ __forceinline__ ARow<R> a_load_row(const ASweepArgs<R> &A, int c, int lane, bool col, bool own, int j) {
    const int Nx = A.Nx, Ny = A.Ny, Nxa = Nx - 1;
    const long Na = (long)Nxa * Ny;
    const long rowoff = (long)Nx * j;
    const R *a = A.ab;
    const R *b = A.ab + Na;
    ARow<R> L;
    L.ap = R(0); L.bc = R(0); L.sa = R(0); L.sb = R(0); L.ra = R(0); L.rb = R(0); L.brc = R(0);
    L.pp = {R(0), R(0)};
    L.pr = {R(0), R(0)};
    if (col) {
        if (c < Nxa && j + 1 < Ny) L.ap = a[c + (long)Nxa * (j + 1)];
        if (j < Ny - 1) {
            L.bc = b[c + rowoff];
            L.pp = ld(A.psi, c + rowoff + Nx);
        }
    }
    L.cell = cell_at(A.mt, Nx - 1, Ny - 1, c, j);
    const long na = c + (long)Nxa * j;
    const long nb = Na + c + rowoff;
    if (own && c < Nxa) { L.sa = A.src[na]; L.ra = A.rhs[na]; }
    if (own && j < Ny - 1) { L.sb = A.src[nb]; L.rb = A.rhs[nb]; }
    if (lane == WAVE - 1 && c + 1 < Nx) {
        if (j < Ny - 1) L.brc = b[c + 1 + rowoff];
        L.pr = ld(A.psi, c + 1 + rowoff);
    }
    return L;
}

// One A sweep over rows [ja, jb) of strip s (both link families of every
// owned node).
template <typename R, bool GAUGE>
__device__ int a_rows(const ASweepArgs<R> &A, int s, int ja, int jb) {
    const int lane = threadIdx.x & (WAVE - 1);
    const int Nx = A.Nx, Ny = A.Ny, Nxa = Nx - 1, Nxc = Nx - 1, Nyc = Ny - 1;
    const long Na = (long)Nxa * Ny;
    const int c = s * STRIP - 1 + lane;
    const bool col = (c >= 0) && (c < Nx);
    const bool own = col && lane >= 1;
    const bool hasa = col && c < Nxa;
    const R *a = A.ab;
    const R *b = A.ab + Na;
    const bool noisy = A.rhs_out != nullptr;
    int rmax = 0;

    // window of this column: a(c, j-1), a(c, j), b(c, j-1), psi(c, j)
    R am = R(0), ac = R(0), bm = R(0);
    R brm = R(0);  // b(c+1, j-1) (lane 63 only; other lanes use shuffles)
    cplx<R> pc = {R(0), R(0)};
    if (col) {
        pc = ld(A.psi, c + (long)Nx * ja);
        if (hasa) ac = a[c + (long)Nxa * ja];
        if (ja > 0) {
            if (hasa) am = a[c + (long)Nxa * (ja - 1)];
            bm = b[c + (long)Nx * (ja - 1)];
        }
    }
    if (lane == WAVE - 1 && c + 1 < Nx && ja > 0) brm = b[c + 1 + (long)Nx * (ja - 1)];
    bool cellm = cell_at(A.mt, Nxc, Nyc, c, ja - 1);
    ARow<R> nx = a_load_row(A, c, lane, col, own, ja);

    for (int j = ja; j < jb; ++j) {
        const ARow<R> cu = nx;
        if (j + 1 < jb) nx = a_load_row(A, c, lane, col, own, j + 1);
        const R ap = cu.ap, bc = cu.bc;
        const cplx<R> pp = cu.pp;
        const bool cellc = cu.cell;
        // right column values
        R brc = shfl_dn1(bc), brm_s = shfl_dn1(bm);
        R pr_re = shfl_dn1(pc.re), pr_im = shfl_dn1(pc.im);
        if (lane == WAVE - 1) {
            brm_s = brm;
            brc = cu.brc;
            pr_re = cu.pr.re;
            pr_im = cu.pr.im;
            brm = brc;
        }
        // left column values
        const R alc = shfl_up1(ac), alp = shfl_up1(ap), blc = shfl_up1(bc);
        const bool cl_c = shfl_up1((int)cellc) != 0;  // cell (c-1, j)

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
        sincos_r(link_phase<R, GAUGE>(A.dx, cu.sa, gauge_a(A.gH, A.gdy, A.gy0, j)), &ssa, &csa);
        sincos_r(link_phase<R, GAUGE>(A.dy, cu.sb, gauge_b(A.gH, A.gdx, A.gx0, c)), &ssb, &csb);

        double r = 0.0;
        if (owna) {
            R rh = R(0), dd = R(1);
            if (j == 0) { rh = A.rh_a; dd = R(2); }
            else if (j + 1 == Ny) { rh = -A.rh_a; dd = R(2); }
            R jl = R(0);
            if (cellm || cellc) {  // cells (c, j-1), (c, j)
                const R x0 = pc.re, y0 = pc.im, x1 = pr_re, y1 = pr_im;
                jl = A.idx * ((x0 * y1 - y0 * x1) * csa - (x0 * x1 + y0 * y1) * ssa);
            }
            R t1 = R(0), t2 = R(0);
            if (j > 0) t1 = A.idy2 * am - A.idxy * bm + A.idxy * brm_s;
            if (j + 1 < Ny) t2 = A.idy2 * ap + A.idxy * bc - A.idxy * brc;
            const R v = (ra + A.dt_rho * (jl + rh) + A.dt_rho_kappa2 * dd * (t1 + t2)) / A.den_a;
            if (noisy && A.jstep == 0) A.rhs_out[na] = ra;
            A.ab_next[na] = v;
            r = fabs(double(v) - double(ac));
        }
        if (ownb) {
            R rh = R(0), dd = R(1);
            if (c == 0) { rh = -A.rh_b; dd = R(2); }
            else if (c + 1 == Nx) { rh = A.rh_b; dd = R(2); }
            R jl = R(0);
            if (cl_c || cellc) {  // cells (c-1, j), (c, j)
                const R x0 = pc.re, y0 = pc.im, x1 = pp.re, y1 = pp.im;
                jl = A.idy * ((x0 * y1 - y0 * x1) * csb - (x0 * x1 + y0 * y1) * ssb);
            }
            R t1 = R(0), t2 = R(0);
            if (c > 0) t1 = A.idx2 * blc - A.idxy * alc + A.idxy * alp;
            if (c + 1 < Nx) t2 = A.idx2 * brc + A.idxy * ac - A.idxy * ap;
            const R v = (rb + A.dt_rho * (jl + rh) + A.dt_rho_kappa2 * dd * (t1 + t2)) / A.den_b;
            if (noisy && A.jstep == 0) A.rhs_out[nb] = rb;
            A.ab_next[nb] = v;
            r = fmax(r, fabs(double(v) - double(bc)));
        }
        if (own) rmax = max(rmax, scaled_residual(r, A.stop_eps));
        // roll
        am = ac;
        ac = ap;
        bm = bc;
        pc = pp;
        cellm = cellc;
    }
    return rmax;
}

// ---- work partition --------------------------------------------------------
// A grid of W waves shares T = nstrips*Ny wave-rows; wave w takes the
// contiguous global rows [w*T/W, (w+1)*T/W), which may span two strips.
__host__ __device__ inline int nstrips_of(int Nx) { return (Nx + STRIP - 1) / STRIP; }

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

// ---- single-sweep kernels ---------------------------------------------------

template <typename R>
__global__ void __launch_bounds__(256) k_psi_sweep(PsiSweepArgs<R> A, int *r2max, long W) {
    __shared__ int lds[4];
    const long w = (long)blockIdx.x * (blockDim.x / WAVE) + threadIdx.x / WAVE;
    int rmax = 0;
    if (w < W) for_rows(A.Nx, A.Ny, w, W, [&](int s, int ja, int jb) { rmax = max(rmax, psi_rows<R, false>(A, s, ja, jb)); });
    rmax = block_max_int(rmax, lds);
    if (threadIdx.x == 0) atomic_max_agent(r2max, rmax);
}

template <typename R>
__global__ void __launch_bounds__(256) k_a_sweep(ASweepArgs<R> A, int *r2max, long W) {
    __shared__ int lds[4];
    const long w = (long)blockIdx.x * (blockDim.x / WAVE) + threadIdx.x / WAVE;
    int rmax = 0;
    if (w < W) for_rows(A.Nx, A.Ny, w, W, [&](int s, int ja, int jb) { rmax = max(rmax, a_rows<R, false>(A, s, ja, jb)); });
    rmax = block_max_int(rmax, lds);
    if (threadIdx.x == 0) atomic_max_agent(r2max, rmax);
}

// ---- persistent solve kernels ----------------------------------------------
//
// One launch runs one Jacobi solve (psi or A) to the reference's stop rule:
// every sweep ends at a grid barrier after which every workgroup reads the
// same residual slots and takes the same branch. Per-step launches are
// enqueued back to back; the buffer holding the current state, the barrier
// epoch and the global sweep counter carry over between launches in device
// memory, so there is no reset launch and no host round trip per step.

constexpr int SLOT_RING = 64;  // residual slots, one per sweep, NGRP ints each

enum { ST_PSI = 0, ST_AB = 1, ST_EPOCH = 2, ST_GSWEEP = 4, ST_ERRSTEP = 6 };

struct TDScratch {
    GridSync gs;
    int *slots;             // SLOT_RING * NGRP
    int *state;             // ST_* fields (64-bit epoch and sweep counters at ST_EPOCH, ST_GSWEEP)
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
    TDScratch sc;
};

__device__ __forceinline__ unsigned long long *state64(int *state, int k) {
    return reinterpret_cast<unsigned long long *>(state + k);
}

// Fold the workgroup's residual into the slot of this sweep, meet at the
// barrier, read the sweep's grid-wide residual. Block 0 clears the slot 32
// sweeps ahead (every block has finished reading it).
__device__ __forceinline__ bool sweep_sync(const TDScratch &sc, int rmax, unsigned long long &epoch,
                                           unsigned long long gsweep, int *lds, int &r_out) {
    rmax = block_max_int(rmax, lds);
    if (threadIdx.x == 0) {
        atomic_max_agent(sc.slots + (gsweep % SLOT_RING) * NGRP + blockIdx.x % NGRP, rmax);
        if (blockIdx.x == 0) {
            int *z = sc.slots + ((gsweep + SLOT_RING / 2) % SLOT_RING) * NGRP;
            for (int k = 0; k < NGRP; ++k) __hip_atomic_store(z + k, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
    if (!grid_barrier(sc.gs, epoch)) return false;
    if (threadIdx.x == 0) {
        int *sl = sc.slots + (gsweep % SLOT_RING) * NGRP;
        int m = 0;
        for (int k = 0; k < NGRP; ++k) m = max(m, ld_acquire(sl + k));
        lds[8] = m;
    }
    __syncthreads();
    r_out = lds[8];
    return true;
}

template <typename R, bool GAUGE>
__global__ void __launch_bounds__(256) k_td_psi_solve(TDSolveArgs<R> T) {
    __shared__ int lds[16];
    const long W = (long)gridDim.x * (blockDim.x / WAVE);
    const long w = (long)blockIdx.x * (blockDim.x / WAVE) + threadIdx.x / WAVE;
    int *st = T.sc.state;
    const int pc = st[ST_PSI], qc = st[ST_AB];
    const unsigned long long e0 = *state64(st, ST_EPOCH), g0 = *state64(st, ST_GSWEEP);
    unsigned long long epoch = e0, gsweep = g0;
    const int o1 = (pc + 1) % 3, o2 = (pc + 2) % 3;
    PsiSweepArgs<R> A = T.psi;
    A.ab = T.Q[qc];
    int in = pc, out = o1, nsw = 0;
    for (int j = 0; j < T.max_sweeps; ++j) {
        A.jstep = (uint32_t)j;
        A.psi = T.P[in];
        A.psi_next = T.P[out];
        A.rhs = (T.rhs_psi && j > 0) ? T.rhs_psi : T.P[pc];
        A.rhs_out = T.rhs_psi;
        int rmax = 0;
        for_rows(A.Nx, A.Ny, w, W, [&](int s, int ja, int jb) { rmax = max(rmax, psi_rows<R, GAUGE>(A, s, ja, jb)); });
        int r = 0;
        if (!sweep_sync(T.sc, rmax, epoch, gsweep, lds, r)) {
            if (blockIdx.x == 0 && threadIdx.x == 0) st[ST_ERRSTEP] = T.step + 1;
            return;
        }
        ++gsweep;
        nsw = j + 1;
        in = out;
        out = (out == o1) ? o2 : o1;
        if (r < 10000) break;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        st[ST_PSI] = in;
        *state64(st, ST_EPOCH) = epoch;
        *state64(st, ST_GSWEEP) = gsweep;
        T.sc.sweeps[2 * T.step] = nsw;
    }
}

template <typename R, bool GAUGE>
__global__ void __launch_bounds__(256) k_td_a_solve(TDSolveArgs<R> T) {
    __shared__ int lds[16];
    const long W = (long)gridDim.x * (blockDim.x / WAVE);
    const long w = (long)blockIdx.x * (blockDim.x / WAVE) + threadIdx.x / WAVE;
    int *st = T.sc.state;
    const int pc = st[ST_PSI], qc = st[ST_AB];
    const unsigned long long e0 = *state64(st, ST_EPOCH), g0 = *state64(st, ST_GSWEEP);
    unsigned long long epoch = e0, gsweep = g0;
    const int o1 = (qc + 1) % 3, o2 = (qc + 2) % 3;
    ASweepArgs<R> A = T.A;
    A.psi = T.P[pc];
    int in = qc, out = o1, nsw = 0;
    for (int j = 0; j < T.max_sweeps; ++j) {
        A.jstep = (uint32_t)j;
        A.ab = T.Q[in];
        A.ab_next = T.Q[out];
        // source A^(2*floor(j/2)): the input on even sweeps, A^n on sweep 1,
        // the output buffer's previous contents on odd j >= 3
        A.src = (j % 2 == 0) ? T.Q[in] : (j == 1 ? T.Q[qc] : T.Q[out]);
        A.rhs = (T.rhs_ab && j > 0) ? T.rhs_ab : T.Q[qc];
        A.rhs_out = T.rhs_ab;
        int rmax = 0;
        for_rows(A.Nx, A.Ny, w, W, [&](int s, int ja, int jb) { rmax = max(rmax, a_rows<R, GAUGE>(A, s, ja, jb)); });
        int r = 0;
        if (!sweep_sync(T.sc, rmax, epoch, gsweep, lds, r)) {
            if (blockIdx.x == 0 && threadIdx.x == 0) st[ST_ERRSTEP] = T.step + 1;
            return;
        }
        ++gsweep;
        nsw = j + 1;
        in = out;
        out = (out == o1) ? o2 : o1;
        if (r < 10000) break;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        st[ST_AB] = in;
        *state64(st, ST_EPOCH) = epoch;
        *state64(st, ST_GSWEEP) = gsweep;
        T.sc.sweeps[2 * T.step + 1] = nsw;
    }
}

// Zero the barrier counters, residual slots and run state (before the first
// solve of a run, or when the grid size changes) and set the buffer indices.
static __global__ void k_td_reset(TDScratch sc, int pc, int qc) {
    const int t = threadIdx.x;
    for (int k = t; k < SLOT_RING * NGRP; k += blockDim.x) sc.slots[k] = 0;
    for (int k = t; k < NGRP * GS_STRIDE; k += blockDim.x) { sc.gs.grp[k] = 0ull; sc.gs.rel[k] = 0ull; }
    if (t == 0) {
        *sc.gs.top = 0ull;
        *sc.gs.err = 0;
        sc.state[ST_PSI] = pc;
        sc.state[ST_AB] = qc;
        *state64(sc.state, ST_EPOCH) = 0ull;
        *state64(sc.state, ST_GSWEEP) = 0ull;
        sc.state[ST_ERRSTEP] = 0;
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
