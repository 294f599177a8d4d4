// Batched non-linear CG (Polak-Ribiere with an exact polynomial line search,
// the algorithm of svirl/solvers/cg.py:452-551) for many small instances.
//
// The A gradient is the exact dF/dA (the reference's Jacobian doubles the
// field part on boundary links, cg.h:179-200, which keeps the gradient norm
// from ever reaching the stop test).
//
// One workgroup owns one instance at a time and runs whole CG iterations on
// it: every reduction (PR sums, gradient norm, the 17 line-search
// coefficients) is a workgroup reduction in FP64, the line search is solved
// by the first wave in FP64, and the phases are separated by workgroup barriers
// only. Instances therefore iterate and stop independently (ragged batch);
// the per-instance state (iteration, betas, initial gradient norm, stop
// flag) lives in device memory so a run is split into launches of a few
// iterations.
#pragma once

#include "glx_fields.hpp"

namespace glx {

template <typename R> struct CGInst {
    Geo g;
    R kappa2, eps, H;         // kappa2 <= 0: kappa = inf (psi only)
    const R *eps_map;         // per node or NULL
    const uint8_t *mt;        // per cell or NULL
    cplx<R> *psi;             // state
    R *ab;                    // state (NULL when kappa = inf)
    cplx<R> *jac[2], *dir[2]; // psi gradient and direction (two each, alternating by iteration)
    R *jacA[2], *dirA[2];     // A gradient and direction
};

// Per-instance state, CG_NSTATE doubles: iteration, stop flag (1
// converged, 2 stalled), beta_psi, beta_A, initial gradient norm, last
// relative gradient norm, last alphas, the iteration the run started at,
// and the stall rule's mark and count. A run starts at iteration CG_IT0
// (0, or where an earlier run stopped) with a steepest-descent step, and
// takes CG_G0 as ||grad F||_0 when it is already set (> 0).
enum {
    CG_IT = 0, CG_DONE = 1, CG_BPSI = 2, CG_BA = 3, CG_G0 = 4, CG_GREL = 5, CG_APSI = 6, CG_AA = 7,
    CG_IT0 = 8, CG_MARK = 9, CG_SINCE = 10, CG_NSTATE = 12
};

// Sum of L doubles per thread over the workgroup, into out (thread 0).
template <int L> __device__ __forceinline__ void block_sums(double (&acc)[L], double *lds, double *out) {
#pragma unroll
    for (int k = 0; k < L; ++k) {
        const double s = block_sum(acc[k], lds);
        if (threadIdx.x == 0) out[k] = s;
        __syncthreads();
    }
}

// FP32: at most 128 VGPRs, so that four workgroups share a CU and a
// 1024-instance batch is resident at once on 256 CUs (at the 164 the
// kernel would otherwise take, three fit and a launch runs in two rounds).
// FP64 keeps the default bound (capping it spills the line search).
template <typename R> struct BatchCGBound { static constexpr int MIN_WG = 1; };
template <> struct BatchCGBound<float> { static constexpr int MIN_WG = 4; };

// Stall rule (stall_iters > 0): an instance whose relative gradient norm
// has not halved within stall_iters iterations stops with flag 2 (the
// FP32-mixed form stalls at the resolution of its gradient).
template <typename R, int L>
__global__ void __launch_bounds__(256, BatchCGBound<R>::MIN_WG) k_batch_cg(const CGInst<R> *insts, double *state, int n, int iters,
                                                  int max_iters, double rtol, int stall_iters) {
    __shared__ double lds[8];
    __shared__ double red[L];
    __shared__ double bc[4];   // beta_psi, beta_A, alpha_psi, alpha_A
    __shared__ int go, s_it;
    for (int k = blockIdx.x; k < n; k += gridDim.x) {
        const CGInst<R> &I = insts[k];
        const Geo &g = I.g;
        const GeoR<R> q(g);
        const R dxdy = q.dx * q.dy, s2 = R(2) * q.dx * q.dy;
        const bool withA = L == 17;
        double *sv = state + CG_NSTATE * k;
        for (int it = 0; it < iters; ++it) {
            if (threadIdx.x == 0) {
                go = (sv[CG_DONE] == 0.0 && sv[CG_IT] < (double)max_iters) ? 1 : 0;
                s_it = (int)sv[CG_IT];
            }
            __syncthreads();
            if (!go) break;
            const int itn = s_it;
            // (selects, not indexing: a run-time index into the instance
            // record would move the whole record to scratch memory)
            const bool odd = (itn & 1) != 0;
            cplx<R> *jac = odd ? I.jac[1] : I.jac[0], *jprev = odd ? I.jac[0] : I.jac[1];
            R *jacA = odd ? I.jacA[1] : I.jacA[0], *jprevA = odd ? I.jacA[0] : I.jacA[1];
            // the direction of the previous iteration and this iteration's
            cplx<R> *dold = odd ? I.dir[1] : I.dir[0], *dnew = odd ? I.dir[0] : I.dir[1];
            R *dAold = odd ? I.dirA[1] : I.dirA[0], *dAnew = odd ? I.dirA[0] : I.dirA[1];
            // 1. gradients, PR sums and the gradient norm
            double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};  // |g|^2, num_psi, den_psi, num_A, den_A
            for (long m = threadIdx.x; m < g.N; m += blockDim.x) {
                int i, j;
                node_ij(g, m, i, j);
                R e, ga, gb;
                cplx<R> gj;
                bool ha, hb;
                grad_node<R, withA>(g, q, I.kappa2, I.eps, I.eps_map, I.H, I.mt, I.psi, (const R *)nullptr, I.ab, m, i,
                                    j, e, gj, ga, gb, ha, hb, R(0.5));
                const cplx<R> gv = {dxdy * gj.re, dxdy * gj.im};
                const cplx<R> gp = ld(jprev, m);
                st(jac, m, gv);
                acc[0] += double(gv.re * gv.re) + double(gv.im * gv.im);
                acc[1] += double(gv.re * (gv.re - gp.re)) + double(gv.im * (gv.im - gp.im));
                acc[2] += double(gp.re * gp.re) + double(gp.im * gp.im);
                if (withA) {
                    if (ha) {
                        const long la = i + (long)g.Nxa * j;
                        const R v = s2 * ga, vp = jprevA[la];
                        jacA[la] = v;
                        acc[0] += double(v * v);
                        acc[3] += double(v * (v - vp));
                        acc[4] += double(vp * vp);
                    }
                    if (hb) {
                        const long lb = g.Na + i + (long)g.Nxb * j;
                        const R v = s2 * gb, vp = jprevA[lb];
                        jacA[lb] = v;
                        acc[0] += double(v * v);
                        acc[3] += double(v * (v - vp));
                        acc[4] += double(vp * vp);
                    }
                }
            }
            block_sums<5>(acc, lds, red);
            if (threadIdx.x == 0) {
                const double gn = sqrt(red[0]);
                const int it0 = (int)sv[CG_IT0];
                if (itn == it0 && !(sv[CG_G0] > 0.0)) sv[CG_G0] = gn;
                const double g0 = sv[CG_G0];
                const double grel = g0 > 0.0 ? gn / g0 : 0.0;
                sv[CG_GREL] = grel;
                if (!(gn > rtol * g0)) {
                    sv[CG_DONE] = 1.0;
                    go = 0;
                } else if (stall_iters > 0 && itn > it0 && !(grel < 0.5 * sv[CG_MARK]) &&
                           sv[CG_SINCE] + 1.0 >= (double)stall_iters) {
                    sv[CG_DONE] = 2.0;
                    go = 0;
                } else {
                    if (itn == it0 || grel < 0.5 * sv[CG_MARK]) {
                        sv[CG_MARK] = grel;
                        sv[CG_SINCE] = 0.0;
                    } else {
                        sv[CG_SINCE] += 1.0;
                    }
                    if (itn > it0) {
                        sv[CG_BPSI] = double(R(fmax(red[1] / red[2], 0.0)));
                        sv[CG_BA] = withA ? double(R(fmax(red[3] / red[4], 0.0))) : 0.0;
                    }
                    bc[0] = sv[CG_BPSI];
                    bc[1] = sv[CG_BA];
                }
            }
            __syncthreads();
            if (!go) break;
            // 2. line-search coefficients of the new direction d = beta d_old
            // - g, which the pass forms where it reads it and writes to dnew
            const DirUpdate<R> U{dold, jac, dAold, jacA, dnew, dAnew, R(bc[0]), R(bc[1])};
            double cacc[L];
#pragma unroll
            for (int c = 0; c < L; ++c) cacc[c] = 0.0;
            double th = 0.0;  // largest link phase step |d dA| (the line search's bound)
            for (long m = threadIdx.x; m < g.N; m += blockDim.x) {
                R C[L];
#pragma unroll
                for (int c = 0; c < L; ++c) C[c] = R(0);
                coef_node<R, L, true>(g, q, withA ? I.kappa2 : R(-1), I.eps, I.eps_map, I.H, I.mt, I.psi, nullptr,
                                      (const R *)nullptr, I.ab, nullptr, m, C, &U, withA ? &th : nullptr);
#pragma unroll
                for (int c = 0; c < L; ++c) cacc[c] += double(C[c]);
            }
            block_sums<L>(cacc, lds, red);
            th = block_max(th, lds);
            th = __shfl(th, 0, WAVE);  // block_max leaves it in thread 0
            if (threadIdx.x < WAVE) {  // the first wave runs the line search
                double c5[5][5] = {};
                const double sc = double(q.dx) * double(q.dy);
                if (withA) {
                    for (int j = 0; j < 5; ++j) {
                        c5[0][j] = sc * red[j];
                        c5[1][j] = sc * red[5 + j];
                        c5[2][j] = sc * red[10 + j];
                    }
                    c5[3][0] = sc * red[15 % L];
                    c5[4][0] = sc * red[16 % L];
                } else {
                    for (int i = 0; i < 5; ++i) c5[i][0] = sc * red[i];
                }
                double x, y;
                quartic2_min_wave(c5, !withA, x, y, alpha_A_bound(th));
                if (threadIdx.x == 0) {
                    sv[CG_APSI] = x;
                    sv[CG_AA] = y;
                    bc[2] = x;
                    bc[3] = y;
                }
            }
            __syncthreads();
            // 3. update
            const R apsi = R(bc[2]), aA = R(bc[3]);
            for (long m = threadIdx.x; m < g.N; m += blockDim.x) {
                const cplx<R> p = ld(I.psi, m), d = ld(dnew, m);
                st(I.psi, m, cplx<R>{apsi * d.re + p.re, apsi * d.im + p.im});
            }
            if (withA)
                for (long m = threadIdx.x; m < g.Nab; m += blockDim.x) I.ab[m] = aA * dAnew[m] + I.ab[m];
            __syncthreads();
            if (threadIdx.x == 0) sv[CG_IT] = double(itn + 1);
            __syncthreads();
        }
        // At both breaks above each wave reads go after a barrier and leaves;
        // without this barrier thread 0 could rewrite go / s_it for the next
        // instance while a slower wave has yet to read them.
        __syncthreads();
    }
}

}  // namespace glx
