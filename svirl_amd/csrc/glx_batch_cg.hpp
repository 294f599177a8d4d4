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
// by thread 0 in FP64, and the phases are separated by workgroup barriers
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
    cplx<R> *jac[2], *dir;    // psi gradient (two, alternating) and direction
    R *jacA[2], *dirA;        // A gradient and direction
};

// Per-instance state, 8 doubles: iteration, stop flag, beta_psi, beta_A,
// initial gradient norm, last relative gradient norm, last alphas.
enum { CG_IT = 0, CG_DONE = 1, CG_BPSI = 2, CG_BA = 3, CG_G0 = 4, CG_GREL = 5, CG_APSI = 6, CG_AA = 7 };

// P(x, y) = sum c[i][j] x^i y^j over i, j < 5 and its derivatives.
__device__ __forceinline__ void poly2(const double (&c)[5][5], double x, double y, double &p, double &px, double &py,
                                      double &pxx, double &pxy, double &pyy) {
    double xi[5] = {1.0, x, x * x, x * x * x, x * x * x * x};
    double yj[5] = {1.0, y, y * y, y * y * y, y * y * y * y};
    p = px = py = pxx = pxy = pyy = 0.0;
    for (int i = 0; i < 5; ++i)
        for (int j = 0; j < 5; ++j) {
            const double cij = c[i][j];
            if (cij == 0.0) continue;
            p += cij * xi[i] * yj[j];
            if (i >= 1) px += cij * i * xi[i - 1] * yj[j];
            if (j >= 1) py += cij * j * xi[i] * yj[j - 1];
            if (i >= 2) pxx += cij * i * (i - 1) * xi[i - 2] * yj[j];
            if (i >= 1 && j >= 1) pxy += cij * i * j * xi[i - 1] * yj[j - 1];
            if (j >= 2) pyy += cij * j * (j - 1) * xi[i] * yj[j - 2];
        }
}

// A local minimiser of P reached from (0, 0) by damped Newton steps
// (gradient steps where the Hessian is not positive definite) with
// Armijo backtracking, in FP64. With one_d, y stays 0 (the psi-only
// quartic). Plays the part of the reference's numpy root / scipy BFGS
// line search (cg.py:227-235, 378-419) on the device.
__device__ __forceinline__ void quartic2_min(const double (&c)[5][5], bool one_d, double &x, double &y) {
    x = 0.0;
    y = 0.0;
    for (int it = 0; it < 80; ++it) {
        double p, px, py, pxx, pxy, pyy;
        poly2(c, x, y, p, px, py, pxx, pxy, pyy);
        if (one_d) { py = 0.0; pxy = 0.0; pyy = 1.0; }
        const double gn = fabs(px) + fabs(py);
        if (!(gn > 0.0)) break;
        double dx, dy;
        const double det = pxx * pyy - pxy * pxy;
        if (pxx > 0.0 && det > 0.0) {
            dx = -(pyy * px - pxy * py) / det;
            dy = -(pxx * py - pxy * px) / det;
        } else {
            dx = -px;
            dy = -py;
        }
        const double slope = px * dx + py * dy;
        if (!(slope < 0.0)) { dx = -px; dy = -py; }
        const double sl = px * dx + py * dy;
        double t = 1.0, pn = 0.0, a, b, c2, d, e2;
        for (int k = 0; k < 60; ++k) {
            poly2(c, x + t * dx, y + t * dy, pn, a, b, c2, d, e2);
            if (pn <= p + 1.0e-4 * t * sl) break;
            t *= 0.5;
        }
        const double sx = t * dx, sy = t * dy;
        x += sx;
        y += sy;
        if (fabs(sx) + fabs(sy) <= 1.0e-15 * (fabs(x) + fabs(y)) || !(pn < p)) break;
    }
}

// Sum of L doubles per thread over the workgroup, into out (thread 0).
template <int L> __device__ __forceinline__ void block_sums(double (&acc)[L], double *lds, double *out) {
    for (int k = 0; k < L; ++k) {
        const double s = block_sum(acc[k], lds);
        if (threadIdx.x == 0) out[k] = s;
        __syncthreads();
    }
}

template <typename R, int L>
__global__ void __launch_bounds__(256) k_batch_cg(const CGInst<R> *insts, double *state, int n, int iters,
                                                  int max_iters, double rtol) {
    __shared__ double lds[8];
    __shared__ double red[L];
    __shared__ double bc[4];   // beta_psi, beta_A, alpha_psi, alpha_A
    __shared__ int go, s_it;
    for (int k = blockIdx.x; k < n; k += gridDim.x) {
        const CGInst<R> I = insts[k];
        const Geo &g = I.g;
        const GeoR<R> q(g);
        const R dxdy = q.dx * q.dy, s2 = R(2) * q.dx * q.dy;
        const bool withA = L == 17;
        double *sv = state + 8 * k;
        for (int it = 0; it < iters; ++it) {
            if (threadIdx.x == 0) {
                go = (sv[CG_DONE] == 0.0 && sv[CG_IT] < (double)max_iters) ? 1 : 0;
                s_it = (int)sv[CG_IT];
            }
            __syncthreads();
            if (!go) break;
            const int itn = s_it;
            cplx<R> *jac = I.jac[itn & 1], *jprev = I.jac[(itn + 1) & 1];
            R *jacA = I.jacA[itn & 1], *jprevA = I.jacA[(itn + 1) & 1];
            // 1. gradients, PR sums and the gradient norm
            double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};  // |g|^2, num_psi, den_psi, num_A, den_A
            for (long m = threadIdx.x; m < g.N; m += blockDim.x) {
                const cplx<R> gj = jac_psi_node(g, q, I.eps, I.eps_map, I.mt, I.psi, (const R *)nullptr, I.ab, m);
                const cplx<R> gv = {dxdy * gj.re, dxdy * gj.im};
                const cplx<R> gp = ld(jprev, m);
                st(jac, m, gv);
                acc[0] += double(gv.re * gv.re) + double(gv.im * gv.im);
                acc[1] += double(gv.re * (gv.re - gp.re)) + double(gv.im * (gv.im - gp.im));
                acc[2] += double(gp.re * gp.re) + double(gp.im * gp.im);
                if (withA) {
                    R ga, gb;
                    bool ha, hb;
                    jac_A_node(g, q, I.kappa2, I.H, I.mt, I.psi, (const R *)nullptr, I.ab, m, ga, gb, ha, hb,
                               R(0.5));
                    int i, j;
                    node_ij(g, m, i, j);
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
                if (itn == 0) sv[CG_G0] = gn;
                const double g0 = sv[CG_G0];
                sv[CG_GREL] = g0 > 0.0 ? gn / g0 : 0.0;
                if (!(gn > rtol * g0)) {
                    sv[CG_DONE] = 1.0;
                    go = 0;
                } else {
                    if (itn > 0) {
                        sv[CG_BPSI] = double(R(fmax(red[1] / red[2], 0.0)));
                        sv[CG_BA] = withA ? double(R(fmax(red[3] / red[4], 0.0))) : 0.0;
                    }
                    bc[0] = sv[CG_BPSI];
                    bc[1] = sv[CG_BA];
                }
            }
            __syncthreads();
            if (!go) break;
            // 2. directions
            const R bpsi = R(bc[0]), bA = R(bc[1]);
            for (long m = threadIdx.x; m < g.N; m += blockDim.x) {
                const cplx<R> d = ld(I.dir, m), gv = ld(jac, m);
                st(I.dir, m, cplx<R>{bpsi * d.re - gv.re, bpsi * d.im - gv.im});
            }
            if (withA)
                for (long m = threadIdx.x; m < g.Nab; m += blockDim.x) I.dirA[m] = bA * I.dirA[m] - jacA[m];
            __syncthreads();
            // 3. line-search coefficients
            double cacc[L];
#pragma unroll
            for (int c = 0; c < L; ++c) cacc[c] = 0.0;
            for (long m = threadIdx.x; m < g.N; m += blockDim.x) {
                R C[L];
#pragma unroll
                for (int c = 0; c < L; ++c) C[c] = R(0);
                coef_node<R, L>(g, q, withA ? I.kappa2 : R(-1), I.eps, I.eps_map, I.H, I.mt, I.psi, I.dir,
                                (const R *)nullptr, I.ab, I.dirA, m, C);
#pragma unroll
                for (int c = 0; c < L; ++c) cacc[c] += double(C[c]);
            }
            block_sums<L>(cacc, lds, red);
            if (threadIdx.x == 0) {
                double c5[5][5] = {};
                const double sc = double(q.dx) * double(q.dy);
                if (withA) {
                    for (int j = 0; j < 5; ++j) {
                        c5[0][j] = sc * red[j];
                        c5[1][j] = sc * red[5 + j];
                        c5[2][j] = sc * red[10 + j];
                    }
                    c5[3][0] = sc * red[15];
                    c5[4][0] = sc * red[16];
                } else {
                    for (int i = 0; i < 5; ++i) c5[i][0] = sc * red[i];
                }
                double x, y;
                quartic2_min(c5, !withA, x, y);
                sv[CG_APSI] = x;
                sv[CG_AA] = y;
                bc[2] = x;
                bc[3] = y;
            }
            __syncthreads();
            // 4. update
            const R apsi = R(bc[2]), aA = R(bc[3]);
            for (long m = threadIdx.x; m < g.N; m += blockDim.x) {
                const cplx<R> p = ld(I.psi, m), d = ld(I.dir, m);
                st(I.psi, m, cplx<R>{apsi * d.re + p.re, apsi * d.im + p.im});
            }
            if (withA)
                for (long m = threadIdx.x; m < g.Nab; m += blockDim.x) I.ab[m] = aA * I.dirA[m] + I.ab[m];
            __syncthreads();
            if (threadIdx.x == 0) sv[CG_IT] = double(itn + 1);
            __syncthreads();
        }
    }
}

}  // namespace glx
