// Observables, CG kernels, reductions and the vortex winding kernel.
//
// One thread per node (or link / cell), 256-thread workgroups, grid-stride
// loops. Every reduction accumulates in FP64 (also for float fields) and is
// deterministic: per-workgroup partials over a fixed grid, then one
// workgroup sums the partials in a fixed order.
//
// Per-node formulas follow svirl/cuda/observables.h and cg.h term by term.
#pragma once

#include <climits>

#include "glx_device.hpp"

namespace glx {

constexpr int RED_BLOCK = 256;
constexpr int RED_MAX_BLOCKS = 1024;

struct Geo {
    int Nx, Ny, Nxa, Nya, Nxb, Nyb, Nxc, Nyc;
    long N, Na, Nab, Nc;
    double dx, dy, idx, idy, idx2, idy2, idxy;
    double inx;  // 1 / Nx (node_ij)
};

inline Geo make_geo(int Nx, int Ny, double dx, double dy) {
    Geo g;
    g.Nx = Nx; g.Ny = Ny;
    g.Nxa = Nx - 1; g.Nya = Ny; g.Nxb = Nx; g.Nyb = Ny - 1; g.Nxc = Nx - 1; g.Nyc = Ny - 1;
    g.N = (long)Nx * Ny; g.Na = (long)g.Nxa * g.Nya; g.Nab = g.Na + (long)g.Nxb * g.Nyb;
    g.Nc = (long)g.Nxc * g.Nyc;
    g.dx = dx; g.dy = dy;
    g.idx = 1.0 / dx; g.idy = 1.0 / dy;
    g.idx2 = 1.0 / (dx * dx); g.idy2 = 1.0 / (dy * dy); g.idxy = 1.0 / (dx * dy);
    g.inx = 1.0 / Nx;
    return g;
}

// (i, j) of node n = i + Nx j without an integer division: for n < 2^30 the
// double product n * (1/Nx) is within 2^-22 of the quotient, so its
// truncation is off by at most one, which one comparison corrects.
__device__ __forceinline__ void node_ij(const Geo &g, long n, int &i, int &j) {
    const int nn = (int)n;
    int jj = (int)((double)nn * g.inx);
    int ii = nn - jj * g.Nx;
    if (ii < 0) {
        ii += g.Nx;
        --jj;
    } else if (ii >= g.Nx) {
        ii -= g.Nx;
        ++jj;
    }
    i = ii;
    j = jj;
}

// Geometry constants converted to the field precision.
template <typename R> struct GeoR {
    R dx, dy, idx, idy, idx2, idy2, idxy;
    __host__ __device__ GeoR() {}
    __host__ __device__ explicit GeoR(const Geo &g)
        : dx(R(g.dx)), dy(R(g.dy)), idx(R(g.idx)), idy(R(g.idy)), idx2(R(g.idx2)), idy2(R(g.idy2)),
          idxy(R(g.idxy)) {}
};

struct Cells { bool mm, mp, pm, pp; };

__device__ __forceinline__ Cells cells_of(const Geo &g, const uint8_t *mt, int i, int j) {
    Cells c;
    c.mm = (i > 0) && (j > 0);
    c.mp = (i > 0) && (j + 1 < g.Ny);
    c.pm = (i + 1 < g.Nx) && (j > 0);
    c.pp = (i + 1 < g.Nx) && (j + 1 < g.Ny);
    if (mt) {
        if (c.mm) c.mm = mt[(i - 1) + (long)g.Nxc * (j - 1)] != 0;
        if (c.mp) c.mp = mt[(i - 1) + (long)g.Nxc * j] != 0;
        if (c.pm) c.pm = mt[i + (long)g.Nxc * (j - 1)] != 0;
        if (c.pp) c.pp = mt[i + (long)g.Nxc * j] != 0;
    }
    return c;
}

template <typename R> __device__ __forceinline__ R pv(const R *abei, const R *ab, long k) {
    R p = R(0);
    if (abei) p += abei[k];
    if (ab) p += ab[k];
    return p;
}

// js(psi0, ph, psi1) = Im(conj psi0 U(ph) psi1) (common.h:66-73)
template <typename R> __device__ __forceinline__ R js(cplx<R> p0, R ph, cplx<R> p1) {
    R s, c;
    sincos_r(ph, &s, &c);
    return (p0.re * p1.im - p0.im * p1.re) * c - (p0.re * p1.re + p0.im * p1.im) * s;
}
// |psi1 U(ph) - psi0|^2 (observables.h:238-247)
template <typename R> __device__ __forceinline__ R g_grad(cplx<R> p0, R ph, cplx<R> p1) {
    R s, c;
    sincos_r(ph, &s, &c);
    const R re = p1.re * c + p1.im * s - p0.re;
    const R im = p1.im * c - p1.re * s - p0.im;
    return re * re + im * im;
}

// ---------------------------------------------------------------- B field
// B of cell (i, j) from abei + ab (observables.h:4-37).
template <typename R>
__device__ __forceinline__ R cell_field(const Geo &g, const GeoR<R> &q, const R *abei, const R *ab, int i, int j) {
    R Bn = R(0);
    for (int s = 0; s < 2; ++s) {
        const R *v = s == 0 ? abei : ab;
        if (!v) continue;
        Bn += q.idx * (v[g.Na + i + 1 + (long)g.Nxb * j] - v[g.Na + i + (long)g.Nxb * j]) -
              q.idy * (v[i + (long)g.Nxa * (j + 1)] - v[i + (long)g.Nxa * j]);
    }
    return Bn;
}

template <typename R>
__global__ void __launch_bounds__(256) k_magnetic_field(Geo g, const R *abei, const R *ab, R *B) {
    const GeoR<R> q(g);
    for (long n = blockIdx.x * (long)blockDim.x + threadIdx.x; n < g.Nc; n += (long)gridDim.x * blockDim.x) {
        const int i = (int)(n % g.Nxc), j = (int)(n / g.Nxc);
        B[n] = cell_field(g, q, abei, ab, i, j);
    }
}

// Ghost-cell curl-curl accumulator of one link (observables.h:66-155,
// cg.h:179-290): H source, abei and ab diagonals, lower then upper
// neighbours (abei unscaled, ab scaled by dd).
template <typename R>
__device__ __forceinline__ R curlcurl_a(const Geo &g, const GeoR<R> &q, R H, const R *abei, const R *ab, int i, int j) {
    const long Na = g.Na;
    const int Nxa = g.Nxa, Nxb = g.Nxb, Nya = g.Nya;
    R c = R(0), dd = R(1);
    if (j == 0) { c -= R(2) * H * q.idy; dd = R(2); }
    else if (j + 1 == Nya) { c += R(2) * H * q.idy; dd = R(2); }
    const long n = i + (long)Nxa * j;
    if (abei) c += R(2) / dd * q.idy2 * abei[n];
    if (ab) c += R(2) * q.idy2 * ab[n];
    if (j > 0) {
        const long jm = j - 1;
        if (abei) c += dd / dd * (-q.idy2 * abei[i + Nxa * jm] + q.idxy * abei[Na + i + Nxb * jm] - q.idxy * abei[Na + i + 1 + Nxb * jm]);
        if (ab) c += dd * (-q.idy2 * ab[i + Nxa * jm] + q.idxy * ab[Na + i + Nxb * jm] - q.idxy * ab[Na + i + 1 + Nxb * jm]);
    }
    if (j + 1 < Nya) {
        const long jp = j + 1;
        if (abei) c += dd / dd * (-q.idy2 * abei[i + Nxa * jp] - q.idxy * abei[Na + i + Nxb * (long)j] + q.idxy * abei[Na + i + 1 + Nxb * (long)j]);
        if (ab) c += dd * (-q.idy2 * ab[i + Nxa * jp] - q.idxy * ab[Na + i + Nxb * (long)j] + q.idxy * ab[Na + i + 1 + Nxb * (long)j]);
    }
    return c;
}

template <typename R>
__device__ __forceinline__ R curlcurl_b(const Geo &g, const GeoR<R> &q, R H, const R *abei, const R *ab, int i, int j) {
    const long Na = g.Na;
    const int Nxa = g.Nxa, Nxb = g.Nxb;
    R c = R(0), dd = R(1);
    if (i == 0) { c += R(2) * H * q.idx; dd = R(2); }
    else if (i + 1 == Nxb) { c -= R(2) * H * q.idx; dd = R(2); }
    const long n = Na + i + (long)Nxb * j;
    if (abei) c += R(2) / dd * q.idx2 * abei[n];
    if (ab) c += R(2) * q.idx2 * ab[n];
    if (i > 0) {
        const int im = i - 1;
        if (abei) c += dd / dd * (-q.idx2 * abei[Na + im + (long)Nxb * j] + q.idxy * abei[im + (long)Nxa * j] - q.idxy * abei[im + (long)Nxa * (j + 1)]);
        if (ab) c += dd * (-q.idx2 * ab[Na + im + (long)Nxb * j] + q.idxy * ab[im + (long)Nxa * j] - q.idxy * ab[im + (long)Nxa * (j + 1)]);
    }
    if (i + 1 < Nxb) {
        const int ip = i + 1;
        if (abei) c += dd / dd * (-q.idx2 * abei[Na + ip + (long)Nxb * j] - q.idxy * abei[i + (long)Nxa * j] + q.idxy * abei[i + (long)Nxa * (j + 1)]);
        if (ab) c += dd * (-q.idx2 * ab[Na + ip + (long)Nxb * j] - q.idxy * ab[i + (long)Nxa * j] + q.idxy * ab[i + (long)Nxa * (j + 1)]);
    }
    return c;
}

// ------------------------------------------------------ current densities
template <typename R>
__global__ void __launch_bounds__(256) k_current_density(Geo g, R kappa2, R H, const R *abei, const R *ab, R *J) {
    const GeoR<R> q(g);
    for (long n = blockIdx.x * (long)blockDim.x + threadIdx.x; n < g.N; n += (long)gridDim.x * blockDim.x) {
        int i, j;
        node_ij(g, n, i, j);
        if (i < g.Nxa && j < g.Nya) J[i + (long)g.Nxa * j] = kappa2 * curlcurl_a(g, q, H, abei, ab, i, j);
        if (i < g.Nxb && j < g.Nyb) J[g.Na + i + (long)g.Nxb * j] = kappa2 * curlcurl_b(g, q, H, abei, ab, i, j);
    }
}

// J_s on the a link (i, j) and the b link (i, j) of node n = (i, j)
// (observables.h:159-235), written where the link exists.
template <typename R>
__device__ __forceinline__ void supercurrent_links(const Geo &g, const GeoR<R> &q, const uint8_t *mt,
                                                   const cplx<R> *psi, const R *abei, const R *ab, long n, int i,
                                                   int j, R *Js) {
    const cplx<R> p0 = ld(psi, n);
    const Cells c = cells_of(g, mt, i, j);
    if (i < g.Nxa && j < g.Nya) {
        const long m = i + (long)g.Nxa * j;
        R jl = R(0);
        if (c.pm || c.pp) {
            const R r = R(0.5) * ((c.pm ? R(1) : R(0)) + (c.pp ? R(1) : R(0)));
            jl = r * q.idx * js(p0, q.dx * pv(abei, ab, m), ld(psi, n + 1));
        }
        Js[m] = jl;
    }
    if (i < g.Nxb && j < g.Nyb) {
        const long m = g.Na + i + (long)g.Nxb * j;
        R jl = R(0);
        if (c.mp || c.pp) {
            const R r = R(0.5) * ((c.mp ? R(1) : R(0)) + (c.pp ? R(1) : R(0)));
            jl = r * q.idy * js(p0, q.dy * pv(abei, ab, m), ld(psi, n + g.Nx));
        }
        Js[m] = jl;
    }
}

template <typename R>
__global__ void __launch_bounds__(256) k_supercurrent_density(Geo g, const uint8_t *mt, const cplx<R> *psi, const R *abei, const R *ab, R *Js) {
    const GeoR<R> q(g);
    for (long n = blockIdx.x * (long)blockDim.x + threadIdx.x; n < g.N; n += (long)gridDim.x * blockDim.x) {
        int i, j;
        node_ij(g, n, i, j);
        supercurrent_links(g, q, mt, psi, abei, ab, n, i, j, Js);
    }
}

// B, J and J_s in one pass over the nodes: node (i, j) writes cell (i, j)
// and its a and b links, so the link values a node's three results share
// are read once. Any of B, J, Js may be NULL (not computed). The values
// are those of k_magnetic_field, k_current_density and
// k_supercurrent_density, to the bit.
template <typename R>
__global__ void __launch_bounds__(256) k_bulk_observables(Geo g, R kappa2, R H, const uint8_t *mt, const cplx<R> *psi,
                                                          const R *abei, const R *ab, R *B, R *J, R *Js) {
    const GeoR<R> q(g);
    for (long n = blockIdx.x * (long)blockDim.x + threadIdx.x; n < g.N; n += (long)gridDim.x * blockDim.x) {
        int i, j;
        node_ij(g, n, i, j);
        if (B && i < g.Nxc && j < g.Nyc) B[i + (long)g.Nxc * j] = cell_field(g, q, abei, ab, i, j);
        if (J) {
            if (i < g.Nxa && j < g.Nya) J[i + (long)g.Nxa * j] = kappa2 * curlcurl_a(g, q, H, abei, ab, i, j);
            if (i < g.Nxb && j < g.Nyb) J[g.Na + i + (long)g.Nxb * j] = kappa2 * curlcurl_b(g, q, H, abei, ab, i, j);
        }
        if (Js) supercurrent_links(g, q, mt, psi, abei, ab, n, i, j, Js);
    }
}

// ---------------------------------------------------------- reductions

// Sums L per-thread accumulators over the workgroup and writes them to
// partials[blockIdx.x*L + k].
template <int L>
__device__ __forceinline__ void block_store_partials(double (&acc)[L], double *partials) {
    __shared__ double lds[RED_BLOCK / WAVE * L];
    const int lane = threadIdx.x & (WAVE - 1), wid = threadIdx.x / WAVE;
    const int nw = blockDim.x / WAVE;
#pragma unroll
    for (int k = 0; k < L; ++k) {
        double v = wave_sum(acc[k]);
        if (lane == 0) lds[wid * L + k] = v;
    }
    __syncthreads();
    if (threadIdx.x < L) {
        double s = 0.0;
        for (int w = 0; w < nw; ++w) s += lds[w * L + threadIdx.x];
        partials[(long)blockIdx.x * L + threadIdx.x] = s;
    }
}

// red[k] = sum_b partials[b*L + k] for k < L into shared memory, one
// workgroup: each thread loads the L values of its blocks together and the
// L sums are reduced over the workgroup at once. Ends with a barrier.
template <int L> __device__ __forceinline__ void reduce_partials(const double *partials, int nb, double *red) {
    __shared__ double lds[RED_BLOCK / WAVE * L];
    const int lane = threadIdx.x & (WAVE - 1), wid = threadIdx.x / WAVE;
    const int nw = blockDim.x / WAVE;
    double acc[L];
#pragma unroll
    for (int k = 0; k < L; ++k) acc[k] = 0.0;
    for (int b = threadIdx.x; b < nb; b += blockDim.x) {
#pragma unroll
        for (int k = 0; k < L; ++k) acc[k] += partials[(long)b * L + k];
    }
#pragma unroll
    for (int k = 0; k < L; ++k) {
        const double v = wave_sum(acc[k]);
        if (lane == 0) lds[wid * L + k] = v;
    }
    __syncthreads();
    if (threadIdx.x < L) {
        double t = 0.0;
        for (int w = 0; w < nw; ++w) t += lds[w * L + threadIdx.x];
        red[threadIdx.x] = t;
    }
    __syncthreads();
}

// out[k] = scale * sum_b partials[b*L + k], one workgroup, fixed order.
static __global__ void __launch_bounds__(256) k_finish_partials(const double *partials, int nb, int L, double scale, double *out) {
    __shared__ double lds[RED_BLOCK / WAVE];
    for (int k = 0; k < L; ++k) {
        double s = 0.0;
        for (int b = threadIdx.x; b < nb; b += blockDim.x) s += partials[(long)b * L + k];
        s = block_sum(s, lds);
        if (threadIdx.x == 0) out[k] = s * scale;
        __syncthreads();
    }
}

template <typename R>
__global__ void __launch_bounds__(256) k_sum(const R *x, long n, double *partials) {
    double acc[1] = {0.0};
    for (long k = blockIdx.x * (long)blockDim.x + threadIdx.x; k < n; k += (long)gridDim.x * blockDim.x) acc[0] += double(x[k]);
    block_store_partials<1>(acc, partials);
}

template <typename R>
__global__ void __launch_bounds__(256) k_sum_v(const R *x, long nv, int ne, double *partials) {
    // one component per launch index k; ne <= 32
    __shared__ double lds[RED_BLOCK / WAVE];
    for (int k = 0; k < ne; ++k) {
        double s = 0.0;
        for (long v = blockIdx.x * (long)blockDim.x + threadIdx.x; v < nv; v += (long)gridDim.x * blockDim.x)
            s += double(x[v * ne + k]);
        s = block_sum(s, lds);
        if (threadIdx.x == 0) partials[(long)blockIdx.x * ne + k] = s;
        __syncthreads();
    }
}

template <typename R>
__global__ void __launch_bounds__(256) k_pr_sums(const R *x, const R *y, long n, double *partials) {
    double acc[2] = {0.0, 0.0};
    for (long k = blockIdx.x * (long)blockDim.x + threadIdx.x; k < n; k += (long)gridDim.x * blockDim.x) {
        const R xv = x[k], yv = y[k];
        acc[0] += double(xv * (xv - yv));
        acc[1] += double(yv * yv);
    }
    block_store_partials<2>(acc, partials);
}

// Polak-Ribiere beta from the two partial sums of k_pr_sums, left in device
// memory (cg.py:422-449, utils.h:139-146): beta = max(num/den, 0) in FP64,
// rounded to the working precision as the reference's real_t scalar.
template <typename R>
__global__ void __launch_bounds__(256) k_pr_beta_finish(const double *partials, int nb, double *beta) {
    __shared__ double lds[RED_BLOCK / WAVE];
    double num = 0.0, den = 0.0;
    for (int b = threadIdx.x; b < nb; b += blockDim.x) num += partials[(long)b * 2];
    num = block_sum(num, lds);
    __syncthreads();
    for (int b = threadIdx.x; b < nb; b += blockDim.x) den += partials[(long)b * 2 + 1];
    den = block_sum(den, lds);
    if (threadIdx.x == 0) *beta = double(R(fmax(num / den, 0.0)));
}

// z = alpha * x - y with alpha read from device memory.
template <typename R>
__global__ void __launch_bounds__(256) k_axmy_dev(const R *x, const R *y, R *z, const double *alpha, long n) {
    const R a = R(*alpha);
    for (long k = blockIdx.x * (long)blockDim.x + threadIdx.x; k < n; k += (long)gridDim.x * blockDim.x)
        z[k] = a * x[k] - y[k];
}

template <typename R>
__global__ void __launch_bounds__(256) k_axpy(const R *x, const R *y, R *z, R alpha, long n, int minus) {
    for (long k = blockIdx.x * (long)blockDim.x + threadIdx.x; k < n; k += (long)gridDim.x * blockDim.x)
        z[k] = minus ? alpha * x[k] - y[k] : alpha * x[k] + y[k];
}

// ----------------------------------------------------------- free energy

// Per-node pseudodensity e (observables.h:250-357), before dx*dy.
template <typename R>
__device__ __forceinline__ R fe_node(const Geo &g, const GeoR<R> &q, R kappa2, R eps, const R *eps_map, R H,
                                     const uint8_t *mt, const cplx<R> *psi, const R *abei, const R *ab, long n) {
    int i, j;
    node_ij(g, n, i, j);
    R e = R(0);
    const Cells c = cells_of(g, mt, i, j);
    if (c.mm || c.mp || c.pm || c.pp) {
        const R f_mm = c.mm ? R(1) : R(0), f_mp = c.mp ? R(1) : R(0), f_pm = c.pm ? R(1) : R(0), f_pp = c.pp ? R(1) : R(0);
        const R r_m0 = R(0.5) * (f_mm + f_mp), r_p0 = R(0.5) * (f_pm + f_pp);
        const R r_0m = R(0.5) * (f_mm + f_pm), r_0p = R(0.5) * (f_mp + f_pp);
        const R gw = R(0.25) * (r_m0 + r_p0 + r_0m + r_0p);
        const cplx<R> p0 = ld(psi, n);
        const R p2 = p0.re * p0.re + p0.im * p0.im;
        const R ee = eps_map ? eps_map[n] : eps;
        e += gw * (R(0.5) * p2 - ee) * p2;
        if (c.pm || c.pp) {
            const R ph = pv(abei, ab, i + (long)g.Nxa * j);
            e += r_p0 * q.idx2 * g_grad(p0, q.dx * ph, ld(psi, n + 1));
        }
        if (c.mp || c.pp) {
            const R ph = pv(abei, ab, g.Na + i + (long)g.Nxb * j);
            e += r_0p * q.idy2 * g_grad(p0, q.dy * ph, ld(psi, n + g.Nx));
        }
    }
    if (kappa2 > R(0) && i < g.Nxc && j < g.Nyc) {
        R dB = -H;
        for (int s = 0; s < 2; ++s) {
            const R *v = s == 0 ? abei : ab;
            if (!v) continue;
            dB += q.idx * (v[g.Na + i + 1 + (long)g.Nxb * j] - v[g.Na + i + (long)g.Nxb * j]) -
                  q.idy * (v[i + (long)g.Nxa * (j + 1)] - v[i + (long)g.Nxa * j]);
        }
        e += kappa2 * dB * dB;
    }
    return e;
}

template <typename R>
__global__ void __launch_bounds__(256) k_free_energy(Geo g, R kappa2, R eps, const R *eps_map, R H, const uint8_t *mt,
                              const cplx<R> *psi, const R *abei, const R *ab, double *partials) {
    const GeoR<R> q(g);
    double acc[1] = {0.0};
    for (long n = blockIdx.x * (long)blockDim.x + threadIdx.x; n < g.N; n += (long)gridDim.x * blockDim.x)
        acc[0] += double(fe_node(g, q, kappa2, eps, eps_map, H, mt, psi, abei, ab, n));
    block_store_partials<1>(acc, partials);
}

// ---------------------------------------------------------------- CG

// dF/dpsi at node n before the dxdy factor (free_energy_jacobian_psi,
// cg.h:15-121).
template <typename R>
__device__ __forceinline__ cplx<R> jac_psi_node(const Geo &g, const GeoR<R> &q, R eps, const R *eps_map,
                                                const uint8_t *mt, const cplx<R> *psi, const R *abei, const R *ab,
                                                long n) {
    int i, j;
    node_ij(g, n, i, j);
    cplx<R> gj = {R(0), R(0)};
    const Cells c = cells_of(g, mt, i, j);
    if (c.mm || c.mp || c.pm || c.pp) {
        const R f_mm = c.mm ? R(1) : R(0), f_mp = c.mp ? R(1) : R(0), f_pm = c.pm ? R(1) : R(0), f_pp = c.pp ? R(1) : R(0);
        const R r_m0 = R(0.5) * (f_mm + f_mp), r_p0 = R(0.5) * (f_pm + f_pp);
        const R r_0m = R(0.5) * (f_mm + f_pm), r_0p = R(0.5) * (f_mp + f_pp);
        const R gw = R(0.25) * (r_m0 + r_p0 + r_0m + r_0p);
        const cplx<R> p0 = ld(psi, n);
        const R ee = eps_map ? eps_map[n] : eps;
        const R p = p0.re * p0.re + p0.im * p0.im - ee;
        const R k = R(2) * gw * p;
        gj.re += k * p0.re;
        gj.im += k * p0.im;
        auto add = [&](R r, R id2, R ph, cplx<R> p1) {
            R s, cs;
            sincos_r(ph, &s, &cs);
            const R tr = R(2) * (p0.re - (p1.re * cs + p1.im * s));
            const R ti = R(2) * (p0.im - (p1.im * cs - p1.re * s));
            const R w = r * id2;
            gj.re += w * tr;
            gj.im += w * ti;
        };
        if (c.mm || c.mp) add(r_m0, q.idx2, -q.dx * pv(abei, ab, (i - 1) + (long)g.Nxa * j), ld(psi, n - 1));
        if (c.pm || c.pp) add(r_p0, q.idx2, q.dx * pv(abei, ab, i + (long)g.Nxa * j), ld(psi, n + 1));
        if (c.mm || c.pm) add(r_0m, q.idy2, -q.dy * pv(abei, ab, g.Na + i + (long)g.Nxb * (j - 1)), ld(psi, n - g.Nx));
        if (c.mp || c.pp) add(r_0p, q.idy2, q.dy * pv(abei, ab, g.Na + i + (long)g.Nxb * j), ld(psi, n + g.Nx));
    }
    return gj;
}

template <typename R>
__global__ void __launch_bounds__(256) k_jacobian_psi(Geo g, R kappa2, R eps, const R *eps_map, R H, const uint8_t *mt,
                               const cplx<R> *psi, const R *abei, const R *ab, cplx<R> *G) {
    const GeoR<R> q(g);
    const R dxdy = q.dx * q.dy;
    for (long n = blockIdx.x * (long)blockDim.x + threadIdx.x; n < g.N; n += (long)gridDim.x * blockDim.x) {
        const cplx<R> gj = jac_psi_node(g, q, eps, eps_map, mt, psi, abei, ab, n);
        cplx<R> o = ld(G, n);
        o.re += dxdy * gj.re;
        o.im += dxdy * gj.im;
        st(G, n, o);
    }
}

// dF/dA of the a link (i, j) and the b link (i, j) at node n, before the
// 2 dxdy factor (free_energy_jacobian_A, cg.h:124-301); has_a / has_b say
// which links exist. The reference's field part on boundary links is twice
// the derivative of F (ghost cell with dd = 2); grad_node's bscale = 1/2
// gives the exact gradient there.
template <typename R>
__device__ __forceinline__ void jac_A_node(const Geo &g, const GeoR<R> &q, R kappa2, R H, const uint8_t *mt,
                                           const cplx<R> *psi, const R *abei, const R *ab, long n, R &ga, R &gb,
                                           bool &has_a, bool &has_b) {
    int i, j;
    node_ij(g, n, i, j);
    const cplx<R> p0 = ld(psi, n);
    const Cells c = cells_of(g, mt, i, j);
    has_a = i < g.Nxa && j < g.Nya;
    has_b = i < g.Nxb && j < g.Nyb;
    ga = gb = R(0);
    if (has_a) {
        const long m = i + (long)g.Nxa * j;
        R gj = curlcurl_a(g, q, H, abei, ab, i, j);
        gj *= kappa2;
        if (c.pm || c.pp) {
            const R r = R(0.5) * ((c.pm ? R(1) : R(0)) + (c.pp ? R(1) : R(0)));
            gj += -r * q.idx * js(p0, q.dx * pv(abei, ab, m), ld(psi, n + 1));
        }
        ga = gj;
    }
    if (has_b) {
        const long m = g.Na + i + (long)g.Nxb * j;
        R gj = curlcurl_b(g, q, H, abei, ab, i, j);
        gj *= kappa2;
        if (c.mp || c.pp) {
            const R r = R(0.5) * ((c.mp ? R(1) : R(0)) + (c.pp ? R(1) : R(0)));
            gj += -r * q.idy * js(p0, q.dy * pv(abei, ab, m), ld(psi, n + g.Nx));
        }
        gb = gj;
    }
}

template <typename R>
__global__ void __launch_bounds__(256) k_jacobian_A(Geo g, R kappa2, R H, const uint8_t *mt, const cplx<R> *psi, const R *abei,
                             const R *ab, R *G) {
    const GeoR<R> q(g);
    const R s2 = R(2) * q.dx * q.dy;
    for (long n = blockIdx.x * (long)blockDim.x + threadIdx.x; n < g.N; n += (long)gridDim.x * blockDim.x) {
        R ga, gb;
        bool has_a, has_b;
        jac_A_node(g, q, kappa2, H, mt, psi, abei, ab, n, ga, gb, has_a, has_b);
        int i, j;
        node_ij(g, n, i, j);
        if (has_a) G[i + (long)g.Nxa * j] += s2 * ga;
        if (has_b) G[g.Na + i + (long)g.Nxb * j] += s2 * gb;
    }
}

// The ten link values around node (i, j) that its free energy and both
// Jacobians read: a(i, j), a(i-1, j), a(i, j-1), a(i, j+1), a(i-1, j+1) and
// b(i, j), b(i, j-1), b(i-1, j), b(i+1, j), b(i+1, j-1). Loaded from clamped
// (always valid) indices, all at once; a value is used only where the
// reference's conditions say the link exists.
template <typename R> struct Links10 {
    R a00, am0, a0m, a0p, amp, b00, b0m, bm0, bp0, bpm;
};

template <typename R>
__device__ __forceinline__ Links10<R> load_links10(const Geo &g, const R *v, int i, int j) {
    const int xa0 = min(i, g.Nxa - 1), xam = max(min(i - 1, g.Nxa - 1), 0);
    const int ya0 = j, yam = max(j - 1, 0), yap = min(j + 1, g.Nya - 1);
    const int xb0 = i, xbm = max(i - 1, 0), xbp = min(i + 1, g.Nxb - 1);
    const int yb0 = min(j, g.Nyb - 1), ybm = max(min(j - 1, g.Nyb - 1), 0);
    const R *b = v + g.Na;
    Links10<R> L;
    L.a00 = v[xa0 + (long)g.Nxa * ya0];
    L.am0 = v[xam + (long)g.Nxa * ya0];
    L.a0m = v[xa0 + (long)g.Nxa * yam];
    L.a0p = v[xa0 + (long)g.Nxa * yap];
    L.amp = v[xam + (long)g.Nxa * yap];
    L.b00 = b[xb0 + (long)g.Nxb * yb0];
    L.b0m = b[xb0 + (long)g.Nxb * ybm];
    L.bm0 = b[xbm + (long)g.Nxb * yb0];
    L.bp0 = b[xbp + (long)g.Nxb * yb0];
    L.bpm = b[xbp + (long)g.Nxb * ybm];
    return L;
}

// curlcurl_a / curlcurl_b on loaded link values (E: abei, B: ab; he, hb:
// present), same terms in the same order.
template <typename R>
__device__ __forceinline__ R curlcurl_a_v(const Geo &g, const GeoR<R> &q, R H, bool he, const Links10<R> &E, bool hb,
                                          const Links10<R> &B, int j) {
    R c = R(0), dd = R(1);
    if (j == 0) { c -= R(2) * H * q.idy; dd = R(2); }
    else if (j + 1 == g.Nya) { c += R(2) * H * q.idy; dd = R(2); }
    if (he) c += R(2) / dd * q.idy2 * E.a00;
    if (hb) c += R(2) * q.idy2 * B.a00;
    if (j > 0) {
        if (he) c += dd / dd * (-q.idy2 * E.a0m + q.idxy * E.b0m - q.idxy * E.bpm);
        if (hb) c += dd * (-q.idy2 * B.a0m + q.idxy * B.b0m - q.idxy * B.bpm);
    }
    if (j + 1 < g.Nya) {
        if (he) c += dd / dd * (-q.idy2 * E.a0p - q.idxy * E.b00 + q.idxy * E.bp0);
        if (hb) c += dd * (-q.idy2 * B.a0p - q.idxy * B.b00 + q.idxy * B.bp0);
    }
    return c;
}

template <typename R>
__device__ __forceinline__ R curlcurl_b_v(const Geo &g, const GeoR<R> &q, R H, bool he, const Links10<R> &E, bool hb,
                                          const Links10<R> &B, int i) {
    R c = R(0), dd = R(1);
    if (i == 0) { c += R(2) * H * q.idx; dd = R(2); }
    else if (i + 1 == g.Nxb) { c -= R(2) * H * q.idx; dd = R(2); }
    if (he) c += R(2) / dd * q.idx2 * E.b00;
    if (hb) c += R(2) * q.idx2 * B.b00;
    if (i > 0) {
        if (he) c += dd / dd * (-q.idx2 * E.bm0 + q.idxy * E.am0 - q.idxy * E.amp);
        if (hb) c += dd * (-q.idx2 * B.bm0 + q.idxy * B.am0 - q.idxy * B.amp);
    }
    if (i + 1 < g.Nxb) {
        if (he) c += dd / dd * (-q.idx2 * E.bp0 - q.idxy * E.a00 + q.idxy * E.a0p);
        if (hb) c += dd * (-q.idx2 * B.bp0 - q.idxy * B.a00 + q.idxy * B.a0p);
    }
    return c;
}

// The free-energy density, dF/dpsi and dF/dA of node n in one evaluation:
// fe_node, jac_psi_node and jac_A_node fused. Every value the
// three read (psi and its four neighbours, the cells, epsilon, the ten link
// values of abei and ab) is loaded once, up front, from clamped indices, so
// the loads of a node go out together instead of one dependent branch at a
// time; the sin/cos of the four link phases are evaluated once. Every
// quantity keeps the operation order of the separate functions, so the
// results are theirs bit for bit. bscale multiplies the field part of dF/dA
// on boundary links: 1 keeps the reference's Jacobian (twice the derivative
// there, ghost cell with dd = 2), 1/2 gives the exact gradient.
//
// The CG state update fused into the gradient pass (StateUpdate, UPD): the
// pass evaluates the updated state s + alpha d (k_axpy_dev's arithmetic)
// wherever it reads psi or ab, and node n writes its own entries (psi(n),
// and the links a(i, j), b(i, j) it owns) to psi_out / ab_out, buffers
// other than the ones read. dab NULL: A is not updated.
template <typename R> struct StateUpdate {
    const cplx<R> *dpsi;
    const R *dab;
    cplx<R> *psi_out;
    R *ab_out;
    R ap, aa;  // alpha_psi, alpha_A rounded to the working precision
};

// Everything grad_eval reads of node (i, j): psi there and at its four
// neighbours, epsilon, the material flags of the four cells, the ten link
// values of abei (E) and ab (B) around it, and the link factors (sin, cos)
// of d A on the right link a(i, j) and the up link b(i, j) (valid where
// that link touches material) and of -d A on the left link a(i-1, j) and
// the down link b(i, j-1) (likewise). grad_node gathers them with loads;
// the row-marching gradient pass takes most of them from its register
// window and the neighbouring lanes.
template <typename R> struct GradIn {
    cplx<R> p0, pl, pr, pd, pu;
    R ee;
    Cells c;
    Links10<R> E, B;
    R sr, cr, su, cu, sl, cl, sd, cd;
};

// The free-energy density, dF/dpsi and dF/dA of node (i, j) from gathered
// values (fe_node, jac_psi_node and jac_A_node fused, each term in the
// separate functions' order).
template <typename R, bool withA>
__device__ __forceinline__ void grad_eval(const Geo &g, const GeoR<R> &q, R kappa2, R H, bool he, bool hb,
                                          const GradIn<R> &in, int i, int j, R &e, cplx<R> &gj, R &ga, R &gb,
                                          bool &has_a, bool &has_b, R bscale) {
    const cplx<R> p0 = in.p0;
    const Cells c = in.c;
    const Links10<R> &E = in.E, &B = in.B;
    const cplx<R> zero = {R(0), R(0)};
    const bool lr = c.pm || c.pp, lu = c.mp || c.pp;  // the right link a(i, j) and the up link b(i, j) are active
    const R sr = in.sr, cr = in.cr, su = in.su, cu = in.cu;
    // |psi1 U - psi0|^2 and Im(conj psi0 U psi1) with U's (sin, cos) given
    auto gg = [&](R s, R cs, cplx<R> p1) {
        const R re = p1.re * cs + p1.im * s - p0.re;
        const R im = p1.im * cs - p1.re * s - p0.im;
        return re * re + im * im;
    };
    auto jsu = [&](R s, R cs, cplx<R> p1) {
        return (p0.re * p1.im - p0.im * p1.re) * cs - (p0.re * p1.re + p0.im * p1.im) * s;
    };
    e = R(0);
    gj = zero;
    if (c.mm || c.mp || c.pm || c.pp) {
        const R f_mm = c.mm ? R(1) : R(0), f_mp = c.mp ? R(1) : R(0), f_pm = c.pm ? R(1) : R(0), f_pp = c.pp ? R(1) : R(0);
        const R r_m0 = R(0.5) * (f_mm + f_mp), r_p0 = R(0.5) * (f_pm + f_pp);
        const R r_0m = R(0.5) * (f_mm + f_pm), r_0p = R(0.5) * (f_mp + f_pp);
        const R gw = R(0.25) * (r_m0 + r_p0 + r_0m + r_0p);
        // fe_node
        const R p2 = p0.re * p0.re + p0.im * p0.im;
        e += gw * (R(0.5) * p2 - in.ee) * p2;
        if (lr) e += r_p0 * q.idx2 * gg(sr, cr, in.pr);
        if (lu) e += r_0p * q.idy2 * gg(su, cu, in.pu);
        // jac_psi_node
        const R p = p0.re * p0.re + p0.im * p0.im - in.ee;
        const R k = R(2) * gw * p;
        gj.re += k * p0.re;
        gj.im += k * p0.im;
        auto add = [&](R r, R id2, R s, R cs, cplx<R> p1) {
            const R tr = R(2) * (p0.re - (p1.re * cs + p1.im * s));
            const R ti = R(2) * (p0.im - (p1.im * cs - p1.re * s));
            const R w = r * id2;
            gj.re += w * tr;
            gj.im += w * ti;
        };
        if (c.mm || c.mp) add(r_m0, q.idx2, in.sl, in.cl, in.pl);
        if (lr) add(r_p0, q.idx2, sr, cr, in.pr);
        if (c.mm || c.pm) add(r_0m, q.idy2, in.sd, in.cd, in.pd);
        if (lu) add(r_0p, q.idy2, su, cu, in.pu);
    }
    // fe_node: the cell's magnetic energy
    if (kappa2 > R(0) && i < g.Nxc && j < g.Nyc) {
        R dB = -H;
        if (he) dB += q.idx * (E.bp0 - E.b00) - q.idy * (E.a0p - E.a00);
        if (hb) dB += q.idx * (B.bp0 - B.b00) - q.idy * (B.a0p - B.a00);
        e += kappa2 * dB * dB;
    }
    // jac_A_node
    has_a = has_b = false;
    ga = gb = R(0);
    if (withA) {
        has_a = i < g.Nxa && j < g.Nya;
        has_b = i < g.Nxb && j < g.Nyb;
        if (has_a) {
            R t = curlcurl_a_v(g, q, H, he, E, hb, B, j);
            if (j == 0 || j + 1 == g.Nya) t *= bscale;
            t *= kappa2;
            if (lr) {
                const R r = R(0.5) * ((c.pm ? R(1) : R(0)) + (c.pp ? R(1) : R(0)));
                t += -r * q.idx * jsu(sr, cr, in.pr);
            }
            ga = t;
        }
        if (has_b) {
            R t = curlcurl_b_v(g, q, H, he, E, hb, B, i);
            if (i == 0 || i + 1 == g.Nxb) t *= bscale;
            t *= kappa2;
            if (lu) {
                const R r = R(0.5) * ((c.mp ? R(1) : R(0)) + (c.pp ? R(1) : R(0)));
                t += -r * q.idy * jsu(su, cu, in.pu);
            }
            gb = t;
        }
    }
}

// Cell material flags of node (i, j) from the mask bytes of its four cells
// (the four flags as read; outside the grid they are ignored).
__device__ __forceinline__ Cells cells_from(const Geo &g, bool has_mt, int i, int j, bool m_mm, bool m_mp, bool m_pm,
                                            bool m_pp) {
    Cells c;
    c.mm = (i > 0) && (j > 0);
    c.mp = (i > 0) && (j + 1 < g.Ny);
    c.pm = (i + 1 < g.Nx) && (j > 0);
    c.pp = (i + 1 < g.Nx) && (j + 1 < g.Ny);
    if (has_mt) {
        c.mm = c.mm && m_mm;
        c.mp = c.mp && m_mp;
        c.pm = c.pm && m_pm;
        c.pp = c.pp && m_pp;
    }
    return c;
}

template <typename R, bool withA, bool UPD = false>
__device__ __forceinline__ void grad_node(const Geo &g, const GeoR<R> &q, R kappa2, R eps, const R *eps_map, R H,
                                          const uint8_t *mt, const cplx<R> *psi, const R *abei, const R *ab, long n,
                                          int i, int j, R &e, cplx<R> &gj, R &ga, R &gb, bool &has_a, bool &has_b,
                                          R bscale = R(1), const StateUpdate<R> *S = nullptr) {
    GradIn<R> in;
    // loads
    const long row = (long)g.Nx * j;
    auto lp = [&](long k) {
        cplx<R> p = ld(psi, k);
        if constexpr (UPD) {
            const cplx<R> d = ld(S->dpsi, k);
            p.re = S->ap * d.re + p.re;
            p.im = S->ap * d.im + p.im;
        }
        return p;
    };
    in.p0 = lp(n);
    in.pl = lp(row + max(i - 1, 0));
    in.pr = lp(row + min(i + 1, g.Nx - 1));
    in.pd = lp((long)g.Nx * max(j - 1, 0) + i);
    in.pu = lp((long)g.Nx * min(j + 1, g.Ny - 1) + i);
    in.ee = eps_map ? eps_map[n] : eps;
    bool m_mm = true, m_mp = true, m_pm = true, m_pp = true;
    if (mt) {
        const int cx0 = min(i, g.Nxc - 1), cxm = max(i - 1, 0), cy0 = min(j, g.Nyc - 1), cym = max(j - 1, 0);
        m_mm = mt[cxm + (long)g.Nxc * cym] != 0;
        m_mp = mt[cxm + (long)g.Nxc * cy0] != 0;
        m_pm = mt[cx0 + (long)g.Nxc * cym] != 0;
        m_pp = mt[cx0 + (long)g.Nxc * cy0] != 0;
    }
    in.c = cells_from(g, mt != nullptr, i, j, m_mm, m_mp, m_pm, m_pp);
    const Cells c = in.c;
    const bool he = abei != nullptr, hb = ab != nullptr;
    Links10<R> &E = in.E, &B = in.B;
    E = Links10<R>{};
    B = Links10<R>{};
    if (he) E = load_links10(g, abei, i, j);
    if (hb) B = load_links10(g, ab, i, j);
    if constexpr (UPD) {
        st(S->psi_out, n, in.p0);
        if (S->dab) {
            const Links10<R> d = load_links10(g, S->dab, i, j);
            const R a = S->aa;
            B.a00 = a * d.a00 + B.a00;
            B.am0 = a * d.am0 + B.am0;
            B.a0m = a * d.a0m + B.a0m;
            B.a0p = a * d.a0p + B.a0p;
            B.amp = a * d.amp + B.amp;
            B.b00 = a * d.b00 + B.b00;
            B.b0m = a * d.b0m + B.b0m;
            B.bm0 = a * d.bm0 + B.bm0;
            B.bp0 = a * d.bp0 + B.bp0;
            B.bpm = a * d.bpm + B.bpm;
            if (i < g.Nxa) S->ab_out[i + (long)g.Nxa * j] = B.a00;
            if (j < g.Nyb) S->ab_out[g.Na + i + (long)g.Nxb * j] = B.b00;
        }
    }
    auto pvv = [&](R ve, R vb) {  // pv(abei, ab, k) of loaded values
        R p = R(0);
        if (he) p += ve;
        if (hb) p += vb;
        return p;
    };
    in.sr = in.su = in.sl = in.sd = R(0);
    in.cr = in.cu = in.cl = in.cd = R(1);
    if (c.pm || c.pp) sincos_r(q.dx * pvv(E.a00, B.a00), &in.sr, &in.cr);
    if (c.mp || c.pp) sincos_r(q.dy * pvv(E.b00, B.b00), &in.su, &in.cu);
    if (c.mm || c.mp) sincos_r(-q.dx * pvv(E.am0, B.am0), &in.sl, &in.cl);
    if (c.mm || c.pm) sincos_r(-q.dy * pvv(E.b0m, B.b0m), &in.sd, &in.cd);
    grad_eval<R, withA>(g, q, kappa2, H, he, hb, in, i, j, e, gj, ga, gb, has_a, has_b, bscale);
}

// One pass of the CG loop over the state (phase 1 of an iteration): writes
// both Jacobians (overwriting, equal to the reference's += into a zeroed
// buffer), and accumulates the free energy of the state and the two pairs
// of Polak-Ribiere sums against the previous Jacobians. partials per
// workgroup: F, num_psi, den_psi, num_A, den_A. withA: kappa < inf.
template <typename R, bool withA, bool UPD = false>
__global__ void __launch_bounds__(256) k_cg_gradient_pass(Geo g, R kappa2, R eps, const R *eps_map, R H, const uint8_t *mt,
                                   const cplx<R> *psi, const R *abei, const R *ab, cplx<R> *jac,
                                   const cplx<R> *jac_prev, R *jacA, const R *jacA_prev, double *partials,
                                   StateUpdate<R> S = {}, const double *alpha = nullptr) {
    const GeoR<R> q(g);
    if constexpr (UPD) {
        S.ap = R(alpha[0]);
        S.aa = R(alpha[1]);
    }
    const R dxdy = q.dx * q.dy, s2 = R(2) * q.dx * q.dy;
    double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    for (long n = blockIdx.x * (long)blockDim.x + threadIdx.x; n < g.N; n += (long)gridDim.x * blockDim.x) {
        int i, j;
        node_ij(g, n, i, j);
        R e, ga, gb;
        cplx<R> gj;
        bool ha, hb;
        grad_node<R, withA, UPD>(g, q, kappa2, eps, eps_map, H, mt, psi, abei, ab, n, i, j, e, gj, ga, gb, ha, hb,
                                 R(1), &S);
        // every load of the node before its stores
        const cplx<R> gp = ld(jac_prev, n);
        const long la = i + (long)g.Nxa * j, lb = g.Na + i + (long)g.Nxb * j;
        const R vpa = (withA && ha) ? jacA_prev[la] : R(0), vpb = (withA && hb) ? jacA_prev[lb] : R(0);
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
                jacA[la] = v;
                acc[3] += double(v * (v - vpa));
                acc[4] += double(vpa * vpa);
            }
            if (hb) {
                const R v = s2 * gb;
                jacA[lb] = v;
                acc[3] += double(v * (v - vpb));
                acc[4] += double(vpb * vpb);
            }
        }
    }
    block_store_partials<5>(acc, partials);
}

// Finish of k_cg_gradient_pass: out[0] = dxdy * F (to be read by the host);
// beta_next[0..1] = max(num/den, 0) for psi and A, in FP64 rounded to the
// working precision (cg.py:422-449), left in device memory for the next
// direction update.
template <typename R>
__global__ void __launch_bounds__(256) k_cg_gradient_finish(const double *partials, int nb, double dxdy, double *out, double *beta_next) {
    __shared__ double v[5];
    reduce_partials<5>(partials, nb, v);
    if (threadIdx.x == 0) {
        out[0] = v[0] * dxdy;
        beta_next[0] = double(R(fmax(v[1] / v[2], 0.0)));
        beta_next[1] = double(R(fmax(v[3] / v[4], 0.0)));
    }
}

// P(x, y) = sum c[i][j] x^i y^j over i, j < 5 and its derivatives.
__device__ __forceinline__ void poly2(const double (&c)[5][5], double x, double y, double &p, double &px, double &py,
                                      double &pxx, double &pxy, double &pyy) {
    double xi[5] = {1.0, x, x * x, x * x * x, x * x * x * x};
    double yj[5] = {1.0, y, y * y, y * y * y, y * y * y * y};
    p = px = py = pxx = pxy = pyy = 0.0;
#pragma unroll
    for (int i = 0; i < 5; ++i)
#pragma unroll
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
//
// The polynomial is exact in alpha_psi but a 4th-order expansion of the
// link factors exp(i alpha_A d dA) in alpha_A (cg.h:477-731), whose quartic
// coefficient is negative wherever neighbouring phases are aligned: far from
// 0 it can fall without bound, and an unconstrained descent then runs off to
// a step where the expansion means nothing (from psi = 1 it took |psi| to
// 1e130). So |alpha_A| is kept within ymax, the step at which the largest
// link phase moves by CG_PHASE_MAX (ymax = CG_PHASE_MAX / max|d dA|; there
// the expansion's relative error is below 3e-4); trial points are projected
// onto |y| <= ymax and must lower the polynomial.
constexpr double CG_PHASE_MAX = 0.5;
__device__ __forceinline__ double alpha_A_bound(double thmax) {
    return thmax > 0.0 ? CG_PHASE_MAX / thmax : HUGE_VAL;
}

__device__ __forceinline__ void quartic2_min(const double (&c)[5][5], bool one_d, double &x, double &y,
                                             double ymax = HUGE_VAL) {
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
        double t = 1.0, pn = 0.0, a, b, c2, d, e2, sx = 0.0, sy = 0.0;
        bool ok = false;
        for (int k = 0; k < 60; ++k) {
            sx = t * dx;
            sy = fmin(fmax(y + t * dy, -ymax), ymax) - y;  // projected onto |y| <= ymax
            poly2(c, x + sx, y + sy, pn, a, b, c2, d, e2);
            if (pn < p && pn <= p + 1.0e-4 * fmin(px * sx + py * sy, 0.0)) {
                ok = true;
                break;
            }
            t *= 0.5;
        }
        if (!ok) break;
        x += sx;
        y += sy;
        if (fabs(sx) + fabs(sy) <= 1.0e-15 * (fabs(x) + fabs(y))) break;
    }
}

// quartic2_min run by a whole wave, every lane with the same c: the
// backtracking steps t = 2^-k, k < 60, are tried at once, lane k taking
// 2^-k, and the first that passes the sufficient-decrease test is taken, as
// the sequential loop would; each lane's arithmetic is that loop's k-th
// iteration, so the result is the same to the bit. All 64 lanes must call.
__device__ __forceinline__ void quartic2_min_wave(const double (&c)[5][5], bool one_d, double &x, double &y,
                                                  double ymax = HUGE_VAL) {
    const int lane = threadIdx.x & (WAVE - 1);
    const double tl = ldexp(1.0, -lane);  // 0.5 applied lane times to 1.0
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
        double pn, a, b, c2, d, e2;
        const double sx = tl * dx;
        const double sy = fmin(fmax(y + tl * dy, -ymax), ymax) - y;
        poly2(c, x + sx, y + sy, pn, a, b, c2, d, e2);
        const bool pass = lane < 60 && pn < p && pn <= p + 1.0e-4 * fmin(px * sx + py * sy, 0.0);
        const unsigned long long m = __ballot(pass);
        if (!m) break;
        const int k = __ffsll((long long)m) - 1;
        const double tx = __shfl(sx, k, WAVE), ty = __shfl(sy, k, WAVE);
        x += tx;
        y += ty;
        if (fabs(tx) + fabs(ty) <= 1.0e-15 * (fabs(x) + fabs(y))) break;
    }
}

// Device line search of the FP32-mixed CG (glx_cg_line_step): the L
// coefficient sums of k_coef (partials), times scale, packed as the
// reference packs them (cg.py:369-373, c[i][0] for the psi-only quartic),
// minimised in FP64 from (0, 0) by quartic2_min; the alphas go to
// alpha[0..1] and the coefficients to c[0..L-1]. One workgroup.
template <int L>
__global__ void __launch_bounds__(256) k_line_search(const double *partials, int nb, double scale, double *c, double *alpha) {
    __shared__ double lds[RED_BLOCK / WAVE];
    __shared__ double red[L];
    reduce_partials<L>(partials, nb, red);
    if (threadIdx.x < L) red[threadIdx.x] *= scale;  // read by thread 0 after block_max's barriers
    // the coefficient pass left each block's largest link phase step after
    // the sums (k_coef, k_coef_dir with L = 17)
    double th = 0.0;
    if (L == 17)
        for (int b = threadIdx.x; b < nb; b += blockDim.x) th = fmax(th, partials[(long)nb * L + b]);
    th = block_max(th, lds);
    th = __shfl(th, 0, WAVE);  // block_max leaves it in thread 0
    if (threadIdx.x < WAVE) {
        double c5[5][5] = {};
        if (L == 17) {
            for (int k = 0; k < 5; ++k) {
                c5[0][k] = red[k];
                c5[1][k] = red[5 + k];
                c5[2][k] = red[10 + k];
            }
            c5[3][0] = red[15 % L];
            c5[4][0] = red[16 % L];
        } else {
            for (int k = 0; k < 5; ++k) c5[k][0] = red[k];
        }
        double x, y;
        quartic2_min_wave(c5, L != 17, x, y, alpha_A_bound(th));
        if (threadIdx.x == 0) {
            alpha[0] = x;
            alpha[1] = y;
            for (int k = 0; k < L; ++k) c[k] = red[k];
        }
    }
}

// z = alpha*x + y with alpha read from device memory and rounded to the
// working precision (the reference's axpy with the host alpha, cg.py:296,
// 535-536).
template <typename R>
__global__ void __launch_bounds__(256) k_axpy_dev(const R *x, const R *y, R *z, const double *alpha, long n) {
    const R a = R(*alpha);
    for (long k = blockIdx.x * (long)blockDim.x + threadIdx.x; k < n; k += (long)gridDim.x * blockDim.x)
        z[k] = a * x[k] + y[k];
}

// Line-search coefficients. L = 5 (cg.h:314-474) or 17 (cg.h:477-731).
// One node's contributions are added to C (before the dxdy factor). The
// reference's kernels read only the scalar eps (eps_map is NULL on the
// drop-in path); the batched CG passes the map. Every value the node reads
// (psi, dpsi and their right and up neighbours, the cells, epsilon, the four
// link values a(i, j), a(i, j+1), b(i, j), b(i+1, j) of abei, ab and dab)
// is loaded up front from clamped indices, and the link's sin/cos is
// evaluated once for all its terms; the arithmetic is the reference's term
// by term.
template <typename R> struct Links4 {
    R a00, a0p, b00, bp0;
};

template <typename R> __device__ __forceinline__ Links4<R> load_links4(const Geo &g, const R *v, int i, int j) {
    const int xa = min(i, g.Nxa - 1), yap = min(j + 1, g.Nya - 1);
    const int xbp = min(i + 1, g.Nxb - 1), yb = min(j, g.Nyb - 1);
    const R *b = v + g.Na;
    Links4<R> L;
    L.a00 = v[xa + (long)g.Nxa * j];
    L.a0p = v[xa + (long)g.Nxa * yap];
    L.b00 = b[i + (long)g.Nxb * yb];
    L.bp0 = b[xbp + (long)g.Nxb * yb];
    return L;
}

// The CG direction update fused into the coefficient pass: the direction
// d = beta * d_old - jac (axmy with the device beta, cg.py:283-290, 491-530)
// is formed wherever coef_node reads it, from d_old and jac, and node n
// writes its own entries (d(n), and the links a(i, j), b(i, j) it owns)
// to d_out, a buffer other than d_old. The arithmetic is k_axmy_dev's.
template <typename R> struct DirUpdate {
    const cplx<R> *dold, *jac;
    const R *dAold, *jacA;
    cplx<R> *dout;
    R *dAout;
    R bp, ba;  // beta_psi, beta_A rounded to the working precision
};

template <typename R>
__device__ __forceinline__ cplx<R> dir_at(const DirUpdate<R> &U, long k) {
    const cplx<R> o = ld(U.dold, k), gk = ld(U.jac, k);
    return {U.bp * o.re - gk.re, U.bp * o.im - gk.im};
}

// x / N for the coefficients' constant divisors: FP64 divides as the
// reference does; FP32 multiplies by the rounded reciprocal (within an ulp
// of the quotient, against about ten instructions for an IEEE division).
template <typename R, int N> __device__ __forceinline__ R div_const(R x) {
    if constexpr (std::is_same<R, float>::value) return x * (1.0f / float(N));
    else return x / R(N);
}

// Everything coef_eval reads of node (i, j): psi and the direction at the
// node and its right and up neighbours, epsilon, the four cell flags, and
// the four link values a(i, j), a(i, j+1), b(i, j), b(i+1, j) of abei (E),
// ab (B) and the A direction (D).
template <typename R> struct CoefIn {
    cplx<R> p0, p_r, p_u, d0, d_r, d_u;
    R eps;
    Cells c;
    Links4<R> E, B, D;
};

// One node's contributions to the L line-search coefficients, added to C
// (before the dxdy factor), from gathered values: the reference's terms in
// its order, the sin/cos of each link phase evaluated once.
template <typename R, int L>
__device__ __forceinline__ void coef_eval(const GeoR<R> &q, R kappa2, R H, bool he, bool hb, const CoefIn<R> &in,
                                          int i, int j, bool cell_ij, R (&C)[L]) {
    const cplx<R> p0 = in.p0, p_r = in.p_r, p_u = in.p_u, d0 = in.d0, d_r = in.d_r, d_u = in.d_u;
    const R eps = in.eps;
    const Cells c = in.c;
    const Links4<R> &E = in.E, &B = in.B, &D = in.D;
    auto pvv = [&](R ve, R vb) {  // pv(abei, ab, k) of loaded values
        R p = R(0);
        if (he) p += ve;
        if (hb) p += vb;
        return p;
    };

    if (c.mm || c.mp || c.pm || c.pp) {
        const R f_mm = c.mm ? R(1) : R(0), f_mp = c.mp ? R(1) : R(0), f_pm = c.pm ? R(1) : R(0), f_pp = c.pp ? R(1) : R(0);
        const R r_m0 = R(0.5) * (f_mm + f_mp), r_p0 = R(0.5) * (f_pm + f_pp);
        const R r_0m = R(0.5) * (f_mm + f_pm), r_0p = R(0.5) * (f_mp + f_pp);
        const R gw = R(0.25) * (r_m0 + r_p0 + r_0m + r_0p);
        const R p2 = p0.re * p0.re + p0.im * p0.im;
        const R d2 = d0.re * d0.re + d0.im * d0.im;
        const R tw = R(2) * (p0.re * d0.re + p0.im * d0.im);
        const R k0 = gw * (R(0.5) * p2 - eps) * p2;
        const R k1 = gw * tw * (p2 - eps);
        const R k2 = gw * (-eps * d2 + R(0.5) * tw * tw + p2 * d2);
        const R k3 = gw * tw * d2;
        const R k4 = gw * R(0.5) * d2 * d2;
        if (L == 5) { C[0] += k0; C[1] += k1; C[2] += k2; C[3] += k3; C[4] += k4; }
        else { C[0] += k0; C[5] += k1; C[10] += k2; C[15 % L] += k3; C[16 % L] += k4; }
        for (int lk = 0; lk < 2; ++lk) {
            const bool on = lk == 0 ? (c.pm || c.pp) : (c.mp || c.pp);
            if (!on) continue;
            const R r = lk == 0 ? r_p0 : r_0p;
            const R id2 = lk == 0 ? q.idx2 : q.idy2;
            const R d = lk == 0 ? q.dx : q.dy;
            R ph = R(0);
            if (he) ph += d * (lk == 0 ? E.a00 : E.b00);
            if (hb) ph += d * (lk == 0 ? B.a00 : B.b00);
            const cplx<R> p1 = lk == 0 ? p_r : p_u, d1 = lk == 0 ? d_r : d_u;
            R s, cs;
            sincos_r(ph, &s, &cs);  // U(-ph) = (cos ph, sin ph)
            auto gc = [&](cplx<R> a0, cplx<R> a1) -> cplx<R> {  // psi1 U(ph) - psi0 (cg.h:304-311)
                return {a1.re * cs + a1.im * s - a0.re, a1.im * cs - a1.re * s - a0.im};
            };
            auto gg = [&](cplx<R> a0, cplx<R> a1) -> R {  // g_grad(a0, ph, a1)
                const R re = a1.re * cs + a1.im * s - a0.re;
                const R im = a1.im * cs - a1.re * s - a0.im;
                return re * re + im * im;
            };
            const cplx<R> zp = gc(p0, p1), zd = gc(d0, d1);
            const R zr = zp.re * zd.re - (-zp.im) * zd.im;  // Re(conj(zp) zd)
            const R w = r * id2;
            if (L == 5) {
                C[2] += w * gg(d0, d1);
                C[1] += w * R(2) * zr;
                C[0] += w * gg(p0, p1);
            } else {
                const R dph = d * (lk == 0 ? D.a00 : D.b00), dph2 = dph * dph;
                C[0] += w * gg(p0, p1);
                C[5] += w * R(2) * zr;
                C[10] += w * gg(d0, d1);
                auto mulU = [&](cplx<R> a, cplx<R> b) -> cplx<R> {  // a * U(-ph) * conj(b)
                    const cplx<R> t = {a.re * cs - a.im * s, a.re * s + a.im * cs};
                    return {t.re * b.re - t.im * (-b.im), t.re * (-b.im) + t.im * b.re};
                };
                const cplx<R> z0 = mulU(p0, p1);
                const cplx<R> za = mulU(p0, d1), zb = mulU(d0, p1);
                const cplx<R> z1 = {za.re + zb.re, za.im + zb.im};
                const cplx<R> z2 = mulU(d0, d1);
                const cplx<R> zz[3] = {z0, z1, z2};
                if constexpr (std::is_same<R, float>::value) {
                    // FP32: the powers of the phase step once per link, then
                    // w Im z and w Re z once per product, so each coefficient
                    // takes one fma (the reference's per-term products
                    // reassociated; FP64 keeps its order)
                    const R k1 = R(2) * dph, k3 = R(-1.0f / 3.0f) * dph2 * dph, k4 = R(-1.0f / 12.0f) * dph2 * dph2;
#pragma unroll
                    for (int m = 0; m < 3; ++m) {
                        const R wi = w * zz[m].im, wr = w * zz[m].re;
                        C[5 * m + 1] = fmaf(wi, k1, C[5 * m + 1]);
                        C[5 * m + 2] = fmaf(wr, dph2, C[5 * m + 2]);
                        C[5 * m + 3] = fmaf(wi, k3, C[5 * m + 3]);
                        C[5 * m + 4] = fmaf(wr, k4, C[5 * m + 4]);
                    }
                } else {
#pragma unroll
                    for (int m = 0; m < 3; ++m) {
                        C[5 * m + 1] += w * R(2) * zz[m].im * dph;
                        C[5 * m + 2] += w * zz[m].re * dph2;
                        C[5 * m + 3] += div_const<R, 3>(-r * id2) * zz[m].im * dph2 * dph;
                        C[5 * m + 4] += div_const<R, 12>(-r * id2) * zz[m].re * dph2 * dph2;
                    }
                }
            }
        }
    }
    if (kappa2 > R(0) && cell_ij) {
        if (L == 5) {
            const R dB = q.idx * (B.bp0 - B.b00) - q.idy * (B.a0p - B.a00) - H;
            C[0] += kappa2 * dB * dB;
        } else {
            const R BH = q.idx * (pvv(E.bp0, B.bp0) - pvv(E.b00, B.b00)) - q.idy * (pvv(E.a0p, B.a0p) - pvv(E.a00, B.a00)) - H;
            const R dB = q.idx * (D.bp0 - D.b00) - q.idy * (D.a0p - D.a00);
            C[0] += kappa2 * BH * BH;
            C[1] += kappa2 * R(2) * BH * dB;
            C[2] += kappa2 * dB * dB;
        }
    }
}

// With L = 17 and th not NULL, *th also takes the largest link phase step
// |d dA| of the node's links a(i, j), b(i, j) (for the line search's bound).
template <typename R, int L, bool DIR = false>
__device__ __forceinline__ void coef_node(const Geo &g, const GeoR<R> &q, R kappa2, R eps_s, const R *eps_map, R H,
                                          const uint8_t *mt, const cplx<R> *psi, const cplx<R> *dpsi, const R *abei,
                                          const R *ab, const R *dab, long n, R (&C)[L],
                                          const DirUpdate<R> *U = nullptr, double *th = nullptr) {
    int i, j;
    node_ij(g, n, i, j);
    CoefIn<R> in;
    // loads
    const long nr = (long)g.Nx * j + min(i + 1, g.Nx - 1), nu = (long)g.Nx * min(j + 1, g.Ny - 1) + i;
    in.p0 = ld(psi, n);
    in.p_r = ld(psi, nr);
    in.p_u = ld(psi, nu);
    if constexpr (DIR) {
        in.d0 = dir_at(*U, n);
        in.d_r = dir_at(*U, nr);
        in.d_u = dir_at(*U, nu);
    } else {
        in.d0 = ld(dpsi, n);
        in.d_r = ld(dpsi, nr);
        in.d_u = ld(dpsi, nu);
    }
    in.eps = eps_map ? eps_map[n] : eps_s;
    bool m_mm = true, m_mp = true, m_pm = true, m_pp = true;
    if (mt) {
        const int cx0 = min(i, g.Nxc - 1), cxm = max(i - 1, 0), cy0 = min(j, g.Nyc - 1), cym = max(j - 1, 0);
        m_mm = mt[cxm + (long)g.Nxc * cym] != 0;
        m_mp = mt[cxm + (long)g.Nxc * cy0] != 0;
        m_pm = mt[cx0 + (long)g.Nxc * cym] != 0;
        m_pp = mt[cx0 + (long)g.Nxc * cy0] != 0;
    }
    in.c = cells_from(g, mt != nullptr, i, j, m_mm, m_mp, m_pm, m_pp);
    const bool he = abei != nullptr, hb = ab != nullptr;
    in.E = Links4<R>{};
    in.B = Links4<R>{};
    in.D = Links4<R>{};
    if (he) in.E = load_links4(g, abei, i, j);
    if (hb) in.B = load_links4(g, ab, i, j);
    if (L == 17) {
        if constexpr (DIR) {
            const Links4<R> o = load_links4(g, U->dAold, i, j), gk = load_links4(g, U->jacA, i, j);
            in.D.a00 = U->ba * o.a00 - gk.a00;
            in.D.a0p = U->ba * o.a0p - gk.a0p;
            in.D.b00 = U->ba * o.b00 - gk.b00;
            in.D.bp0 = U->ba * o.bp0 - gk.bp0;
        } else {
            in.D = load_links4(g, dab, i, j);
        }
        if (th) *th = fmax(*th, fmax(fabs(double(q.dx * in.D.a00)), fabs(double(q.dy * in.D.b00))));
    }
    if constexpr (DIR) {
        st(U->dout, n, in.d0);
        if (L == 17) {
            if (i < g.Nxa) U->dAout[i + (long)g.Nxa * j] = in.D.a00;
            if (j < g.Nyb) U->dAout[g.Na + i + (long)g.Nxb * j] = in.D.b00;
        }
    }
    coef_eval<R, L>(q, kappa2, H, he, hb, in, i, j, i < g.Nxc && j < g.Nyc, C);
}

// Partials: the L sums of block b at partials[b L + k]; with L = 17 the
// block's largest link phase step after all sums, at partials[nb L + b].
__device__ __forceinline__ void block_store_phase_max(double th, double *partials, int L) {
    __shared__ double lds[RED_BLOCK / WAVE];
    th = block_max(th, lds);
    if (threadIdx.x == 0) partials[(long)gridDim.x * L + blockIdx.x] = th;
}

template <typename R, int L>
__global__ void __launch_bounds__(256) k_coef(Geo g, R kappa2, R eps, R H, const uint8_t *mt, const cplx<R> *psi, const cplx<R> *dpsi,
                       const R *abei, const R *ab, const R *dab, double *partials) {
    const GeoR<R> q(g);
    double acc[L];
#pragma unroll
    for (int k = 0; k < L; ++k) acc[k] = 0.0;
    double th = 0.0;
    for (long n = blockIdx.x * (long)blockDim.x + threadIdx.x; n < g.N; n += (long)gridDim.x * blockDim.x) {
        R C[L];
#pragma unroll
        for (int k = 0; k < L; ++k) C[k] = R(0);
        coef_node<R, L>(g, q, kappa2, eps, (const R *)nullptr, H, mt, psi, dpsi, abei, ab, dab, n, C, nullptr,
                        L == 17 ? &th : nullptr);
#pragma unroll
        for (int k = 0; k < L; ++k) acc[k] += double(C[k]);
    }
    block_store_partials<L>(acc, partials);
    if (L == 17) block_store_phase_max(th, partials, L);
}

// k_coef with the direction update fused in (DirUpdate): the coefficients
// of F(psi + a d_psi, A + b d_A) for the new direction, which is written to
// U.dout / U.dAout on the way. beta[0..1] are the device betas.
template <typename R, int L>
__global__ void __launch_bounds__(256) k_coef_dir(Geo g, R kappa2, R eps, R H, const uint8_t *mt, const cplx<R> *psi,
                                                  const R *abei, const R *ab, DirUpdate<R> U, const double *beta,
                                                  double *partials) {
    const GeoR<R> q(g);
    U.bp = R(beta[0]);
    U.ba = R(beta[1]);
    double acc[L];
#pragma unroll
    for (int k = 0; k < L; ++k) acc[k] = 0.0;
    double th = 0.0;
    for (long n = blockIdx.x * (long)blockDim.x + threadIdx.x; n < g.N; n += (long)gridDim.x * blockDim.x) {
        R C[L];
#pragma unroll
        for (int k = 0; k < L; ++k) C[k] = R(0);
        coef_node<R, L, true>(g, q, kappa2, eps, (const R *)nullptr, H, mt, psi, nullptr, abei, ab, nullptr, n, C, &U,
                              L == 17 ? &th : nullptr);
#pragma unroll
        for (int k = 0; k < L; ++k) acc[k] += double(C[k]);
    }
    block_store_partials<L>(acc, partials);
    if (L == 17) block_store_phase_max(th, partials, L);
}

// --------------------------------------------------------------- vortices

// fmod(x, b) for b > 0 from an estimate of the truncated quotient: the
// remainder x - n b of the right n is a representable number, so one FMA
// gives it exactly; n = trunc(x * inv_b), inv_b the rounded 1 / b, is off by
// at most one for |x / b| < 2^20, which the sign and size of the FMA's
// result reveal (a wrong n leaves a nonzero remainder of the wrong sign or
// at least b in size). Larger or non-finite x take the library fmod.
__device__ __forceinline__ double fmod_pos(double x, double b, double inv_b) {
    double n = trunc(x * inv_b);
    if (!(fabs(n) < 1048576.0)) return fmod(x, b);
    double r = fma(-n, b, x);
    if (x >= 0.0) {
        if (r < 0.0) n -= 1.0;
        else if (r >= b) n += 1.0;
        else return r;
    } else {
        if (r > 0.0) n += 1.0;
        else if (r <= -b) n -= 1.0;
        else return r;
    }
    return fma(-n, b, x);
}

// np.mod(x, 2 pi) for float64 (fmod, shifted into [0, 2 pi)).
__device__ __forceinline__ double np_mod_2pi(double x) {
    const double b = 6.283185307179586;  // 2.0*np.pi
    double m = fmod_pos(x, b, 0.15915494309189535);  // 1 / b rounded
    if (m != 0.0) {
        if (m < 0.0) m += b;
    } else {
        m = 0.0;
    }
    return m;
}

constexpr int VREC = 14;

// Winding number of one cell (vortex_detector.py:55-72) from the phases of
// its corners (i, j), (i+1, j), (i, j+1), (i+1, j+1) and its links a(i, j),
// a(i, j+1), b(i, j), b(i+1, j), term by term in the reference's order.
__device__ __forceinline__ double cell_winding(const Geo &g, double dxdyH, double t00, double tp0, double t0p,
                                               double tpp, double a0, double a1, double b0, double bp) {
    const double pi = 3.141592653589793;
    return -(0.5 / pi) * (np_mod_2pi(tp0 - t00 - g.dx * a0 + pi) + np_mod_2pi(tpp - tp0 - g.dy * bp + pi) +
                          np_mod_2pi(t0p - tpp + g.dx * a1 + pi) + np_mod_2pi(t00 - t0p + g.dy * b0 + pi) - 4.0 * pi +
                          dxdyH);
}

// The reference keeps cells with |v| > 0.5 whose v is within 0.1 of an
// integer (vortex_detector.py:67-72).
__device__ __forceinline__ bool winding_flag(double v) { return fabs(v) > 0.5 && fabs(v - rint(v)) < 0.1; }

// Vortex cells without an atomic: (1) k_vortex_flags evaluates every cell
// and stores, per strip row, the 64-bit mask of the flagged cells (at
// s Nyc + j); (2) two small launches turn the masks into each strip row's
// first rank in cell-id order (k_vortex_rowprefix, k_vortex_rowscan); (3)
// k_vortex_emit walks the masks again and writes
// a record for each set bit at its rank, re-evaluating only those cells
// (the same expression on the same values, so the same v). Records come
// out in cell-id order, the reference's loop order.
//
// Cells are cut into strips of 63 columns; in (1) a wave marches down a
// range of cell rows of one strip, lane l holding node column i = 63 s + l
// of the two node rows j, j + 1 that bound cell row j (phase, psi and the a
// link of each) and the b link of row j; lane l + 1 supplies the right-hand
// corner values by a lane shift, so lanes 0..62 evaluate their cells and
// each node's phase is evaluated once. Bounds: node rows j + 1 <= Ny - 1
// and b rows j <= Ny - 2 since j < Nyc; columns past the grid are clamped
// to the last one for their loads and never evaluate a cell.
constexpr int VSTRIP = WAVE - 1;
constexpr long VORTEX_MAX_WAVES = 32768;  // waves of the flag and emit launches

inline long vortex_rows(const Geo &g) { return (long)((g.Nxc + VSTRIP - 1) / VSTRIP) * g.Nyc; }

inline long vortex_waves(const Geo &g) {
    long W = (vortex_rows(g) + 15) / 16;  // about 16 rows per wave
    if (W > VORTEX_MAX_WAVES) W = VORTEX_MAX_WAVES;
    return W < 1 ? 1 : W;
}

// The (strip, cell row range) items of wave w, handed out row-major as the
// row-marching sweeps do: the cell rows are cut into S = max(1, W / ns)
// segments and item k = (segment, strip) goes to wave k mod U, U =
// min(W, S ns), so waves of consecutive index march the same rows of
// adjacent strips. f(s, ja, jb) per item, in a fixed order.
template <typename F> __device__ __forceinline__ void vortex_items(const Geo &g, long w, long W, F &&f) {
    const long ns = (g.Nxc + VSTRIP - 1) / VSTRIP;
    const long S = W >= ns ? W / ns : 1;
    const long U = W < S * ns ? W : S * ns;
    for (long k = w; k < S * ns; k += U) {
        const long seg = k / ns;
        const int ja = (int)((seg * g.Nyc) / S), jb = (int)(((seg + 1) * g.Nyc) / S);
        if (ja < jb) f((int)(k % ns), ja, jb);
    }
}

// Rows are taken VROWS at a time: their loads are issued together and
// their phases evaluated together, so a wave waits for memory once per
// VROWS rows.
constexpr int VROWS = 8;

// The flag test in FP32 with a guard band. Phases are FP32 atan2f of the
// node values, each np.mod term a float remainder, and v is summed in
// float: against the FP64 evaluation each term is off by a few 1e-6 rad at
// most while every link term |dx a| stays below 1e3 (beyond, the cell is
// decided in FP64), v by about 1e-6. A cell whose
// terms all lie more than VGUARD from the wrap point of np.mod (0 and 2 pi)
// and whose |v| and |v - round(v)| lie more than VGUARD from the thresholds
// 0.5 and 0.1 gets the same flag in FP64; every other cell, and every cell
// with a corner whose magnitude float cannot resolve (nonzero below 1e-30
// or above 1e30), is decided by the FP64 expression itself.
constexpr float VGUARD = 1e-3f;

template <typename R> __device__ __forceinline__ bool phase_unresolved(const cplx<R> &p) {
    const R m = fmax(fabs(p.re), fabs(p.im));
    return m != R(0) && !(m >= R(1e-30) && m <= R(1e30));
}

__device__ __forceinline__ float wrap_2pi_guarded(float x, bool &amb) {
    const float two_pi = 6.28318530717958647692f;
    const float m = x - two_pi * floorf(x * 0.159154943091895335769f);
    amb = amb || !(m > VGUARD && m < two_pi - VGUARD);
    return m;
}

template <typename R>
__global__ void __launch_bounds__(256) k_vortex_flags(Geo g, const cplx<R> *psi, const R *ab, double H,
                                                      unsigned long long *bits, long W) {
    const double dxdyH = g.dx * g.dy * H;
    const float pi_f = 3.14159265358979323846f;
    const float dxf = float(g.dx), dyf = float(g.dy), dxdyHf = float(dxdyH);
    const long w = (long)xcd_block() * (blockDim.x / WAVE) + wave_id();
    if (w >= W) return;  // wave-uniform
    const int lane = threadIdx.x & (WAVE - 1);
    const R *bl = ab + g.Na;
    vortex_items(g, w, W, [&](int s, int ja, int jb) {
        const int i = s * VSTRIP + lane;
        const bool col = i < g.Nx, cell = lane < VSTRIP && i < g.Nxc;
        const int ic = min(i, g.Nx - 1), iac = min(i, g.Nxa - 1);
        cplx<R> p0 = ld(psi, ic + (long)g.Nx * ja);
        float f0 = col ? atan2f(float(p0.im), float(p0.re)) : 0.f;
        bool u0 = col && phase_unresolved(p0);
        R a0 = ab[iac + (long)g.Nxa * ja];
        for (int j = ja; j < jb; j += VROWS) {
            // rows j + k (clamped to jb - 1 for the loads: node row <= Nyc,
            // b row <= Nyc - 1)
            cplx<R> p1[VROWS];
            R a1[VROWS], b0[VROWS];
#pragma unroll
            for (int k = 0; k < VROWS; ++k) {
                const int jj = min(j + k, jb - 1);
                p1[k] = ld(psi, ic + (long)g.Nx * (jj + 1));
                a1[k] = ab[iac + (long)g.Nxa * (jj + 1)];
                b0[k] = bl[ic + (long)g.Nxb * jj];
            }
            float f1[VROWS];
#pragma unroll
            for (int k = 0; k < VROWS; ++k) f1[k] = col ? atan2f(float(p1[k].im), float(p1[k].re)) : 0.f;
#pragma unroll
            for (int k = 0; k < VROWS; ++k) {
                if (j + k < jb) {  // wave-uniform
                    const bool u1 = col && phase_unresolved(p1[k]);
                    const float f0r = shfl_dn1(f0), f1r = shfl_dn1(f1[k]);
                    const R b0r = shfl_dn1(b0[k]);
                    const bool ur = shfl_dn1((int)(u0 || u1)) != 0;
                    const R p0r_re = shfl_dn1(p0.re), p0r_im = shfl_dn1(p0.im);
                    const R p1r_re = shfl_dn1(p1[k].re), p1r_im = shfl_dn1(p1[k].im);
                    bool flag = false;
                    if (cell) {
                        bool amb = u0 || u1 || ur;
                        const float sum = wrap_2pi_guarded(f0r - f0 - dxf * float(a0) + pi_f, amb) +
                                          wrap_2pi_guarded(f1r - f0r - dyf * float(b0r) + pi_f, amb) +
                                          wrap_2pi_guarded(f1[k] - f1r + dxf * float(a1[k]) + pi_f, amb) +
                                          wrap_2pi_guarded(f0 - f1[k] + dyf * float(b0[k]) + pi_f, amb);
                        // link terms beyond 1e3 rad carry float errors near the guard band
                        amb = amb || !(fmaxf(fmaxf(fabsf(dxf * float(a0)), fabsf(dyf * float(b0r))),
                                             fmaxf(fabsf(dxf * float(a1[k])), fabsf(dyf * float(b0[k])))) < 1e3f);
                        const float v = -(0.5f / pi_f) * (sum - 4.0f * pi_f + dxdyHf);
                        const float av = fabsf(v), dv = fabsf(v - rintf(v));
                        amb = amb || !(fabsf(av - 0.5f) > VGUARD) || !(fabsf(dv - 0.1f) > VGUARD);
                        if (!amb) {
                            flag = av > 0.5f && dv < 0.1f;
                        } else {  // the exact test, as k_vortex_emit evaluates it
                            const double t00 = atan2(double(p0.im), double(p0.re));
                            const double tp0 = atan2(double(p0r_im), double(p0r_re));
                            const double t0p = atan2(double(p1[k].im), double(p1[k].re));
                            const double tpp = atan2(double(p1r_im), double(p1r_re));
                            flag = winding_flag(cell_winding(g, dxdyH, t00, tp0, t0p, tpp, double(a0), double(a1[k]),
                                                             double(b0[k]), double(b0r)));
                        }
                    }
                    const unsigned long long m = __ballot(flag);
                    if (lane == 0) bits[(long)s * g.Nyc + j + k] = m;
                    p0 = p1[k];
                    f0 = f1[k];
                    u0 = u1;
                    a0 = a1[k];
                }
            }
        }
    });
}

// The rank of each strip row's first flagged cell in cell-id order (row
// by row, strips left to right within a row), from the masks, in two
// small launches: k_vortex_rowprefix stores, per strip row, the flagged
// cells of the strips to its left in the same cell row (pre[s Nyc + j],
// the layout of the masks) and per cell row their total; k_vortex_rowscan
// turns the totals into the rank of each cell row's first cell and stores
// the grand total. A cell's rank is then rowoff[j] + pre[s Nyc + j] + its
// place among the strip row's set bits. In k_vortex_rowprefix a workgroup
// takes 64 rows, each of its four waves a contiguous quarter of the strips
// (a wave reads 64 consecutive rows of a strip), and the quarters' sums,
// exchanged through LDS, give each wave the offset of its first strip.
static __global__ void __launch_bounds__(256) k_vortex_rowprefix(const unsigned long long *bits, long ns, int Nyc,
                                                                 int *pre, int *rowcnt) {
    __shared__ int part[4][WAVE];
    const int r = threadIdx.x & (WAVE - 1), q = threadIdx.x / WAVE;
    const long j = (long)blockIdx.x * WAVE + r;
    const long s0 = q * ns / 4, s1 = (q + 1) * ns / 4;
    int c = 0;
    if (j < Nyc)
        for (long s = s0; s < s1; ++s) c += (int)__popcll(bits[s * Nyc + j]);
    part[q][r] = c;
    __syncthreads();
    if (j < Nyc) {
        int run = 0;
        for (int k = 0; k < q; ++k) run += part[k][r];
        for (long s = s0; s < s1; ++s) {
            pre[s * Nyc + j] = run;
            run += (int)__popcll(bits[s * Nyc + j]);
        }
        if (q == 0) rowcnt[j] = part[0][r] + part[1][r] + part[2][r] + part[3][r];
    }
}

// In place, one workgroup of 1024 threads, each a contiguous run of rows.
// Counts and offsets stay below 2^30 (the cells of a grid).
static __global__ void __launch_bounds__(1024) k_vortex_rowscan(int *rowcnt, int Nyc, int *total) {
    __shared__ long part[1024];
    const int t = threadIdx.x;
    const long C = ((long)Nyc + 1023) / 1024, a = t * C, b = min((long)Nyc, a + C);
    long sum = 0;
    for (long j = a; j < b; ++j) sum += rowcnt[j];
    part[t] = sum;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {  // inclusive scan of the run sums
        const long v = t >= off ? part[t - off] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    long run = part[t] - sum;
    for (long j = a; j < b; ++j) {
        const int c = rowcnt[j];
        rowcnt[j] = (int)run;
        run += c;
    }
    if (t == 1023) *total = (int)min(part[1023], (long)INT_MAX);
}

// Exclusive prefix sum of v over the lanes of a wave.
__device__ __forceinline__ int wave_excl_scan(int v) {
    const int lane = threadIdx.x & (WAVE - 1);
    int x = v;
#pragma unroll
    for (int d = 1; d < WAVE; d <<= 1) {
        const int y = __shfl_up(x, d, WAVE);
        if (lane >= d) x += y;
    }
    return x - v;
}

// Position of the n-th (from 0) set bit of m; m has more than n set bits.
__device__ __forceinline__ int nth_set_bit(unsigned long long m, int n) {
    int pos = 0;
#pragma unroll
    for (int w = 32; w >= 1; w >>= 1) {
        const unsigned long long lo = m & ((1ull << w) - 1ull);
        const int c = (int)__popcll(lo);
        if (n >= c) {
            n -= c;
            m >>= w;
            pos += w;
        } else {
            m = lo;
        }
    }
    return pos;
}

// Records of the flagged cells in cell-id order: (cell n, v, psi at the
// four corners counter-clockwise from (i, j), a(i, j), a(i, j+1), b(i, j),
// b(i+1, j)) with abi added to the links when given. A cell's slot is its
// cell row's rank (rowoff) plus the cells of the strips to its left in that
// row (pre) plus its place among its strip row's set bits. Slots past
// `capacity` are not written.
template <typename R>
__global__ void __launch_bounds__(256) k_vortex_emit(Geo g, const cplx<R> *psi, const R *ab, const R *abi, double H,
                                                     const unsigned long long *bits, const int *pre,
                                                     const int *rowoff, double *rec, int capacity, long W) {
    const double dxdyH = g.dx * g.dy * H;
    const long w = (long)xcd_block() * (blockDim.x / WAVE) + wave_id();
    if (w >= W) return;  // wave-uniform
    const int lane = threadIdx.x & (WAVE - 1);
    vortex_items(g, w, W, [&](int s, int ja, int jb) {
        const long r0 = (long)s * g.Nyc + ja, r1 = (long)s * g.Nyc + jb;
        for (long rc = r0; rc < r1; rc += WAVE) {
            // the masks of up to 64 rows in one load, lane k holding row rc + k
            const unsigned long long mk = rc + lane < r1 ? bits[rc + lane] : 0ull;
            // rank of the row's first cell: its cell row's offset plus the
            // strips to its left in that row
            const int pk = rc + lane < r1 ? pre[rc + lane] + rowoff[ja + (rc - r0) + lane] : 0;
            const int ck = (int)__popcll(mk);
            const int ek = wave_excl_scan(ck);  // flagged cells in the rows before row rc + lane
            const int total = __shfl(ek + ck, WAVE - 1, WAVE);
            // the flagged cells of these rows, one per lane: cell f lies in
            // the last row k with ek <= f, at the (f - ek)-th set bit
            for (int f0 = 0; f0 < total; f0 += WAVE) {
                const int f = f0 + lane;
                int k = 0;
#pragma unroll
                for (int step = WAVE / 2; step >= 1; step >>= 1) {
                    const int e = __shfl(ek, min(k + step, WAVE - 1), WAVE);
                    if (k + step < WAVE && e <= f) k += step;
                }
                const unsigned long long m = __shfl(mk, k, WAVE);
                const int e = __shfl(ek, k, WAVE);
                const long q = (long)__shfl(pk, k, WAVE) + (f - e);
                if (f < total) {
                    const int j = ja + (int)(rc - r0) + k;
                    const int i = s * VSTRIP + nth_set_bit(m, f - e);  // i < Nxc, j < Nyc: a cell
                    const long n00 = i + (long)g.Nx * j, n0p = n00 + g.Nx;
                    const cplx<R> p00 = ld(psi, n00), pp0 = ld(psi, n00 + 1), p0p = ld(psi, n0p),
                                  ppp = ld(psi, n0p + 1);
                    const long ka00 = i + (long)g.Nxa * j, ka0p = ka00 + g.Nxa;
                    const long kb00 = g.Na + i + (long)g.Nxb * j, kbp0 = kb00 + 1;
                    const double a0 = double(ab[ka00]), a1 = double(ab[ka0p]), b0 = double(ab[kb00]),
                                 bp = double(ab[kbp0]);
                    const double v = cell_winding(g, dxdyH, atan2(double(p00.im), double(p00.re)),
                                                  atan2(double(pp0.im), double(pp0.re)),
                                                  atan2(double(p0p.im), double(p0p.re)),
                                                  atan2(double(ppp.im), double(ppp.re)), a0, a1, b0, bp);
                    if (q < capacity) {
                        double *o = rec + q * VREC;
                        o[0] = double(i + (long)g.Nxc * j);
                        o[1] = v;
                        o[2] = double(p00.re); o[3] = double(p00.im);
                        o[4] = double(pp0.re); o[5] = double(pp0.im);
                        o[6] = double(ppp.re); o[7] = double(ppp.im);
                        o[8] = double(p0p.re); o[9] = double(p0p.im);
                        o[10] = a0 + (abi ? double(abi[ka00]) : 0.0);
                        o[11] = a1 + (abi ? double(abi[ka0p]) : 0.0);
                        o[12] = b0 + (abi ? double(abi[kb00]) : 0.0);
                        o[13] = bp + (abi ? double(abi[kbp0]) : 0.0);
                    }
                }
            }
        }
    });
}

// Sub-cell vortex position in each flagged cell, one thread per record of
// k_vortex_emit (vortex_detector.py:103-147, with find_zero,
// find_zero_line and find_intersection of :75-101): the corner values are
// turned into a common gauge, the zero lines of Re and Im are taken where
// the linearised field changes sign along the cell border, and the vortex
// is their intersection. Writes (x, y, round(v), keep) per record; keep is
// 1 when both zero lines have exactly two points, the lines are not
// parallel (angle folded into [0, pi/2) above 1e-10) and the point lies in
// the cell's 3x3 neighbourhood, as the reference's tests. FP64 throughout,
// with the reference's operation order.
// find_zero(x2, y2, f2, x1, y1, f1) when f changes sign from point 1 to 2
// (also the closing edge: find_zero(x1, y1, f1, x4, y4, f4) is this with
// point 4 first); the first two points are kept, k counts them all.
__device__ __forceinline__ void zero_on_border(double x1, double y1, double f1, double x2, double y2, double f2,
                                               int &k, double (&px)[2], double (&py)[2]) {
    if (f2 * f1 < -1e-10) {
        if (k < 2) {
            px[k] = (f1 * x2 - x1 * f2) / (f1 - f2);
            py[k] = (f1 * y2 - y1 * f2) / (f1 - f2);
        }
        ++k;
    }
}

// psi * exp(-i th) with NumPy's complex product and exp (exp(0) = 1 exactly).
__device__ __forceinline__ void rotate_corner(double re, double im, double th, double &ore, double &oim) {
    double s, c;
    sincos(th, &s, &c);
    const double d = -s;
    ore = re * c - im * d;
    oim = re * d + im * c;
}

template <typename R>
__global__ void __launch_bounds__(256) k_vortex_triangulate(Geo g, const double *rec, int count, double *xyv) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= count) return;
    const double *r = rec + (long)k * VREC;
    const long n = (long)r[0];
    const long i = n % g.Nxc, j = n / g.Nxc;
    const double x = g.dx * double(i), y = g.dy * double(j);
    const double inta_00 = g.dx * r[10], inta_0p = g.dx * r[11];
    const double intb_00 = g.dy * r[12], intb_p0 = g.dy * r[13];
    // corners 00, p0, pp, 0p in the gauge of corner 00
    double re[4], im[4];
    re[0] = r[2];
    im[0] = r[3];
    rotate_corner(r[4], r[5], 0.75 * inta_00 + 0.25 * (intb_00 + inta_0p - intb_p0), re[1], im[1]);
    rotate_corner(r[6], r[7], 0.5 * (inta_00 + intb_p0) + 0.5 * (intb_00 + inta_0p), re[2], im[2]);
    rotate_corner(r[8], r[9], 0.75 * intb_00 + 0.25 * (inta_00 + intb_p0 - inta_0p), re[3], im[3]);
    const double cx[4] = {x, x + g.dx, x + g.dx, x}, cy[4] = {y, y, y + g.dy, y + g.dy};
    double rx[2] = {0.0, 0.0}, ry[2] = {0.0, 0.0}, qx[2] = {0.0, 0.0}, qy[2] = {0.0, 0.0};
    int nr = 0, nq = 0;
    // edges 1-2, 2-3, 3-4, 4-1 of the border, in the reference's order
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const int a = e, b = (e + 1) & 3;
        zero_on_border(cx[a], cy[a], re[a], cx[b], cy[b], re[b], nr, rx, ry);
        zero_on_border(cx[a], cy[a], im[a], cx[b], cy[b], im[b], nq, qx, qy);
    }
    double ix = 0.0, iy = 0.0, keep = 0.0;
    if (nr == 2 && nq == 2) {
        const double l1x1 = rx[0], l1y1 = ry[0], l1x2 = rx[1], l1y2 = ry[1];
        const double l2x1 = qx[0], l2y1 = qy[0], l2x2 = qx[1], l2y2 = qy[1];
        const double D = (l1x1 - l1x2) * (l2y1 - l2y2) - (l1y1 - l1y2) * (l2x1 - l2x2);
        double ph = atan2(D, (l1y1 - l1y2) * (l2y1 - l2y2) - (l1x1 - l1x2) * (l2x1 - l2x2));
        ph = fmod(fabs(ph), 1.5707963267948966);  // np.mod(|ph|, 0.5*np.pi)
        if (fabs(ph) > 1e-10) {
            ix = ((l1x1 * l1y2 - l1y1 * l1x2) * (l2x1 - l2x2) - (l1x1 - l1x2) * (l2x1 * l2y2 - l2y1 * l2x2)) / D;
            iy = ((l1x1 * l1y2 - l1y1 * l1x2) * (l2y1 - l2y2) - (l1y1 - l1y2) * (l2x1 * l2y2 - l2y1 * l2x2)) / D;
            if (x - g.dx < ix && ix < x + 2.0 * g.dx && y - g.dy < iy && iy < y + 2.0 * g.dy) keep = 1.0;
        }
    }
    double *o = xyv + 4L * k;
    o[0] = ix;
    o[1] = iy;
    o[2] = rint(r[1]);
    o[3] = keep;
}

template <typename R>
__global__ void __launch_bounds__(256) k_phase_lock(cplx<R> *psi, int lockN, const int *ns) {
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l < lockN) {
        const int n = ns[l];
        cplx<R> p = ld(psi, n);
        p.re = hypot(p.re, p.im);
        p.im = R(0);
        st(psi, n, p);
    }
}

}  // namespace glx
