// Device-side building blocks shared by every glx kernel (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "glx_limits.hpp"

namespace glx {

constexpr int WAVE = 64;
// Columns owned by one wave in the row-marching kernels: lanes 0 and 63 carry
// the columns left and right of the strip (halos), lanes 1..62 own 62
// consecutive columns.
constexpr int STRIP = WAVE - 2;

template <typename R> struct cplx { R re, im; };

// Index of the wave within its workgroup, as a wave-uniform (scalar) value.
__device__ __forceinline__ int wave_id() { return __builtin_amdgcn_readfirstlane(threadIdx.x / WAVE); }

template <typename R> __device__ __forceinline__ cplx<R> ld(const cplx<R> *p, long n);
template <> __device__ __forceinline__ cplx<double> ld(const cplx<double> *p, long n) {
    double2 v = reinterpret_cast<const double2 *>(p)[n];
    return {v.x, v.y};
}
template <> __device__ __forceinline__ cplx<float> ld(const cplx<float> *p, long n) {
    float2 v = reinterpret_cast<const float2 *>(p)[n];
    return {v.x, v.y};
}
template <typename R> __device__ __forceinline__ void st(cplx<R> *p, long n, cplx<R> v);
template <> __device__ __forceinline__ void st(cplx<double> *p, long n, cplx<double> v) {
    reinterpret_cast<double2 *>(p)[n] = make_double2(v.re, v.im);
}
template <> __device__ __forceinline__ void st(cplx<float> *p, long n, cplx<float> v) {
    reinterpret_cast<float2 *>(p)[n] = make_float2(v.re, v.im);
}

// A wave-uniform value kept in a vector register: the row kernels hold many
// loop-invariant scalars, and the scalar register file overflows into
// lane spills inside the hot loops if all of them live there.
template <typename T> __device__ __forceinline__ T in_vgpr(T x) {
    asm("" : "+v"(x));
    return x;
}

// Element at byte offset off (a per-lane 32-bit value) from a wave-uniform
// row pointer: the compiler keeps base + lane offset per array in registers
// and adds the scalar row offset with one vector instruction per access.
template <typename T> __device__ __forceinline__ T at(const T *row, uint32_t off) {
    return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(row) + off);
}
template <typename T> __device__ __forceinline__ void st_at(T *row, uint32_t off, T v) {
    *reinterpret_cast<T *>(reinterpret_cast<char *>(row) + off) = v;
}
template <typename R> __device__ __forceinline__ cplx<R> ld_at(const cplx<R> *row, uint32_t off) {
    return ld(reinterpret_cast<const cplx<R> *>(reinterpret_cast<const char *>(row) + off), 0);
}
template <typename R> __device__ __forceinline__ void st_at(cplx<R> *row, uint32_t off, cplx<R> v) {
    st(reinterpret_cast<cplx<R> *>(reinterpret_cast<char *>(row) + off), 0, v);
}

// sin and cos of one link phase. Link phases d*A are tens of radians at
// most, so a bounded-range method keeps the register footprint small:
// Cody-Waite reduction by pi/2 with a three-part constant (each leading part
// has few enough significant bits that k*part is exact, for |k| < 2^20 in
// double and |k| < 2^12 in float), then the Taylor series of sin and cos on
// [-pi/4, pi/4] in Horner form. The result is exactly odd/even symmetric:
// sincos_r(-x) = (-sin x, cos x), so U(-x) = conj(U(x)) holds bitwise.
__host__ __device__ __forceinline__ void sincos_r(double x, double *s, double *c) {
    const double k = rint(x * 0.6366197723675814);  // x * 2/pi
    double r = fma(-k, 1.5707963267341256, x);      // pi/2 = p1 + p2 + p3
    r = fma(-k, 6.077100506303966e-11, r);
    r = fma(-k, 2.0222662487959506e-21, r);
    const double z = r * r;
    double ps = 1.0 / 355687428096000.0;             // 1/17!
    ps = fma(ps, z, -1.0 / 1307674368000.0);         // 1/15!
    ps = fma(ps, z, 1.0 / 6227020800.0);             // 1/13!
    ps = fma(ps, z, -1.0 / 39916800.0);              // 1/11!
    ps = fma(ps, z, 1.0 / 362880.0);                 // 1/9!
    ps = fma(ps, z, -1.0 / 5040.0);                  // 1/7!
    ps = fma(ps, z, 1.0 / 120.0);                    // 1/5!
    ps = fma(ps, z, -1.0 / 6.0);                     // 1/3!
    const double sr = fma(r * z, ps, r);
    double pc = -1.0 / 6402373705728000.0;           // 1/18!
    pc = fma(pc, z, 1.0 / 20922789888000.0);         // 1/16!
    pc = fma(pc, z, -1.0 / 87178291200.0);           // 1/14!
    pc = fma(pc, z, 1.0 / 479001600.0);              // 1/12!
    pc = fma(pc, z, -1.0 / 3628800.0);               // 1/10!
    pc = fma(pc, z, 1.0 / 40320.0);                  // 1/8!
    pc = fma(pc, z, -1.0 / 720.0);                   // 1/6!
    pc = fma(pc, z, 1.0 / 24.0);                     // 1/4!
    pc = fma(pc, z, -0.5);                           // 1/2!
    const double cr = fma(z, pc, 1.0);
    const int q = (int)((long long)k & 3);
    const double sv = (q & 1) ? cr : sr, cv = (q & 1) ? sr : cr;
    *s = (q & 2) ? -sv : sv;
    *c = ((q + 1) & 2) ? -cv : cv;
}

__host__ __device__ __forceinline__ void sincos_r(float x, float *s, float *c) {
    const float k = rintf(x * 0.63661975f);
    float r = fmaf(-k, 1.5703125f, x);
    r = fmaf(-k, 4.837512969970703e-4f, r);
    r = fmaf(-k, 7.549790126404332e-08f, r);
    // Taylor terms through r^9 (sin) and r^10 (cos): the first omitted terms,
    // (pi/4)^11/11! = 1.7e-9 and (pi/4)^12/12! = 1.1e-10, are far below half
    // an FP32 ulp
    const float z = r * r;
    float ps = 1.0f / 362880.0f;
    ps = fmaf(ps, z, -1.0f / 5040.0f);
    ps = fmaf(ps, z, 1.0f / 120.0f);
    ps = fmaf(ps, z, -1.0f / 6.0f);
    const float sr = fmaf(r * z, ps, r);
    float pc = -1.0f / 3628800.0f;
    pc = fmaf(pc, z, 1.0f / 40320.0f);
    pc = fmaf(pc, z, -1.0f / 720.0f);
    pc = fmaf(pc, z, 1.0f / 24.0f);
    pc = fmaf(pc, z, -0.5f);
    const float cr = fmaf(z, pc, 1.0f);
    // quadrant q = k mod 4: swap on odd q, sign of sin from bit 1 of q, sign
    // of cos from bit 1 of q + 1, applied to the sign bit directly
    const uint32_t q = (uint32_t)(int)k;
    const bool odd = (q & 1u) != 0u;
    const float sv = odd ? cr : sr, cv = odd ? sr : cr;
    *s = __builtin_bit_cast(float, __builtin_bit_cast(uint32_t, sv) ^ ((q << 30) & 0x80000000u));
    *c = __builtin_bit_cast(float, __builtin_bit_cast(uint32_t, cv) ^ (((q + 1u) << 30) & 0x80000000u));
}

// Thomas Wang's 32-bit integer hash and the two Langevin streams built on
// it (the Langevin noise of svirl/cuda/common.h:36-63).
__device__ __forceinline__ uint32_t wang_hash(uint32_t s) {
    s = (s ^ 61u) ^ (s >> 16);
    s *= 9u;
    s = s ^ (s >> 4);
    s *= 0x27d4eb2du;
    s = s ^ (s >> 15);
    return s;
}
template <typename R> __device__ __forceinline__ R noise_1(uint32_t n, uint32_t t) {
    return R(2.3283064365386962890625e-10 * double(R(wang_hash(71u * n + 9887u * t))));
}
template <typename R> __device__ __forceinline__ R noise_2(uint32_t n, uint32_t t) {
    return R(2.3283064365386962890625e-10 * double(R(wang_hash(73u * n + 9901u * t + 1u))));
}

// Scaled Jacobi residual of the stop rule (svirl/cuda/td.h:123-131):
// int(min(1e4 * r / stop_eps, 1e8)) with the quotient in double and, as in
// the reference, stored in the working precision before the conversion.
// r is the largest |delta| (non-decreasing in r, so it may be applied to a
// maximum); stop_eps is already rounded to the working precision.
template <typename R> __device__ __forceinline__ int scaled_residual(R r, double stop_eps);
template <> __device__ __forceinline__ int scaled_residual<double>(double r, double stop_eps) {
    double v = 1.0e4 * r / stop_eps;
    if (v > 1.0e8) v = 1.0e8;
    return int(v);
}
template <> __device__ __forceinline__ int scaled_residual<float>(float r, double stop_eps) {
    float v = float(1.0e4 * double(r) / stop_eps);
    if (v > 1.0e8f) v = 1.0e8f;
    return int(v);
}

// Lane i receives lane i-1 (up) or lane i+1 (down) of a wave64, by DPP
// whole-wave shifts; the edge lane that has no source gets 0 (bound_ctrl:
// the hardware writes the zero, so no register has to be cleared first).
__device__ __forceinline__ int dpp_up1(int v) { return __builtin_amdgcn_mov_dpp(v, 0x138, 0xF, 0xF, true); }
__device__ __forceinline__ int dpp_dn1(int v) { return __builtin_amdgcn_mov_dpp(v, 0x130, 0xF, 0xF, true); }
__device__ __forceinline__ int shfl_up1(int v) { return dpp_up1(v); }
__device__ __forceinline__ int shfl_dn1(int v) { return dpp_dn1(v); }
__device__ __forceinline__ float shfl_up1(float v) { return __int_as_float(dpp_up1(__float_as_int(v))); }
__device__ __forceinline__ float shfl_dn1(float v) { return __int_as_float(dpp_dn1(__float_as_int(v))); }
__device__ __forceinline__ double shfl_up1(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = dpp_up1((int)(b & 0xffffffffll)), hi = dpp_up1((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
__device__ __forceinline__ double shfl_dn1(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = dpp_dn1((int)(b & 0xffffffffll)), hi = dpp_dn1((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// ---- wave64 / workgroup reductions ---------------------------------------

__device__ __forceinline__ int wave_max_int(int v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = max(v, __shfl_xor(v, off, WAVE));
    return v;
}
// Workgroups are dispatched to the NXCD accelerator dies round-robin by
// blockIdx, and each die has its own L2. The row-marching kernels number
// their work by xcd_block(), which gives the workgroups of one die a contiguous
// range, so that waves sharing halo rows, halo columns and cache lines of
// adjacent strips mostly sit on the same die and share its L2.
constexpr int NXCD = 8;
__device__ __forceinline__ int xcd_block() {
    const int G = gridDim.x, b = blockIdx.x;
    if (G % NXCD) return b;
    return (b % NXCD) * (G / NXCD) + b / NXCD;
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, WAVE);
    return v;
}

// Sum of one double per thread over the workgroup; result valid in thread 0.
// `lds` needs blockDim.x/64 doubles.
__device__ __forceinline__ double block_sum(double v, double *lds) {
    v = wave_sum(v);
    const int lane = threadIdx.x & (WAVE - 1), wid = threadIdx.x / WAVE;
    const int nw = (blockDim.x + WAVE - 1) / WAVE;
    __syncthreads();
    if (lane == 0) lds[wid] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0)
        for (int k = 0; k < nw; ++k) s += lds[k];
    return s;
}

// Maximum of one non-negative double per thread over the workgroup; result
// valid in thread 0. `lds` needs blockDim.x/64 doubles.
__device__ __forceinline__ double block_max(double v, double *lds) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, WAVE));
    const int lane = threadIdx.x & (WAVE - 1), wid = threadIdx.x / WAVE;
    const int nw = (blockDim.x + WAVE - 1) / WAVE;
    __syncthreads();
    if (lane == 0) lds[wid] = v;
    __syncthreads();
    double m = 0.0;
    if (threadIdx.x == 0)
        for (int k = 0; k < nw; ++k) m = fmax(m, lds[k]);
    return m;
}

// Maximum of one int per thread over the workgroup; valid in thread 0.
__device__ __forceinline__ int block_max_int(int v, int *lds) {
    v = wave_max_int(v);
    const int lane = threadIdx.x & (WAVE - 1), wid = threadIdx.x / WAVE;
    const int nw = (blockDim.x + WAVE - 1) / WAVE;
    __syncthreads();
    if (lane == 0) lds[wid] = v;
    __syncthreads();
    int m = 0;
    if (threadIdx.x == 0)
        for (int k = 0; k < nw; ++k) m = max(m, lds[k]);
    return m;
}

// ---- agent-scope atomics ---------------------------------------------------

__device__ __forceinline__ void atomic_max_agent(int *p, int v) {
    __hip_atomic_fetch_max(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// ---- grid barrier with a max-reduction, for persistent kernels ----------
//
// Arrival: each workgroup makes its stores visible device-wide (release
// fence), then writes (epoch << 32 | value) into its own word of `arrive`.
// Workgroup 0 polls all arrival words with plain device-coherent loads (all
// of its threads in parallel), takes the max of the values once every word
// carries the current epoch, and writes (epoch << 32 | max) into NREL release
// words; every other workgroup polls one release word. No two workgroups
// contend for one atomic, polls never invalidate the caches (one acquire
// fence follows a successful poll), and the cost is a few memory round trips
// whatever the number of workgroups. A poll that waits longer than
// BARRIER_TIMEOUT_TICKS sets `err`, which every poller also watches, so a
// grid that is not co-resident fails instead of hanging.
constexpr unsigned long long BARRIER_TIMEOUT_TICKS = 100ull * 1000 * 1000 * 5;  // 5 s at 100 MHz

struct GridSync {
    unsigned long long *arrive;  // MAX_GRID_BLOCKS words
    unsigned long long *rel;     // NREL words, REL_STRIDE apart
    int *err;                    // 1 flag
};

__device__ __forceinline__ unsigned long long ld_relaxed(const unsigned long long *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ int ld_relaxed(const int *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// A team: the workgroups [b0, b0 + nb) that sweep one grid together (the
// whole launch for a single instance, a slice of it for one instance of a
// batch). Its master is workgroup b0; its release words are rel[k *
// REL_STRIDE], k < nrel.
struct Team {
    int b0, nb, nrel;
    unsigned long long *rel;
};

// Call from every thread of every workgroup of the team the same number of
// times with the same `epoch` (barriers the team has passed; incremented
// here). `v` is the workgroup's value (thread 0's is used): NF fields of
// 32 / NF bits, at most 10 (three 10-bit fields at bits 0-9, 10-19, 20-29;
// four 8-bit fields). On success returns true and sets `vmax` (all threads)
// to the field-wise max over the team; returns false when a poll timed out
// anywhere in the grid. A one-workgroup team synchronises in the workgroup
// only.
template <int NF = 3>
__device__ __forceinline__ bool team_sync_max(const GridSync &gs, const Team &tm, unsigned long long &epoch, int v,
                                              int &vmax) {
    constexpr int FB = NF >= 4 ? 8 : 10;
    constexpr unsigned long long FM = (1ull << FB) - 1ull;
    __shared__ int s_red[16];
    __shared__ int s_ok, s_max;
    ++epoch;
    if (tm.nb == 1) {
        __syncthreads();
        if (threadIdx.x == 0) s_max = v;
        __syncthreads();
        vmax = s_max;
        return true;
    }
    __builtin_amdgcn_s_waitcnt(0);  // this wave's stores have left the wave
    __syncthreads();
    const unsigned long long tag = epoch << 32;
    const int local = (int)blockIdx.x - tm.b0;
    if (threadIdx.x == 0) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        __hip_atomic_store(gs.arrive + blockIdx.x, tag | (unsigned)v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    if (local == 0) {
        const unsigned long long t0 = wall_clock64();
        int ok = 1, mf[NF];
        for (;;) {
            int all = 1;
#pragma unroll
            for (int f = 0; f < NF; ++f) mf[f] = 0;
            for (int k = threadIdx.x; k < tm.nb; k += blockDim.x) {
                const unsigned long long a = ld_relaxed(gs.arrive + tm.b0 + k);
                if ((a >> 32) != epoch) all = 0;
                else {
#pragma unroll
                    for (int f = 0; f < NF; ++f) mf[f] = max(mf[f], (int)(unsigned)((a >> (FB * f)) & FM));
                }
            }
            int abort = 0;
            if (threadIdx.x == 0) abort = (wall_clock64() - t0 > BARRIER_TIMEOUT_TICKS) || ld_relaxed(gs.err);
            if (__syncthreads_or(abort)) { ok = 0; break; }
            if (__syncthreads_and(all)) break;
            __builtin_amdgcn_s_sleep(1);
        }
        if (ok) {  // uniform over the workgroup
            int packed = 0;
#pragma unroll
            for (int f = 0; f < NF; ++f) packed |= block_max_int(mf[f], s_red) << (FB * f);
            if (threadIdx.x == 0) s_max = packed;
            __syncthreads();
            if (threadIdx.x < tm.nrel) {
                __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "agent");
                __hip_atomic_store(tm.rel + threadIdx.x * REL_STRIDE, tag | (unsigned)s_max, __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
            }
        } else if (threadIdx.x == 0) {
            __hip_atomic_store(gs.err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        if (threadIdx.x == 0) s_ok = ok;
    } else if (threadIdx.x == 0) {
        const unsigned long long *r = tm.rel + (local % tm.nrel) * REL_STRIDE;
        const unsigned long long t0 = wall_clock64();
        int ok = 1;
        unsigned long long a;
        while (((a = ld_relaxed(r)) >> 32) != epoch) {
            if (ld_relaxed(gs.err)) { ok = 0; break; }
            if (wall_clock64() - t0 > BARRIER_TIMEOUT_TICKS) {
                __hip_atomic_store(gs.err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                ok = 0;
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
        s_ok = ok;
        s_max = (int)(unsigned)(a & 0xffffffffull);
    }
    if (threadIdx.x == 0) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    __syncthreads();
    vmax = s_max;
    return s_ok != 0;
}

// The whole launch as one team.
__device__ __forceinline__ bool grid_sync_max(const GridSync &gs, unsigned long long &epoch, int v, int &vmax) {
    const Team tm{0, (int)gridDim.x, NREL, gs.rel};
    return team_sync_max(gs, tm, epoch, v, vmax);
}

}  // namespace glx
