// Launcher bodies, instantiated once per precision by glx_f64.hip and
// glx_f32.hip.
#pragma once

#include <algorithm>
#include <mutex>
#include <unordered_map>
#include <vector>

#include "glx_fields.hpp"
#include "glx_launch.hpp"
#include "glx_td.hpp"
#include "glx_batch_cg.hpp"
#include "glx_cg_rows.hpp"

namespace glx {

// Waves of `fn` (256-thread workgroups) resident on the chip at once, from
// its occupancy, queried once per kernel; 0 if the query fails.
static long resident_waves(const void *fn, int num_cus) {
    static std::mutex mu;
    static std::unordered_map<const void *, int> per_cu;
    int per = 0;
    {
        std::lock_guard<std::mutex> lk(mu);
        auto it = per_cu.find(fn);
        if (it != per_cu.end()) {
            per = it->second;
        } else {
            if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, fn, RED_BLOCK, 0) != hipSuccess) per = 0;
            per_cu[fn] = per;
        }
    }
    return (long)per * num_cus * (RED_BLOCK / WAVE);
}

// Waves of a row-marching pass cut to whole rounds of the chip: W is
// lowered to a multiple of the waves the kernel keeps resident (at most
// max_per_cu workgroups per CU when that is given), so that no round runs
// a fraction of the chip.
static long fit_row_waves(const void *fn, long W, int num_cus, int max_per_cu = 0) {
    long res = resident_waves(fn, num_cus);
    if (max_per_cu > 0) res = std::min(res, (long)max_per_cu * num_cus * (RED_BLOCK / WAVE));
    if (res <= 0 || W <= res) return W;
    return (W / res) * res;
}


// Nodes are addressed with 32-bit offsets: Nx * Ny is capped at 2^30.
inline bool grid_ok(const glx_grid &g) {
    return g.Nx >= 2 && g.Ny >= 2 && (long)g.Nx * g.Ny <= (1L << 30) && g.dx > 0.0 && g.dy > 0.0;
}

inline int grid_blocks(long work, int block = RED_BLOCK, int cap = 65536) {
    long b = (work + block - 1) / block;
    if (b < 1) b = 1;
    if (b > cap) b = cap;
    return (int)b;
}

#define GLX_LAUNCH_CHECK(what) \
    do { hipError_t _e = hipGetLastError(); if (_e != hipSuccess) return check_hip(_e, what); } while (0)

// Copy `n` doubles of e.result to the host after the stream drains.
inline int fetch_result(const Env &e, double *out, int n) {
    hipError_t st = hipMemcpyAsync(out, e.result, sizeof(double) * n, hipMemcpyDeviceToHost, e.stream);
    if (st != hipSuccess) return check_hip(st, "hipMemcpyAsync(result)");
    return check_hip(hipStreamSynchronize(e.stream), "hipStreamSynchronize");
}

// Waves used by a single-sweep launch: about 16 rows per wave.
inline long sweep_waves(const glx_grid &g) {
    const long T = (long)nstrips_of<1>(g.Nx) * g.Ny;
    long W = (T + 15) / 16;
    return W < 1 ? 1 : W;
}

template <typename R>
static PsiSweepArgs<R> make_psi_args(const glx_grid &g, double dt, double eps, const void *eps_map,
                                     const uint8_t *mt, const void *ab, double stop_eps, double langevin_c) {
    const Geo G = make_geo(g.Nx, g.Ny, g.dx, g.dy);
    PsiSweepArgs<R> A{};
    A.Nx = g.Nx; A.Ny = g.Ny;
    A.dx = R(g.dx); A.dy = R(g.dy); A.idx2 = R(G.idx2); A.idy2 = R(G.idy2);
    A.dt = R(dt); A.eps = R(eps);
    A.eps_map = (const R *)eps_map;
    A.mt = mt;
    A.ab = (const R *)ab;
    A.langevin_c = R(langevin_c);
    A.stop_eps = double(R(stop_eps));
    return A;
}

template <typename R>
static ASweepArgs<R> make_a_args(const glx_grid &g, double dt, double kappa2, double rho, double H,
                                 const uint8_t *mt, double stop_eps, double langevin_c) {
    const Geo G = make_geo(g.Nx, g.Ny, g.dx, g.dy);
    ASweepArgs<R> A{};
    A.Nx = g.Nx; A.Ny = g.Ny;
    A.dx = R(g.dx); A.dy = R(g.dy); A.idx = R(G.idx); A.idy = R(G.idy);
    A.idx2 = R(G.idx2); A.idy2 = R(G.idy2); A.idxy = R(G.idxy);
    // scalar constants in the reference's expression order (td.h:338, 376, 403, 422, 449)
    const double dt_rho = double(R(dt)) * double(R(rho));
    const double dt_rho_R = double(R(dt_rho));
    const double dt_rho_kappa2 = dt_rho_R * double(R(kappa2));
    A.dt_rho = R(dt_rho);
    A.dt_rho_kappa2 = R(dt_rho_kappa2);
    const double k2 = double(R(kappa2)), HH = double(R(H));
    A.rh_a = R(2.0 * k2 * HH * double(A.idy));
    A.rh_b = R(2.0 * k2 * HH * double(A.idx));
    const double drk = double(A.dt_rho_kappa2);
    A.den_a = R(1.0 + 2.0 * drk * double(A.idy2));
    A.den_b = R(1.0 + 2.0 * drk * double(A.idx2));
    A.iden_a = R(1.0 / double(A.den_a));
    A.iden_b = R(1.0 / double(A.den_b));
    A.mt = mt;
    A.langevin_c = R(langevin_c);
    A.stop_eps = double(R(stop_eps));
    return A;
}

template <typename R>
int L<R>::psi_sweep(const Env &e, glx_grid g, double dt, double eps, const void *eps_map, const uint8_t *mt,
                    const void *ab, void *psi_rhs, const void *psi, void *psi_next, double langevin_c,
                    uint32_t jstep, uint32_t rand_t, double stop_eps, int32_t *r2max) {
    if (!grid_ok(g) || !psi_rhs || !psi || !psi_next || !r2max || !ab || stop_eps <= 0.0)
        return set_error(GLX_E_ARG, "glx_td_psi_sweep: invalid arguments");
    PsiSweepArgs<R> A = make_psi_args<R>(g, dt, eps, eps_map, mt, ab, stop_eps, langevin_c);
    A.psi = (const cplx<R> *)psi;
    A.psi_next = (cplx<R> *)psi_next;
    A.rhs = (const cplx<R> *)psi_rhs;
    A.rhs_out = (langevin_c > 1.0e-32 && jstep == 0) ? (cplx<R> *)psi_rhs : nullptr;
    A.jstep = jstep;
    A.rand_t = rand_t;
    const long W = sweep_waves(g);
    hipLaunchKernelGGL(k_psi_sweep<R>, dim3((unsigned)((W + 3) / 4)), dim3(256), 0, e.stream, A, (int *)r2max, W);
    GLX_LAUNCH_CHECK("k_psi_sweep");
    return GLX_OK;
}

template <typename R>
int L<R>::a_sweep(const Env &e, glx_grid g, double dt, double kappa2, double rho, double H, const uint8_t *mt,
                  const void *psi, const void *src, void *ab_rhs, const void *ab, void *ab_next, double langevin_c,
                  uint32_t jstep, uint32_t rand_t, double stop_eps, int32_t *r2max) {
    if (!grid_ok(g) || !psi || !src || !ab_rhs || !ab || !ab_next || !r2max || stop_eps <= 0.0)
        return set_error(GLX_E_ARG, "glx_td_a_sweep: invalid arguments");
    if (ab == ab_next) return set_error(GLX_E_ARG, "glx_td_a_sweep: ab and ab_next must differ");
    ASweepArgs<R> A = make_a_args<R>(g, dt, kappa2, rho, H, mt, stop_eps, langevin_c);
    A.psi = (const cplx<R> *)psi;
    A.src = (const R *)src;
    A.rhs = (const R *)ab_rhs;
    A.rhs_out = (langevin_c > 1.0e-32 && jstep == 0) ? (R *)ab_rhs : nullptr;
    A.ab = (const R *)ab;
    A.ab_next = (R *)ab_next;
    A.jstep = jstep;
    A.rand_t = rand_t;
    const long W = sweep_waves(g);
    hipLaunchKernelGGL(k_a_sweep<R>, dim3((unsigned)((W + 3) / 4)), dim3(256), 0, e.stream, A, (int *)r2max, W);
    GLX_LAUNCH_CHECK("k_a_sweep");
    return GLX_OK;
}

// The FP32-mixed run stores A as A - A_gauge(H) while it evolves A (exact
// change of variable: the symmetric gauge is a fixed point of the A sweep
// with psi = 0), so that the iterates stay small and float-resolved. FP64
// runs, and FP32 runs that do not evolve A, keep the plain form (A is then
// left bit-for-bit untouched).
template <typename R> struct RunForm { static constexpr bool has_gauge = false; };
template <> struct RunForm<float> { static constexpr bool has_gauge = true; };

// Co-resident workgroups per CU of the psi solve kernel and of the A solve
// kernels (the smaller of the plain and fixed-source forms).
template <typename R, bool G> static int solve_occupancy2(int *psi_per_cu, int *a_per_cu) {
    int a = 0, b = 0, c = 1 << 20;
    hipError_t st = hipOccupancyMaxActiveBlocksPerMultiprocessor(&a, k_td_psi_solve<R, G>, 256, 0);
    if (st == hipSuccess) st = hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, k_td_a_solve<R, G>, 256, 0);
    if constexpr (!G)  // the fixed-source A solve runs in the plain form only
        if (st == hipSuccess) st = hipOccupancyMaxActiveBlocksPerMultiprocessor(&c, k_td_a_solve<R, G, true>, 256, 0);
    if (st != hipSuccess) return check_hip(st, "hipOccupancyMaxActiveBlocksPerMultiprocessor");
    *psi_per_cu = a;
    *a_per_cu = b < c ? b : c;
    return GLX_OK;
}

template <typename R, bool G> static int solve_occupancy(int *per_cu) {
    int p = 0, q = 0;
    int rc = solve_occupancy2<R, G>(&p, &q);
    if (rc) return rc;
    *per_cu = p < q ? p : q;
    return GLX_OK;
}

// The co-resident limits of the two persistent solves separately (each
// launch only needs its own grid resident): workgroups on the chip.
template <typename R> static int solve_grids(const Env &e, int32_t *psi_max, int32_t *a_max) {
    static int pp = 0, pa = 0;
    if (pp <= 0 || pa <= 0) {
        int p = 0, a = 0;
        int rc = solve_occupancy2<R, false>(&p, &a);
        if (rc) return rc;
        if constexpr (RunForm<R>::has_gauge) {
            int pg = 0, ag = 0;
            rc = solve_occupancy2<R, true>(&pg, &ag);
            if (rc) return rc;
            if (pg < p) p = pg;
            if (ag < a) a = ag;
        }
        pp = p;
        pa = a;
    }
    *psi_max = std::min(pp * e.num_cus, MAX_GRID_BLOCKS);
    *a_max = std::min(pa * e.num_cus, MAX_GRID_BLOCKS);
    return GLX_OK;
}

template <typename R> int L<R>::td_max_grid(const Env &e, int32_t wg_size, int32_t *max_wg) {
    if (wg_size != 256 || !max_wg) return set_error(GLX_E_ARG, "glx_td_max_grid: wg_size must be 256");
    static int per_cu = 0;  // co-resident workgroups per CU of every solve kernel (queried once)
    if (per_cu <= 0) {
        int m = 0;
        int rc = solve_occupancy<R, false>(&m);
        if (rc) return rc;
        if constexpr (RunForm<R>::has_gauge) {
            int mg = 0;
            rc = solve_occupancy<R, true>(&mg);
            if (rc) return rc;
            if (mg < m) m = mg;
        }
        per_cu = m;
    }
    int m = per_cu * e.num_cus;
    *max_wg = m < MAX_GRID_BLOCKS ? m : MAX_GRID_BLOCKS;
    return GLX_OK;
}

template <typename R, bool G>
static int td_run_form(const Env &e, glx_grid g, const glx_td_params *p, const void *eps_map, const uint8_t *mt,
                       void *const P[3], int32_t *pc, void *const Q[3], int32_t *qc, void *rhs_psi, void *rhs_ab,
                       int32_t nsteps, int32_t *sweeps, int32_t wg, int32_t wg_a) {

    TDSolveArgs<R> T{};
    T.psi = make_psi_args<R>(g, p->dt, p->eps, eps_map, mt, nullptr, p->stop_psi, p->langevin_psi);
    T.A = make_a_args<R>(g, p->dt, p->kappa2, p->rho, p->H, mt, p->stop_A, p->langevin_A);
    const double Lx = g.dx * (g.Nx - 1), Ly = g.dy * (g.Ny - 1);
    const double gH = G ? p->H : 0.0;
    T.psi.gH = T.A.gH = gH;
    T.psi.gdx = T.A.gdx = g.dx;
    T.psi.gdy = T.A.gdy = g.dy;
    T.psi.gx0 = T.A.gx0 = 0.5 * Lx;
    T.psi.gy0 = T.A.gy0 = 0.5 * Ly;
    if (G) {  // the gauge absorbs the boundary H source of the A sweep
        T.A.rh_a = R(0);
        T.A.rh_b = R(0);
    }
    for (int k = 0; k < 3; ++k) { T.P[k] = (cplx<R> *)P[k]; T.Q[k] = (R *)Q[k]; }
    T.rhs_psi = p->langevin_psi > 1.0e-32 ? (cplx<R> *)rhs_psi : nullptr;
    T.rhs_ab = (p->solveA && p->langevin_A > 1.0e-32) ? (R *)rhs_ab : nullptr;
    T.max_sweeps = p->max_sweeps;
    T.a_src = (const R *)p->a_src;
    T.sc.gs.arrive = e.arrive;
    T.sc.gs.rel = e.rel;
    T.sc.gs.err = e.err;
    T.sc.state = e.state;
    T.sc.sweeps = e.sweeps;
    const long Nab = (long)(g.Nx - 1) * g.Ny + (long)g.Nx * (g.Ny - 1);

    hipLaunchKernelGGL(k_td_reset, dim3(1), dim3(256), 0, e.stream, T.sc, (int)*pc, (int)*qc, e.abort_next ? 1 : 0);
    GLX_LAUNCH_CHECK("k_td_reset");
    if (G) {
        hipLaunchKernelGGL(k_gauge_shift<R>, dim3(grid_blocks(Nab)), dim3(256), 0, e.stream, T.Q[*qc], g.Nx, g.Ny, gH,
                           g.dx, g.dy, 0.5 * Lx, 0.5 * Ly, -1.0);
        GLX_LAUNCH_CHECK("k_gauge_shift");
    }
    hipError_t est = hipEventRecord(e.kev0, e.stream);
    if (est != hipSuccess) return check_hip(est, "hipEventRecord");
    uint32_t rt = p->rand_t;
    const bool do_psi = p->eqn != 2, do_A = p->solveA && p->eqn != 1;
    const bool prof = e.td_profile && e.tev && e.tev_cap >= 4 * nsteps;
    int nev = 0;
    for (int step = 0; step < nsteps; ++step) {
        T.step = step;
        if (do_psi) {
            T.psi.rand_t = rt++;
            if (prof) (void)hipEventRecord(e.tev[nev++], e.stream);
            hipLaunchKernelGGL((k_td_psi_solve<R, G>), dim3((unsigned)wg), dim3(256), 0, e.stream, T);
            GLX_LAUNCH_CHECK("k_td_psi_solve");
            if (prof) (void)hipEventRecord(e.tev[nev++], e.stream);
        }
        if (do_A) {
            T.A.rand_t = rt++;
            if (prof) (void)hipEventRecord(e.tev[nev++], e.stream);
            bool fixed = false;
            if constexpr (!G) {
                if (T.a_src) {
                    hipLaunchKernelGGL((k_td_a_solve<R, G, true>), dim3((unsigned)wg_a), dim3(256), 0, e.stream, T);
                    fixed = true;
                }
            }
            if (!fixed) hipLaunchKernelGGL((k_td_a_solve<R, G>), dim3((unsigned)wg_a), dim3(256), 0, e.stream, T);
            GLX_LAUNCH_CHECK("k_td_a_solve");
            if (prof) (void)hipEventRecord(e.tev[nev++], e.stream);
        }
    }
    est = hipEventRecord(e.kev1, e.stream);
    if (est != hipSuccess) return check_hip(est, "hipEventRecord");
    if (G) {
        hipLaunchKernelGGL(k_gauge_shift_state<R>, dim3(grid_blocks(Nab)), dim3(256), 0, e.stream, T.Q[0], T.Q[1], T.Q[2],
                           e.state, g.Nx, g.Ny, gH, g.dx, g.dy, 0.5 * Lx, 0.5 * Ly, 1.0);
        GLX_LAUNCH_CHECK("k_gauge_shift_state");
    }
    int host_state[ST_WORDS] = {0};
    int host_err = 0;
    hipError_t st = hipMemcpyAsync(host_state, e.state, sizeof(int) * ST_WORDS, hipMemcpyDeviceToHost, e.stream);
    if (st == hipSuccess) st = hipMemcpyAsync(&host_err, e.err, sizeof(int), hipMemcpyDeviceToHost, e.stream);
    if (st == hipSuccess) st = hipMemcpyAsync(sweeps, e.sweeps, sizeof(int) * 2 * nsteps, hipMemcpyDeviceToHost, e.stream);
    if (st == hipSuccess) st = hipStreamSynchronize(e.stream);
    if (st != hipSuccess) return check_hip(st, "glx_td_run readback");
    float kms = 0.f;
    if (hipEventElapsedTime(&kms, e.kev0, e.kev1) == hipSuccess && e.last_kernel_ms) *e.last_kernel_ms = kms;
    if (e.solve_ms && e.solve_n) {
        e.solve_ms[0] = e.solve_ms[1] = 0.0;
        e.solve_n[0] = e.solve_n[1] = 0;
        if (prof) {
            int k = 0;
            for (int step = 0; step < nsteps; ++step)
                for (int kind = 0; kind < 2; ++kind) {
                    if ((kind == 0 && !do_psi) || (kind == 1 && !do_A)) continue;
                    float ms = 0.f;
                    if (hipEventElapsedTime(&ms, e.tev[k], e.tev[k + 1]) == hipSuccess) {
                        e.solve_ms[kind] += ms;
                        e.solve_n[kind] += 1;
                    }
                    k += 2;
                }
        }
    }
    if (host_err) return set_error(GLX_E_TIMEOUT, "glx_td_run: grid barrier timed out in step %d of %d", host_state[ST_ERRSTEP], nsteps);
    if (!do_psi)
        for (int k = 0; k < nsteps; ++k) sweeps[2 * k] = 0;
    if (!do_A)
        for (int k = 0; k < nsteps; ++k) sweeps[2 * k + 1] = 0;
    *pc = host_state[ST_PSI];
    *qc = host_state[ST_AB];
    if (e.passes)
        for (int k = 0; k < 5; ++k) e.passes[k] = host_state[ST_PASS_PSI + k];
    return GLX_OK;
}


template <typename R>
int L<R>::td_run(const Env &e, glx_grid g, const glx_td_params *p, const void *eps_map, const uint8_t *mt,
                 void *const P[3], int32_t *pc, void *const Q[3], int32_t *qc, void *rhs_psi, void *rhs_ab,
                 int32_t nsteps, int32_t *sweeps, int32_t wg, int32_t wg_size) {
    if (!grid_ok(g) || !p || !P || !pc || !sweeps || nsteps < 0 || *pc < 0 || *pc > 2)
        return set_error(GLX_E_ARG, "glx_td_run: invalid arguments");
    if (!P[0] || !P[1] || !P[2]) return set_error(GLX_E_ARG, "glx_td_run: psi buffers");
    if (!Q || !qc || *qc < 0 || *qc > 2 || !Q[0] || !Q[1] || !Q[2])
        return set_error(GLX_E_ARG, "glx_td_run: vector potential buffers");
    if (p->max_sweeps < 1 || p->stop_psi <= 0.0 || p->stop_A <= 0.0 || p->eqn < 0 || p->eqn > 2)
        return set_error(GLX_E_ARG, "glx_td_run: invalid parameters");
    if (p->langevin_psi > 1.0e-32 && !rhs_psi) return set_error(GLX_E_ARG, "glx_td_run: Langevin psi needs rhs_psi");
    if (p->solveA && p->langevin_A > 1.0e-32 && !rhs_ab) return set_error(GLX_E_ARG, "glx_td_run: Langevin A needs rhs_ab");
    if (nsteps == 0) return GLX_OK;
    if (nsteps > e.sweeps_cap) return set_error(GLX_E_ARG, "glx_td_run: nsteps %d exceeds %d per launch", nsteps, e.sweeps_cap);
    if (wg_size == 0) wg_size = 256;
    int32_t maxg = 0;
    int rc = td_max_grid(e, wg_size, &maxg);
    if (rc) return rc;
    // Default grids, the measured optima at 4097^2: two workgroups per CU
    // for the psi solve, three for the A solve, each within its own kernel's
    // co-resident limit (maxg, the smaller of the two, bounds an explicit wg,
    // which applies to both).
    int32_t wg_a = wg;
    if (wg <= 0) {
        int32_t psi_max = 0, a_max = 0;
        rc = solve_grids<R>(e, &psi_max, &a_max);
        if (rc) return rc;
        wg = std::min(psi_max, 2 * e.num_cus);
        wg_a = std::min(a_max, 3 * e.num_cus);
    }
    if (wg > maxg) return set_error(GLX_E_ARG, "glx_td_run: grid %d exceeds co-resident limit %d", wg, maxg);
    const bool evolves_A = p->solveA && p->eqn != 1;
    if (p->a_src && (p->a_src == Q[0] || p->a_src == Q[1] || p->a_src == Q[2]))
        return set_error(GLX_E_ARG, "glx_td_run: a_src must not be one of the rotating A buffers");
    // a fixed source is given in the plain form, so the run keeps that form
    if constexpr (RunForm<R>::has_gauge)
        if (evolves_A && !p->a_src)
            return td_run_form<R, true>(e, g, p, eps_map, mt, P, pc, Q, qc, rhs_psi, rhs_ab, nsteps, sweeps, wg, wg_a);
    return td_run_form<R, false>(e, g, p, eps_map, mt, P, pc, Q, qc, rhs_psi, rhs_ab, nsteps, sweeps, wg, wg_a);
}

// ---- batched TD ---------------------------------------------------------------

// Per-instance solve arguments, as td_run_form builds them for one grid.
template <typename R>
static TDSolveArgs<R> batch_args(const glx_batch_instance &I, bool gauge_form) {
    const glx_grid &g = I.g;
    const glx_td_params *p = &I.p;
    TDSolveArgs<R> T{};
    T.psi = make_psi_args<R>(g, p->dt, p->eps, I.eps_map, I.mt, nullptr, p->stop_psi, 0.0);
    T.A = make_a_args<R>(g, p->dt, p->kappa2, p->rho, p->H, I.mt, p->stop_A, 0.0);
    const bool G = gauge_form && p->solveA;
    const double gH = G ? p->H : 0.0;
    T.psi.gH = T.A.gH = gH;
    T.psi.gdx = T.A.gdx = g.dx;
    T.psi.gdy = T.A.gdy = g.dy;
    T.psi.gx0 = T.A.gx0 = 0.5 * g.dx * (g.Nx - 1);
    T.psi.gy0 = T.A.gy0 = 0.5 * g.dy * (g.Ny - 1);
    if (G) {
        T.A.rh_a = R(0);
        T.A.rh_b = R(0);
    }
    for (int k = 0; k < 3; ++k) { T.P[k] = (cplx<R> *)I.psi[k]; T.Q[k] = (R *)I.ab[k]; }
    T.rhs_psi = nullptr;
    T.rhs_ab = nullptr;
    T.max_sweeps = p->max_sweeps;
    T.solveA = p->solveA ? 1 : 0;
    return T;
}

// Teams of consecutive workgroups, sized to each instance's work (its
// wave-rows times its relative sweep count): the largest scale f with
// sum_i clamp(floor(f*work_i), 1, cap_i) <= G, where cap_i keeps at least
// min_rows rows per wave.
static void plan_teams(const std::vector<long> &rows, const std::vector<double> &work, int G, int min_rows,
                       std::vector<int> &nb) {
    const size_t n = rows.size();
    nb.assign(n, 1);
    auto total = [&](double f, std::vector<int> *out) {
        long t = 0;
        for (size_t i = 0; i < n; ++i) {
            const long cap = std::max(1L, rows[i] / (4L * min_rows));
            long k = (long)(f * work[i]);
            k = std::min(std::max(k, 1L), cap);
            if (out) (*out)[i] = (int)k;
            t += k;
        }
        return t;
    };
    double lo = 0.0, hi = 1.0;
    while (total(hi, nullptr) <= G && hi < 1e12) hi *= 2.0;
    for (int it = 0; it < 100; ++it) {
        const double mid = 0.5 * (lo + hi);
        if (total(mid, nullptr) <= G) lo = mid;
        else hi = mid;
    }
    total(lo, &nb);
}

template <typename R, bool G, int KB> static int batch_occupancy(int *per_cu) {
    hipError_t st = hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, k_batch_td_step<R, G, KB>, 256, 0);
    return st == hipSuccess ? GLX_OK : check_hip(st, "hipOccupancyMaxActiveBlocksPerMultiprocessor");
}

template <typename R, bool G, int KB>
static int batch_run_form(const Env &e, glx_batch_instance *inst, int32_t n, int32_t nsteps, int32_t *sweeps,
                          int32_t use_graph, int32_t wg_per_cu, int32_t min_rows, const double *weight,
                          const Workspace &ws) {
    int per_cu = 0;
    int rc = batch_occupancy<R, G, KB>(&per_cu);
    if (rc) return rc;
    if (wg_per_cu > 0 && wg_per_cu < per_cu) per_cu = wg_per_cu;
    const int Gb = std::min(per_cu * e.num_cus, MAX_GRID_BLOCKS);
    if (n > Gb) return set_error(GLX_E_ARG, "%s: %d instances exceed %d co-resident workgroups", ws.what, n, Gb);
    std::vector<long> rows(n);
    std::vector<double> work(n);
    for (int i = 0; i < n; ++i) {
        rows[i] = (long)nstrips_of<1>(inst[i].g.Nx) * inst[i].g.Ny;
        work[i] = (double)rows[i] * ((weight && weight[i] > 0.0) ? weight[i] : 1.0);
    }
    std::vector<int> nb;
    plan_teams(rows, work, Gb, min_rows > 0 ? min_rows : 16, nb);
    std::vector<int> team_of_block;
    std::vector<int2> teams(n);
    for (int i = 0; i < n; ++i) {
        teams[i] = make_int2((int)team_of_block.size(), nb[i]);
        for (int k = 0; k < nb[i]; ++k) team_of_block.push_back(i);
    }
    const int grid = (int)team_of_block.size();
    std::vector<TDSolveArgs<R>> args(n);
    std::vector<int> state(8 * (size_t)n, 0);
    for (int i = 0; i < n; ++i) {
        args[i] = batch_args<R>(inst[i], G);
        state[8 * i + ST_PSI] = inst[i].psi_cur;
        state[8 * i + ST_AB] = inst[i].ab_cur;
    }
    // one device block: args | team_of_block | teams | rel | state | sweeps
    auto al = [](size_t b) { return (b + 255) & ~size_t(255); };
    const size_t s_args = al(sizeof(TDSolveArgs<R>) * n), s_tob = al(sizeof(int) * grid), s_teams = al(sizeof(int2) * n);
    const size_t s_rel = al(sizeof(unsigned long long) * REL_STRIDE * n), s_state = al(sizeof(int) * 8 * n);
    const size_t s_sw = al(sizeof(int) * 2 * (size_t)n * nsteps);
    const size_t need = s_args + s_tob + s_teams + s_rel + s_state + s_sw;
    char *blk = nullptr;
    hipError_t st = hipSuccess;
    if (ws.ptr) {
        // caller-provided device workspace (the no-allocation entry point)
        if (need > ws.bytes)
            return set_error(GLX_E_ARG, "%s: %d instances x %d steps need %zu bytes of scratch, above its %zu-byte maximum",
                             ws.what, n, nsteps, need, ws.bytes);
        blk = (char *)ws.ptr;
    } else {
        st = hipMalloc((void **)&blk, need);
        if (st != hipSuccess) return check_hip(st, "hipMalloc(batch)");
    }
    struct Free { char *p; ~Free() { if (p) (void)hipFree(p); } } guard{ws.ptr ? nullptr : blk};
    char *q = blk;
    BatchArgs<R> B{};
    B.args = (const TDSolveArgs<R> *)q; q += s_args;
    B.team_of_block = (const int *)q; q += s_tob;
    B.teams = (const int2 *)q; q += s_teams;
    B.rel = (unsigned long long *)q; q += s_rel;
    B.state = (int *)q; q += s_state;
    B.sweeps = (int *)q;
    B.gs.arrive = e.arrive;
    B.gs.rel = e.rel;
    B.gs.err = e.err;
    B.n = n;
    st = hipMemcpyAsync((void *)B.args, args.data(), sizeof(TDSolveArgs<R>) * n, hipMemcpyHostToDevice, e.stream);
    if (st == hipSuccess)
        st = hipMemcpyAsync((void *)B.team_of_block, team_of_block.data(), sizeof(int) * grid, hipMemcpyHostToDevice, e.stream);
    if (st == hipSuccess) st = hipMemcpyAsync((void *)B.teams, teams.data(), sizeof(int2) * n, hipMemcpyHostToDevice, e.stream);
    if (st == hipSuccess) st = hipMemcpyAsync(B.state, state.data(), sizeof(int) * 8 * n, hipMemcpyHostToDevice, e.stream);
    if (st != hipSuccess) return check_hip(st, "glx_batch_td_run upload");
    hipLaunchKernelGGL(k_batch_reset, dim3(std::max(1, grid_blocks(std::max(grid, n)))), dim3(256), 0, e.stream, e.arrive,
                       grid, B.rel, n, e.err, e.abort_next ? 1 : 0);
    GLX_LAUNCH_CHECK("k_batch_reset");
    if (G)
        for (int i = 0; i < n; ++i) {
            if (!inst[i].p.solveA) continue;
            const glx_grid &g = inst[i].g;
            const long Nab = (long)(g.Nx - 1) * g.Ny + (long)g.Nx * (g.Ny - 1);
            hipLaunchKernelGGL(k_gauge_shift<R>, dim3(grid_blocks(Nab)), dim3(256), 0, e.stream,
                               (R *)inst[i].ab[inst[i].ab_cur], g.Nx, g.Ny, inst[i].p.H, g.dx, g.dy,
                               0.5 * g.dx * (g.Nx - 1), 0.5 * g.dy * (g.Ny - 1), -1.0);
            GLX_LAUNCH_CHECK("k_gauge_shift");
        }
    st = hipEventRecord(e.kev0, e.stream);
    if (st != hipSuccess) return check_hip(st, "hipEventRecord");
    int launched = 0;
    if (use_graph && nsteps > 1 && !ws.ptr) {
        const int K = std::min(nsteps, 64);
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
        st = hipStreamBeginCapture(e.stream, hipStreamCaptureModeThreadLocal);
        if (st != hipSuccess) return check_hip(st, "hipStreamBeginCapture");
        for (int k = 0; k < K; ++k) hipLaunchKernelGGL((k_batch_td_step<R, G, KB>), dim3(grid), dim3(256), 0, e.stream, B);
        st = hipStreamEndCapture(e.stream, &graph);
        if (st != hipSuccess) return check_hip(st, "hipStreamEndCapture");
        st = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        if (st != hipSuccess) {
            (void)hipGraphDestroy(graph);
            return check_hip(st, "hipGraphInstantiate");
        }
        for (; launched + K <= nsteps && st == hipSuccess; launched += K) st = hipGraphLaunch(exec, e.stream);
        (void)hipGraphExecDestroy(exec);
        (void)hipGraphDestroy(graph);
        if (st != hipSuccess) return check_hip(st, "hipGraphLaunch");
    }
    for (; launched < nsteps; ++launched) {
        hipLaunchKernelGGL((k_batch_td_step<R, G, KB>), dim3(grid), dim3(256), 0, e.stream, B);
        GLX_LAUNCH_CHECK("k_batch_td_step");
    }
    st = hipEventRecord(e.kev1, e.stream);
    if (st != hipSuccess) return check_hip(st, "hipEventRecord");
    if (G)
        for (int i = 0; i < n; ++i) {
            if (!inst[i].p.solveA) continue;
            const glx_grid &g = inst[i].g;
            const long Nab = (long)(g.Nx - 1) * g.Ny + (long)g.Nx * (g.Ny - 1);
            hipLaunchKernelGGL(k_gauge_shift_state<R>, dim3(grid_blocks(Nab)), dim3(256), 0, e.stream, (R *)inst[i].ab[0],
                               (R *)inst[i].ab[1], (R *)inst[i].ab[2], B.state + 8 * i, g.Nx, g.Ny, inst[i].p.H, g.dx,
                               g.dy, 0.5 * g.dx * (g.Nx - 1), 0.5 * g.dy * (g.Ny - 1), 1.0);
            GLX_LAUNCH_CHECK("k_gauge_shift_state");
        }
    int host_err = 0;
    std::vector<int> sw_host(sweeps ? 2 * (size_t)n * nsteps : 0);
    st = hipMemcpyAsync(state.data(), B.state, sizeof(int) * 8 * n, hipMemcpyDeviceToHost, e.stream);
    if (st == hipSuccess) st = hipMemcpyAsync(&host_err, e.err, sizeof(int), hipMemcpyDeviceToHost, e.stream);
    if (st == hipSuccess && sweeps)
        st = hipMemcpyAsync(sw_host.data(), B.sweeps, sizeof(int) * sw_host.size(), hipMemcpyDeviceToHost, e.stream);
    if (st == hipSuccess) st = hipStreamSynchronize(e.stream);
    if (st != hipSuccess) return check_hip(st, "glx_batch_td_run readback");
    float kms = 0.f;
    if (hipEventElapsedTime(&kms, e.kev0, e.kev1) == hipSuccess && e.last_kernel_ms) *e.last_kernel_ms = kms;
    if (host_err) return set_error(GLX_E_TIMEOUT, "%s: a team barrier timed out", ws.what);
    for (int i = 0; i < n; ++i) {
        inst[i].psi_cur = state[8 * i + ST_PSI];
        inst[i].ab_cur = state[8 * i + ST_AB];
    }
    if (sweeps) std::copy(sw_host.begin(), sw_host.end(), sweeps);
    return GLX_OK;
}

template <typename R>
int L<R>::batch_td_run(const Env &e, glx_batch_instance *inst, int32_t n, int32_t nsteps, int32_t *sweeps,
                       int32_t use_graph, int32_t wg_per_cu, int32_t min_rows, const double *weight,
                       int32_t psi_depth, const Workspace &ws) {
    if (!inst || n < 1 || nsteps < 0) return set_error(GLX_E_ARG, "%s: invalid arguments", ws.what);
    if (psi_depth == 0) psi_depth = 3;
    if (psi_depth < 2 || psi_depth > 4) return set_error(GLX_E_ARG, "%s: psi_depth %d not in 2..4", ws.what, psi_depth);
    if ((long)n * nsteps > (1L << 26)) return set_error(GLX_E_ARG, "%s: n*nsteps too large", ws.what);
    bool any_A = false;
    for (int i = 0; i < n; ++i) {
        const glx_batch_instance &I = inst[i];
        const glx_td_params &p = I.p;
        if (!grid_ok(I.g)) return set_error(GLX_E_ARG, "%s: instance %d: invalid grid", ws.what, i);
        if (p.max_sweeps < 1 || p.stop_psi <= 0.0 || p.stop_A <= 0.0 || p.eqn != 0)
            return set_error(GLX_E_ARG, "%s: instance %d: invalid parameters", ws.what, i);
        if (p.langevin_psi > 1.0e-32 || p.langevin_A > 1.0e-32)
            return set_error(GLX_E_ARG, "%s: instance %d: Langevin noise is not supported in batches", ws.what, i);
        if (p.a_src)
            return set_error(GLX_E_ARG, "%s: instance %d: a fixed A source is not supported in batches", ws.what, i);
        if (I.psi_cur < 0 || I.psi_cur > 2 || I.ab_cur < 0 || I.ab_cur > 2)
            return set_error(GLX_E_ARG, "%s: instance %d: buffer index", ws.what, i);
        for (int k = 0; k < 3; ++k)
            if (!I.psi[k] || !I.ab[k]) return set_error(GLX_E_ARG, "%s: instance %d: NULL buffer", ws.what, i);
        any_A = any_A || p.solveA;
    }
    if (nsteps == 0) return GLX_OK;
    auto run = [&](auto g, auto kb) {
        return batch_run_form<R, decltype(g)::value, decltype(kb)::value>(e, inst, n, nsteps, sweeps, use_graph,
                                                                         wg_per_cu, min_rows, weight, ws);
    };
    auto by_depth = [&](auto g) {
        if (psi_depth == 2) return run(g, std::integral_constant<int, 2>{});
        if (psi_depth == 4) return run(g, std::integral_constant<int, 4>{});
        return run(g, std::integral_constant<int, 3>{});
    };
    if constexpr (RunForm<R>::has_gauge)
        if (any_A) return by_depth(std::true_type{});
    return by_depth(std::false_type{});
}

// ---- batched CG -----------------------------------------------------------------

template <typename R, int LL>
static int batch_cg_form(const Env &e, glx_batch_cg_instance *inst, int32_t n, int32_t max_iters, double rtol,
                         int32_t per_launch, int32_t stall_iters, const double *init_state) {
    const bool withA = LL == 17;
    // scratch per instance: jac[2], dir[2] (N complex each) and, with A, jacA[2], dirA[2] (Nab reals each)
    std::vector<size_t> off(n + 1, 0);
    auto al = [](size_t b) { return (b + 255) & ~size_t(255); };
    for (int i = 0; i < n; ++i) {
        const glx_grid &g = inst[i].g;
        const size_t N = (size_t)g.Nx * g.Ny, Nab = (size_t)(g.Nx - 1) * g.Ny + (size_t)g.Nx * (g.Ny - 1);
        off[i + 1] = off[i] + 4 * al(N * sizeof(cplx<R>)) + (withA ? 4 * al(Nab * sizeof(R)) : 0);
    }
    const size_t s_inst = al(sizeof(CGInst<R>) * n), s_state = al(sizeof(double) * CG_NSTATE * n);
    char *blk = nullptr;
    hipError_t st = hipMalloc((void **)&blk, s_inst + s_state + off[n]);
    if (st != hipSuccess) return check_hip(st, "hipMalloc(batch cg)");
    struct Free { char *p; ~Free() { if (p) (void)hipFree(p); } } guard{blk};
    CGInst<R> *d_inst = (CGInst<R> *)blk;
    double *d_state = (double *)(blk + s_inst);
    char *work = blk + s_inst + s_state;
    std::vector<CGInst<R>> h(n);
    for (int i = 0; i < n; ++i) {
        const glx_batch_cg_instance &I = inst[i];
        const glx_grid &g = I.g;
        const size_t N = (size_t)g.Nx * g.Ny, Nab = (size_t)(g.Nx - 1) * g.Ny + (size_t)g.Nx * (g.Ny - 1);
        CGInst<R> &c = h[i];
        c.g = make_geo(g.Nx, g.Ny, g.dx, g.dy);
        c.kappa2 = R(I.kappa2);
        c.eps = R(I.eps);
        c.H = R(I.H);
        c.eps_map = (const R *)I.eps_map;
        c.mt = I.mt;
        c.psi = (cplx<R> *)I.psi;
        c.ab = withA ? (R *)I.ab : nullptr;
        char *w = work + off[i];
        for (int k = 0; k < 2; ++k) { c.jac[k] = (cplx<R> *)w; w += al(N * sizeof(cplx<R>)); }
        for (int k = 0; k < 2; ++k) { c.dir[k] = (cplx<R> *)w; w += al(N * sizeof(cplx<R>)); }
        for (int k = 0; k < 2; ++k) { c.jacA[k] = withA ? (R *)w : nullptr; if (withA) w += al(Nab * sizeof(R)); }
        for (int k = 0; k < 2; ++k) { c.dirA[k] = withA ? (R *)w : nullptr; if (withA) w += al(Nab * sizeof(R)); }
    }
    st = hipMemcpyAsync(d_inst, h.data(), sizeof(CGInst<R>) * n, hipMemcpyHostToDevice, e.stream);
    if (st == hipSuccess) st = hipMemsetAsync(d_state, 0, s_state + off[n], e.stream);
    if (st == hipSuccess && init_state)
        st = hipMemcpyAsync(d_state, init_state, sizeof(double) * CG_NSTATE * n, hipMemcpyHostToDevice, e.stream);
    if (st != hipSuccess) return check_hip(st, "glx_batch_cg_run upload");
    int per_cu = 0;
    st = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_batch_cg<R, LL>, 256, 0);
    if (st != hipSuccess) return check_hip(st, "hipOccupancyMaxActiveBlocksPerMultiprocessor");
    const int grid = std::max(1, std::min(n, per_cu * e.num_cus));
    const int chunk = per_launch > 0 ? per_launch : 50;
    int it_min = max_iters;  // fewest iterations any instance starts with
    for (int i = 0; i < n; ++i) it_min = std::min(it_min, init_state ? (int)init_state[CG_NSTATE * i + CG_IT] : 0);
    std::vector<double> hs(CG_NSTATE * (size_t)n);
    if (init_state) std::copy(init_state, init_state + CG_NSTATE * (size_t)n, hs.begin());
    else std::fill(hs.begin(), hs.end(), 0.0);
    st = hipEventRecord(e.kev0, e.stream);
    if (st != hipSuccess) return check_hip(st, "hipEventRecord");
    for (int done = it_min; done < max_iters; done += chunk) {
        hipLaunchKernelGGL((k_batch_cg<R, LL>), dim3(grid), dim3(256), 0, e.stream, d_inst, d_state, n,
                           std::min(chunk, max_iters - done), max_iters, rtol, stall_iters);
        GLX_LAUNCH_CHECK("k_batch_cg");
        st = hipMemcpyAsync(hs.data(), d_state, sizeof(double) * CG_NSTATE * n, hipMemcpyDeviceToHost, e.stream);
        if (st == hipSuccess) st = hipStreamSynchronize(e.stream);
        if (st != hipSuccess) return check_hip(st, "glx_batch_cg_run readback");
        bool all = true;
        for (int i = 0; i < n && all; ++i)
            all = hs[CG_NSTATE * i + CG_DONE] != 0.0 || hs[CG_NSTATE * i + CG_IT] >= max_iters;
        if (all) break;
    }
    st = hipEventRecord(e.kev1, e.stream);
    if (st == hipSuccess) st = hipStreamSynchronize(e.stream);
    if (st != hipSuccess) return check_hip(st, "hipEventRecord");
    float kms = 0.f;
    if (hipEventElapsedTime(&kms, e.kev0, e.kev1) == hipSuccess && e.last_kernel_ms) *e.last_kernel_ms = kms;
    for (int i = 0; i < n; ++i) {
        const double *sv = hs.data() + CG_NSTATE * i;
        inst[i].iters = (int32_t)sv[CG_IT];
        inst[i].converged = sv[CG_DONE] == 1.0 ? 1 : 0;
        inst[i].stalled = sv[CG_DONE] == 2.0 ? 1 : 0;
        inst[i].grad_rel = sv[CG_GREL];
        inst[i].grad0 = sv[CG_G0];
        inst[i].fp64_iters = 0;
    }
    return GLX_OK;
}

template <typename D, typename S> __global__ void __launch_bounds__(256) k_convert(D *dst, const S *src, long n) {
    for (long k = blockIdx.x * (long)blockDim.x + threadIdx.x; k < n; k += (long)gridDim.x * blockDim.x)
        dst[k] = D(src[k]);
}

template <typename D, typename S> static int convert(const Env &e, D *dst, const S *src, long n) {
    if (n <= 0) return GLX_OK;
    hipLaunchKernelGGL((k_convert<D, S>), dim3(grid_blocks(n)), dim3(256), 0, e.stream, dst, src, n);
    GLX_LAUNCH_CHECK("k_convert");
    return GLX_OK;
}

// FP64 finish of the instances an FP32-mixed run left stalled: each is
// copied into FP64 (psi, A and the epsilon map), continues from its
// iteration count with a steepest-descent step and the FP32 run's
// ||grad F||_0, to rtol or max_iters in all, and its FP64 result is
// rounded back into the FP32 state.
template <int LL>
static int batch_cg_finish64(const Env &e, glx_batch_cg_instance *inst, int32_t n, int32_t max_iters, double rtol,
                             int32_t per_launch) {
    const bool withA = LL == 17;
    std::vector<int> idx;
    for (int i = 0; i < n; ++i)
        if (inst[i].stalled && inst[i].iters < max_iters) idx.push_back(i);
    if (idx.empty()) return GLX_OK;
    const int m = (int)idx.size();
    auto al = [](size_t b) { return (b + 255) & ~size_t(255); };
    auto sizes = [](const glx_grid &g, size_t &N, size_t &Nab) {
        N = (size_t)g.Nx * g.Ny;
        Nab = (size_t)(g.Nx - 1) * g.Ny + (size_t)g.Nx * (g.Ny - 1);
    };
    std::vector<size_t> off(m + 1, 0);
    for (int k = 0; k < m; ++k) {
        size_t N, Nab;
        sizes(inst[idx[k]].g, N, Nab);
        off[k + 1] = off[k] + al(N * 16) + (withA ? al(Nab * 8) : 0) + (inst[idx[k]].eps_map ? al(N * 8) : 0);
    }
    char *blk = nullptr;
    hipError_t st = hipMalloc((void **)&blk, off[m]);
    if (st != hipSuccess) return check_hip(st, "hipMalloc(fp64 finish)");
    struct Free { char *p; ~Free() { if (p) (void)hipFree(p); } } guard{blk};
    std::vector<glx_batch_cg_instance> sub(m);
    std::vector<double> init(CG_NSTATE * (size_t)m, 0.0);
    for (int k = 0; k < m; ++k) {
        const glx_batch_cg_instance &I = inst[idx[k]];
        size_t N, Nab;
        sizes(I.g, N, Nab);
        char *w = blk + off[k];
        glx_batch_cg_instance &J = sub[k];
        J = I;
        J.psi = w;
        w += al(N * 16);
        int rc = convert(e, (double *)J.psi, (const float *)I.psi, 2 * (long)N);
        if (!rc && withA) {
            J.ab = w;
            w += al(Nab * 8);
            rc = convert(e, (double *)J.ab, (const float *)I.ab, (long)Nab);
        }
        if (!rc && I.eps_map) {
            J.eps_map = w;
            rc = convert(e, (double *)w, (const float *)I.eps_map, (long)N);
        }
        if (rc) return rc;
        double *sv = init.data() + CG_NSTATE * k;
        sv[CG_IT] = sv[CG_IT0] = I.iters;
        sv[CG_G0] = I.grad0;
    }
    const float kms32 = e.last_kernel_ms ? *e.last_kernel_ms : 0.f;
    int rc = batch_cg_form<double, LL>(e, sub.data(), m, max_iters, rtol, per_launch, 0, init.data());
    if (rc) return rc;
    if (e.last_kernel_ms) *e.last_kernel_ms += kms32;
    for (int k = 0; k < m; ++k) {
        glx_batch_cg_instance &I = inst[idx[k]];
        const glx_batch_cg_instance &J = sub[k];
        size_t N, Nab;
        sizes(I.g, N, Nab);
        rc = convert(e, (float *)I.psi, (const double *)J.psi, 2 * (long)N);
        if (!rc && withA) rc = convert(e, (float *)I.ab, (const double *)J.ab, (long)Nab);
        if (rc) return rc;
        I.fp64_iters = J.iters - I.iters;
        I.iters = J.iters;
        I.converged = J.converged;
        I.stalled = 0;
        I.grad_rel = J.grad_rel;
    }
    st = hipStreamSynchronize(e.stream);  // before the FP64 block is freed
    return st == hipSuccess ? GLX_OK : check_hip(st, "glx_batch_cg_run fp64 finish");
}

template <typename R>
int L<R>::batch_cg_run(const Env &e, glx_batch_cg_instance *inst, int32_t n, int32_t max_iters, double rtol,
                       int32_t iters_per_launch, int32_t stall_iters, int32_t fp64_finish) {
    if (!inst || n < 1 || max_iters < 0 || !(rtol >= 0.0) || stall_iters < 0)
        return set_error(GLX_E_ARG, "glx_batch_cg_run: invalid arguments");
    if (fp64_finish && (sizeof(R) != sizeof(float) || stall_iters < 1))
        return set_error(GLX_E_ARG, "glx_batch_cg_run: the FP64 finish needs an FP32-mixed batch and a stall rule");
    const bool withA = inst[0].kappa2 > 0.0;
    for (int i = 0; i < n; ++i) {
        const glx_batch_cg_instance &I = inst[i];
        if (!grid_ok(I.g) || !I.psi) return set_error(GLX_E_ARG, "glx_batch_cg_run: instance %d: invalid grid or psi", i);
        if ((I.kappa2 > 0.0) != withA)
            return set_error(GLX_E_ARG, "glx_batch_cg_run: instances must all have kappa < inf or all kappa = inf");
        if (withA && !I.ab) return set_error(GLX_E_ARG, "glx_batch_cg_run: instance %d: NULL vector potential", i);
    }
    for (int i = 0; i < n; ++i) {
        inst[i].iters = inst[i].converged = inst[i].stalled = inst[i].fp64_iters = 0;
        inst[i].grad_rel = inst[i].grad0 = 0.0;
    }
    if (max_iters == 0) return GLX_OK;
    int rc = withA ? batch_cg_form<R, 17>(e, inst, n, max_iters, rtol, iters_per_launch, stall_iters, nullptr)
                   : batch_cg_form<R, 5>(e, inst, n, max_iters, rtol, iters_per_launch, stall_iters, nullptr);
    if (rc || !fp64_finish) return rc;
    return withA ? batch_cg_finish64<17>(e, inst, n, max_iters, rtol, iters_per_launch)
                 : batch_cg_finish64<5>(e, inst, n, max_iters, rtol, iters_per_launch);
}

template <typename R> int L<R>::phase_lock(const Env &e, void *psi, int32_t n, const int32_t *ns) {
    if (n <= 0) return GLX_OK;
    if (!psi || !ns) return set_error(GLX_E_ARG, "glx_td_phase_lock: NULL pointer");
    hipLaunchKernelGGL(k_phase_lock<R>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, e.stream, (cplx<R> *)psi, n, ns);
    GLX_LAUNCH_CHECK("k_phase_lock");
    return GLX_OK;
}

template <typename R>
int L<R>::magnetic_field(const Env &e, glx_grid g, const void *abei, const void *ab, void *B) {
    if (!grid_ok(g) || !B) return set_error(GLX_E_ARG, "glx_magnetic_field: invalid arguments");
    const Geo G = make_geo(g.Nx, g.Ny, g.dx, g.dy);
    hipLaunchKernelGGL(k_magnetic_field<R>, dim3(grid_blocks(G.Nc)), dim3(256), 0, e.stream, G, (const R *)abei, (const R *)ab, (R *)B);
    GLX_LAUNCH_CHECK("k_magnetic_field");
    return GLX_OK;
}

template <typename R>
int L<R>::current_density(const Env &e, glx_grid g, double kappa2, double H, const void *abei, const void *ab, void *J) {
    if (!grid_ok(g) || !J) return set_error(GLX_E_ARG, "glx_current_density: invalid arguments");
    const Geo G = make_geo(g.Nx, g.Ny, g.dx, g.dy);
    hipLaunchKernelGGL(k_current_density<R>, dim3(grid_blocks(G.N)), dim3(256), 0, e.stream, G, R(kappa2), R(H),
                       (const R *)abei, (const R *)ab, (R *)J);
    GLX_LAUNCH_CHECK("k_current_density");
    return GLX_OK;
}

template <typename R>
int L<R>::supercurrent_density(const Env &e, glx_grid g, const uint8_t *mt, const void *psi, const void *abei,
                               const void *ab, void *Js) {
    if (!grid_ok(g) || !psi || !Js) return set_error(GLX_E_ARG, "glx_supercurrent_density: invalid arguments");
    const Geo G = make_geo(g.Nx, g.Ny, g.dx, g.dy);
    hipLaunchKernelGGL(k_supercurrent_density<R>, dim3(grid_blocks(G.N)), dim3(256), 0, e.stream, G, mt,
                       (const cplx<R> *)psi, (const R *)abei, (const R *)ab, (R *)Js);
    GLX_LAUNCH_CHECK("k_supercurrent_density");
    return GLX_OK;
}

template <typename R>
int L<R>::bulk_observables(const Env &e, glx_grid g, double kappa2, double H, const uint8_t *mt, const void *psi,
                           const void *abei, const void *ab, void *B, void *J, void *Js) {
    if (!grid_ok(g) || (Js && !psi)) return set_error(GLX_E_ARG, "glx_bulk_observables: invalid arguments");
    if (!B && !J && !Js) return GLX_OK;
    const Geo G = make_geo(g.Nx, g.Ny, g.dx, g.dy);
    hipLaunchKernelGGL(k_bulk_observables<R>, dim3(grid_blocks(G.N)), dim3(256), 0, e.stream, G, R(kappa2), R(H), mt,
                       (const cplx<R> *)psi, (const R *)abei, (const R *)ab, (R *)B, (R *)J, (R *)Js);
    GLX_LAUNCH_CHECK("k_bulk_observables");
    return GLX_OK;
}

template <typename R>
int L<R>::free_energy(const Env &e, glx_grid g, double kappa2, double eps, const void *eps_map, double H,
                      const uint8_t *mt, const void *psi, const void *abei, const void *ab, double *F) {
    if (!grid_ok(g) || !psi || !F) return set_error(GLX_E_ARG, "glx_free_energy: invalid arguments");
    if (kappa2 > 0.0 && !abei && !ab) return set_error(GLX_E_ARG, "glx_free_energy: finite kappa needs a vector potential");
    const Geo G = make_geo(g.Nx, g.Ny, g.dx, g.dy);
    const int nb = grid_blocks(G.N, RED_BLOCK, RED_MAX_BLOCKS);
    hipLaunchKernelGGL(k_free_energy<R>, dim3(nb), dim3(RED_BLOCK), 0, e.stream, G, R(kappa2), R(eps), (const R *)eps_map,
                       R(H), mt, (const cplx<R> *)psi, (const R *)abei, (const R *)ab, e.partials);
    GLX_LAUNCH_CHECK("k_free_energy");
    const double dxdy = double(R(g.dx)) * double(R(g.dy));
    hipLaunchKernelGGL(k_finish_partials, dim3(1), dim3(RED_BLOCK), 0, e.stream, e.partials, nb, 1, dxdy, e.result);
    GLX_LAUNCH_CHECK("k_finish_partials");
    return fetch_result(e, F, 1);
}

// The library's growable device buffer, at least `bytes` large. Growing
// frees the old buffer after the device has finished with it (hipFree
// waits for the device).
static int grow_workspace(const Env &e, size_t bytes, void **out) {
    if (*e.grow_bytes < bytes) {
        if (*e.grow) {
            hipError_t st = hipFree(*e.grow);
            *e.grow = nullptr;
            *e.grow_bytes = 0;
            if (st != hipSuccess) return check_hip(st, "hipFree(workspace)");
        }
        hipError_t st = hipMalloc(e.grow, bytes);
        if (st != hipSuccess) {
            *e.grow = nullptr;
            return check_hip(st, "hipMalloc(workspace)");
        }
        *e.grow_bytes = bytes;
    }
    *out = *e.grow;
    return GLX_OK;
}

template <typename R>
int L<R>::vortex_cells(const Env &e, glx_grid g, const void *psi, const void *ab, const void *abi, double H,
                       double *rec, int32_t capacity, int32_t *count) {
    if (!grid_ok(g) || !psi || !ab || !count || capacity < 0 || (capacity > 0 && !rec))
        return set_error(GLX_E_ARG, "glx_vortex_cells: invalid arguments");
    const Geo G = make_geo(g.Nx, g.Ny, g.dx, g.dy);
    const long T = vortex_rows(G), W = vortex_waves(G), ns = T / G.Nyc;
    // workspace: the masks (8 B per strip row), the ranks (4 B per strip
    // row) and the per-row counts and offsets (4 B per cell row)
    void *ws = nullptr;
    int rc = grow_workspace(e, (sizeof(unsigned long long) + sizeof(int)) * (size_t)T + sizeof(int) * (size_t)G.Nyc, &ws);
    if (rc) return rc;
    unsigned long long *bits = (unsigned long long *)ws;
    int *pre = (int *)(bits + T), *rowcnt = pre + T;
    int *dcount = (int *)e.result;
    const unsigned nb = (unsigned)((W + 3) / 4), nr = (unsigned)((G.Nyc + WAVE - 1) / WAVE);
    hipLaunchKernelGGL(k_vortex_flags<R>, dim3(nb), dim3(256), 0, e.stream, G, (const cplx<R> *)psi, (const R *)ab,
                       double(R(H)), bits, W);
    GLX_LAUNCH_CHECK("k_vortex_flags");
    hipLaunchKernelGGL(k_vortex_rowprefix, dim3(nr), dim3(256), 0, e.stream, (const unsigned long long *)bits, ns,
                       G.Nyc, pre, rowcnt);
    GLX_LAUNCH_CHECK("k_vortex_rowprefix");
    hipLaunchKernelGGL(k_vortex_rowscan, dim3(1), dim3(1024), 0, e.stream, rowcnt, G.Nyc, dcount);
    GLX_LAUNCH_CHECK("k_vortex_rowscan");
    hipLaunchKernelGGL(k_vortex_emit<R>, dim3(nb), dim3(256), 0, e.stream, G, (const cplx<R> *)psi, (const R *)ab,
                       (const R *)abi, double(R(H)), (const unsigned long long *)bits, (const int *)pre,
                       (const int *)rowcnt, rec, (int)capacity, W);
    GLX_LAUNCH_CHECK("k_vortex_emit");
    int c = 0;
    hipError_t st = hipMemcpyAsync(&c, dcount, sizeof(int), hipMemcpyDeviceToHost, e.stream);
    if (st == hipSuccess) st = hipStreamSynchronize(e.stream);
    if (st != hipSuccess) return check_hip(st, "glx_vortex_cells readback");
    *count = c;
    if (c > capacity) return set_error(GLX_E_CAPACITY, "glx_vortex_cells: %d cells flagged, capacity %d", c, capacity);
    return GLX_OK;
}

template <typename R>
int L<R>::vortex_triangulate(const Env &e, glx_grid g, const double *rec, int32_t count, double *xyv) {
    if (!grid_ok(g) || count < 0 || (count > 0 && (!rec || !xyv)))
        return set_error(GLX_E_ARG, "glx_vortex_triangulate: invalid arguments");
    if (count == 0) return GLX_OK;
    const Geo G = make_geo(g.Nx, g.Ny, g.dx, g.dy);
    hipLaunchKernelGGL(k_vortex_triangulate<R>, dim3((count + 255) / 256), dim3(256), 0, e.stream, G, rec, (int)count,
                       xyv);
    GLX_LAUNCH_CHECK("k_vortex_triangulate");
    return GLX_OK;
}

template <typename R> int L<R>::sum(const Env &e, const void *x, int64_t n, int32_t bs, double *out) {
    if (!x || !out || n < 0) return set_error(GLX_E_ARG, "glx_sum: invalid arguments");
    if (bs == 0) bs = RED_BLOCK;
    if (bs != RED_BLOCK && (bs % 64 != 0 || bs > 1024 || bs < 64))
        return set_error(GLX_E_ARG, "glx_sum: block_size must be a multiple of 64 in [64, 1024]");
    const int nb = grid_blocks(n, RED_BLOCK, RED_MAX_BLOCKS);
    hipLaunchKernelGGL(k_sum<R>, dim3(nb), dim3(RED_BLOCK), 0, e.stream, (const R *)x, (long)n, e.partials);
    GLX_LAUNCH_CHECK("k_sum");
    hipLaunchKernelGGL(k_finish_partials, dim3(1), dim3(RED_BLOCK), 0, e.stream, e.partials, nb, 1, 1.0, e.result);
    GLX_LAUNCH_CHECK("k_finish_partials");
    return fetch_result(e, out, 1);
}

template <typename R> int L<R>::sum_v(const Env &e, const void *x, int64_t nv, int32_t ne, int32_t bs, double *out) {
    if (!x || !out || nv < 0 || ne < 1 || ne > 32) return set_error(GLX_E_ARG, "glx_sum_v: invalid arguments");
    if (bs != 0 && (bs % 64 != 0 || bs > 1024 || bs < 64))
        return set_error(GLX_E_ARG, "glx_sum_v: block_size must be a multiple of 64 in [64, 1024]");
    const int nb = grid_blocks(nv, RED_BLOCK, RED_MAX_BLOCKS);
    hipLaunchKernelGGL(k_sum_v<R>, dim3(nb), dim3(RED_BLOCK), 0, e.stream, (const R *)x, (long)nv, (int)ne, e.partials);
    GLX_LAUNCH_CHECK("k_sum_v");
    hipLaunchKernelGGL(k_finish_partials, dim3(1), dim3(RED_BLOCK), 0, e.stream, e.partials, nb, (int)ne, 1.0, e.result);
    GLX_LAUNCH_CHECK("k_finish_partials");
    return fetch_result(e, out, ne);
}

template <typename R> int L<R>::pr_sums(const Env &e, const void *x, const void *y, int64_t n, double *out) {
    if (!x || !y || !out || n < 0) return set_error(GLX_E_ARG, "glx_pr_sums: invalid arguments");
    const int nb = grid_blocks(n, RED_BLOCK, RED_MAX_BLOCKS);
    hipLaunchKernelGGL(k_pr_sums<R>, dim3(nb), dim3(RED_BLOCK), 0, e.stream, (const R *)x, (const R *)y, (long)n, e.partials);
    GLX_LAUNCH_CHECK("k_pr_sums");
    hipLaunchKernelGGL(k_finish_partials, dim3(1), dim3(RED_BLOCK), 0, e.stream, e.partials, nb, 2, 1.0, e.result);
    GLX_LAUNCH_CHECK("k_finish_partials");
    return fetch_result(e, out, 2);
}

template <typename R>
int L<R>::jacobian_psi(const Env &e, glx_grid g, double kappa2, double eps, const void *eps_map, double H,
                       const uint8_t *mt, const void *psi, const void *abei, const void *ab, void *Gp) {
    if (!grid_ok(g) || !psi || !Gp) return set_error(GLX_E_ARG, "glx_cg_jacobian_psi: invalid arguments");
    const Geo G = make_geo(g.Nx, g.Ny, g.dx, g.dy);
    hipLaunchKernelGGL(k_jacobian_psi<R>, dim3(grid_blocks(G.N)), dim3(256), 0, e.stream, G, R(kappa2), R(eps),
                       (const R *)eps_map, R(H), mt, (const cplx<R> *)psi, (const R *)abei, (const R *)ab, (cplx<R> *)Gp);
    GLX_LAUNCH_CHECK("k_jacobian_psi");
    return GLX_OK;
}

template <typename R>
int L<R>::jacobian_A(const Env &e, glx_grid g, double kappa2, double H, const uint8_t *mt, const void *psi,
                     const void *abei, const void *ab, void *Gp) {
    if (!grid_ok(g) || !psi || !Gp) return set_error(GLX_E_ARG, "glx_cg_jacobian_A: invalid arguments");
    const Geo G = make_geo(g.Nx, g.Ny, g.dx, g.dy);
    hipLaunchKernelGGL(k_jacobian_A<R>, dim3(grid_blocks(G.N)), dim3(256), 0, e.stream, G, R(kappa2), R(H), mt,
                       (const cplx<R> *)psi, (const R *)abei, (const R *)ab, (R *)Gp);
    GLX_LAUNCH_CHECK("k_jacobian_A");
    return GLX_OK;
}

template <typename R>
int L<R>::coef_psi(const Env &e, glx_grid g, double kappa2, double eps, double H, const uint8_t *mt, const void *psi,
                   const void *dpsi, const void *abei, const void *ab, double *c5) {
    if (!grid_ok(g) || !psi || !dpsi || !c5) return set_error(GLX_E_ARG, "glx_cg_coef_psi: invalid arguments");
    if (kappa2 > 0.0 && !ab) return set_error(GLX_E_ARG, "glx_cg_coef_psi: finite kappa needs ab");
    const Geo G = make_geo(g.Nx, g.Ny, g.dx, g.dy);
    const int nb = grid_blocks(G.N, RED_BLOCK, RED_MAX_BLOCKS);
    hipLaunchKernelGGL((k_coef<R, 5>), dim3(nb), dim3(RED_BLOCK), 0, e.stream, G, R(kappa2), R(eps), R(H), mt,
                       (const cplx<R> *)psi, (const cplx<R> *)dpsi, (const R *)abei, (const R *)ab, (const R *)nullptr, e.partials);
    GLX_LAUNCH_CHECK("k_coef5");
    const double dxdy = double(R(g.dx)) * double(R(g.dy));
    hipLaunchKernelGGL(k_finish_partials, dim3(1), dim3(RED_BLOCK), 0, e.stream, e.partials, nb, 5, dxdy, e.result);
    GLX_LAUNCH_CHECK("k_finish_partials");
    return fetch_result(e, c5, 5);
}

template <typename R>
int L<R>::coef(const Env &e, glx_grid g, double kappa2, double eps, double H, const uint8_t *mt, const void *psi,
               const void *dpsi, const void *abei, const void *ab, const void *dab, double *c17) {
    if (!grid_ok(g) || !psi || !dpsi || !dab || !c17) return set_error(GLX_E_ARG, "glx_cg_coef: invalid arguments");
    const Geo G = make_geo(g.Nx, g.Ny, g.dx, g.dy);
    const int nb = grid_blocks(G.N, RED_BLOCK, RED_MAX_BLOCKS);
    hipLaunchKernelGGL((k_coef<R, 17>), dim3(nb), dim3(RED_BLOCK), 0, e.stream, G, R(kappa2), R(eps), R(H), mt,
                       (const cplx<R> *)psi, (const cplx<R> *)dpsi, (const R *)abei, (const R *)ab, (const R *)dab, e.partials);
    GLX_LAUNCH_CHECK("k_coef17");
    const double dxdy = double(R(g.dx)) * double(R(g.dy));
    hipLaunchKernelGGL(k_finish_partials, dim3(1), dim3(RED_BLOCK), 0, e.stream, e.partials, nb, 17, dxdy, e.result);
    GLX_LAUNCH_CHECK("k_finish_partials");
    return fetch_result(e, c17, 17);
}

// The gradient pass, optionally with the state update fused in (S, alpha:
// UPD; S.dpsi NULL: no update).
template <typename R>
static int gradient_pass_launch(const Env &e, glx_grid g, double kappa2, double eps, const void *eps_map, double H,
                                const uint8_t *mt, const void *psi, const void *abei, const void *ab, void *jac,
                                const void *jac_prev, void *jacA, const void *jacA_prev, const StateUpdate<R> &S,
                                const double *alpha, int *nb_out) {
    const bool withA = kappa2 > 0.0;  // the psi Jacobian reads A in both cases
    const bool upd = S.dpsi != nullptr;
    const Geo G = make_geo(g.Nx, g.Ny, g.dx, g.dy);
    const void *fn = withA ? (upd ? (const void *)k_cg_grad_rows<R, true, true> : (const void *)k_cg_grad_rows<R, true, false>)
                           : (upd ? (const void *)k_cg_grad_rows<R, false, true> : (const void *)k_cg_grad_rows<R, false, false>);
    // two workgroups per CU, not the three the gradient pass's registers
    // allow: measured on a 4097^2 grid with holes and an eps map, a CG
    // iteration takes 0.70 ms this way against 0.72 ms with three (its
    // traffic per node barely changes, so the gain is in how the passes
    // overlap their loads, not in bytes)
    const long W = fit_row_waves(fn, cg_row_waves(g.Nx, g.Ny), e.num_cus, 2);
    const int nb = (int)((W + 3) / 4);
    *nb_out = nb;
    const R k2 = R(kappa2), ep = R(eps), h = R(H);
    const R *em = (const R *)eps_map, *ae = (const R *)abei, *a = (const R *)ab;
    const cplx<R> *p = (const cplx<R> *)psi, *jp = (const cplx<R> *)jac_prev;
    cplx<R> *jc = (cplx<R> *)jac;
    R *jA = withA ? (R *)jacA : nullptr;
    const R *jAp = withA ? (const R *)jacA_prev : nullptr;
    if (withA && upd)
        hipLaunchKernelGGL((k_cg_grad_rows<R, true, true>), dim3(nb), dim3(RED_BLOCK), 0, e.stream, G, k2, ep, em, h, mt,
                           p, ae, a, jc, jp, jA, jAp, e.partials, S, alpha, W);
    else if (withA)
        hipLaunchKernelGGL((k_cg_grad_rows<R, true, false>), dim3(nb), dim3(RED_BLOCK), 0, e.stream, G, k2, ep, em, h,
                           mt, p, ae, a, jc, jp, jA, jAp, e.partials, S, alpha, W);
    else if (upd)
        hipLaunchKernelGGL((k_cg_grad_rows<R, false, true>), dim3(nb), dim3(RED_BLOCK), 0, e.stream, G, k2, ep, em, h,
                           mt, p, ae, a, jc, jp, jA, jAp, e.partials, S, alpha, W);
    else
        hipLaunchKernelGGL((k_cg_grad_rows<R, false, false>), dim3(nb), dim3(RED_BLOCK), 0, e.stream, G, k2, ep, em, h,
                           mt, p, ae, a, jc, jp, jA, jAp, e.partials, S, alpha, W);
    GLX_LAUNCH_CHECK("k_cg_grad_rows");
    return GLX_OK;
}

template <typename R>
int L<R>::cg_update_gradient_pass(const Env &e, glx_grid g, double kappa2, double eps, const void *eps_map, double H,
                                  const uint8_t *mt, const void *psi, const void *abei, const void *ab,
                                  const void *dpsi, const void *dab, const double *alpha_dev, void *psi_out,
                                  void *ab_out, void *jac, const void *jac_prev, void *jacA, const void *jacA_prev,
                                  double *beta_next_dev, double *F) {
    const bool withA = kappa2 > 0.0;
    if (!grid_ok(g) || !psi || !jac || !jac_prev || !beta_next_dev || !dpsi || !alpha_dev || !psi_out)
        return set_error(GLX_E_ARG, "glx_cg_update_gradient_pass: invalid arguments");
    if (withA && (!(abei || ab) || !jacA || !jacA_prev))
        return set_error(GLX_E_ARG, "glx_cg_update_gradient_pass: finite kappa needs the vector potential and its Jacobians");
    if (dab && (!ab || !ab_out))
        return set_error(GLX_E_ARG, "glx_cg_update_gradient_pass: an A direction needs ab and ab_out");
    if (psi_out == psi || psi_out == dpsi || (dab && (ab_out == ab || ab_out == dab)))
        return set_error(GLX_E_ARG, "glx_cg_update_gradient_pass: the updated state needs buffers of its own");
    const StateUpdate<R> S{(const cplx<R> *)dpsi, (const R *)dab, (cplx<R> *)psi_out, (R *)ab_out, R(0), R(0)};
    int nb = 0;
    int rc = gradient_pass_launch<R>(e, g, kappa2, eps, eps_map, H, mt, psi, abei, ab, jac, jac_prev, jacA, jacA_prev,
                                     S, alpha_dev, &nb);
    if (rc) return rc;
    const double dxdy = double(R(g.dx)) * double(R(g.dy));
    hipLaunchKernelGGL(k_cg_gradient_finish<R>, dim3(1), dim3(RED_BLOCK), 0, e.stream, e.partials, nb, dxdy, e.result,
                       beta_next_dev);
    GLX_LAUNCH_CHECK("k_cg_gradient_finish");
    if (F) return fetch_result(e, F, 1);
    // asynchronous: the energy lands in pinned memory; glx_cg_energy_wait
    // waits for this copy only, not for work enqueued after it
    hipError_t st = hipMemcpyAsync(e.energy_host, e.result, sizeof(double), hipMemcpyDeviceToHost, e.stream);
    if (st == hipSuccess) st = hipEventRecord(e.ev_energy, e.stream);
    return check_hip(st, "glx_cg_update_gradient_pass energy readback");
}

template <typename R>
int L<R>::cg_gradient_pass(const Env &e, glx_grid g, double kappa2, double eps, const void *eps_map, double H,
                           const uint8_t *mt, const void *psi, const void *abei, const void *ab, void *jac,
                           const void *jac_prev, void *jacA, const void *jacA_prev, double *beta_next_dev, double *F) {
    const bool withA = kappa2 > 0.0;  // the psi Jacobian reads A in both cases
    if (!grid_ok(g) || !psi || !jac || !jac_prev || !beta_next_dev || !F)
        return set_error(GLX_E_ARG, "glx_cg_gradient_pass: invalid arguments");
    if (withA && (!(abei || ab) || !jacA || !jacA_prev))
        return set_error(GLX_E_ARG, "glx_cg_gradient_pass: finite kappa needs the vector potential and its Jacobians");
    int nb = 0;
    int rc = gradient_pass_launch<R>(e, g, kappa2, eps, eps_map, H, mt, psi, abei, ab, jac, jac_prev, jacA, jacA_prev,
                                     StateUpdate<R>{}, nullptr, &nb);
    if (rc) return rc;
    const double dxdy = double(R(g.dx)) * double(R(g.dy));
    hipLaunchKernelGGL(k_cg_gradient_finish<R>, dim3(1), dim3(RED_BLOCK), 0, e.stream, e.partials, nb, dxdy, e.result,
                       beta_next_dev);
    GLX_LAUNCH_CHECK("k_cg_gradient_finish");
    return fetch_result(e, F, 1);
}

template <typename R> int L<R>::pr_beta(const Env &e, const void *x, const void *y, int64_t n, double *beta_dev) {
    if (!x || !y || !beta_dev || n < 0) return set_error(GLX_E_ARG, "glx_cg_pr_beta: invalid arguments");
    const int nb = grid_blocks(n, RED_BLOCK, RED_MAX_BLOCKS);
    hipLaunchKernelGGL(k_pr_sums<R>, dim3(nb), dim3(RED_BLOCK), 0, e.stream, (const R *)x, (const R *)y, (long)n, e.partials);
    GLX_LAUNCH_CHECK("k_pr_sums");
    hipLaunchKernelGGL(k_pr_beta_finish<R>, dim3(1), dim3(RED_BLOCK), 0, e.stream, e.partials, nb, beta_dev);
    GLX_LAUNCH_CHECK("k_pr_beta_finish");
    return GLX_OK;
}

template <typename R>
int L<R>::axmy_dev(const Env &e, const void *x, const void *y, void *z, const double *alpha_dev, int64_t n) {
    if (!x || !y || !z || !alpha_dev || n < 0) return set_error(GLX_E_ARG, "glx_axmy_dev: invalid arguments");
    if (n == 0) return GLX_OK;
    hipLaunchKernelGGL(k_axmy_dev<R>, dim3(grid_blocks(n)), dim3(256), 0, e.stream, (const R *)x, (const R *)y, (R *)z,
                       alpha_dev, (long)n);
    GLX_LAUNCH_CHECK("k_axmy_dev");
    return GLX_OK;
}

template <typename R>
int L<R>::cg_line_step(const Env &e, glx_grid g, double kappa2, double eps, double H, const uint8_t *mt, void *psi,
                       const void *dpsi, const void *abei, void *ab, const void *dab, double *alpha_dev) {
    const bool withA = dab != nullptr;
    if (!grid_ok(g) || !psi || !dpsi || !alpha_dev) return set_error(GLX_E_ARG, "glx_cg_line_step: invalid arguments");
    if (withA && !ab) return set_error(GLX_E_ARG, "glx_cg_line_step: a vector potential direction needs ab");
    if (kappa2 > 0.0 && !(abei || ab)) return set_error(GLX_E_ARG, "glx_cg_line_step: finite kappa needs ab");
    const Geo G = make_geo(g.Nx, g.Ny, g.dx, g.dy);
    const int nb = grid_blocks(G.N, RED_BLOCK, RED_MAX_BLOCKS);
    const double dxdy = double(R(g.dx)) * double(R(g.dy));
    if (withA) {
        hipLaunchKernelGGL((k_coef<R, 17>), dim3(nb), dim3(RED_BLOCK), 0, e.stream, G, R(kappa2), R(eps), R(H), mt,
                           (const cplx<R> *)psi, (const cplx<R> *)dpsi, (const R *)abei, (const R *)ab, (const R *)dab,
                           e.partials);
        GLX_LAUNCH_CHECK("k_coef17");
        hipLaunchKernelGGL(k_line_search<17>, dim3(1), dim3(RED_BLOCK), 0, e.stream, e.partials, nb, dxdy, e.result,
                           alpha_dev);
    } else {
        hipLaunchKernelGGL((k_coef<R, 5>), dim3(nb), dim3(RED_BLOCK), 0, e.stream, G, R(kappa2), R(eps), R(H), mt,
                           (const cplx<R> *)psi, (const cplx<R> *)dpsi, (const R *)abei, (const R *)ab,
                           (const R *)nullptr, e.partials);
        GLX_LAUNCH_CHECK("k_coef5");
        hipLaunchKernelGGL(k_line_search<5>, dim3(1), dim3(RED_BLOCK), 0, e.stream, e.partials, nb, dxdy, e.result,
                           alpha_dev);
    }
    GLX_LAUNCH_CHECK("k_line_search");
    const long n2 = 2 * G.N;
    hipLaunchKernelGGL(k_axpy_dev<R>, dim3(grid_blocks(n2)), dim3(256), 0, e.stream, (const R *)dpsi, (const R *)psi,
                       (R *)psi, (const double *)alpha_dev, n2);
    GLX_LAUNCH_CHECK("k_axpy_dev");
    if (withA) {
        hipLaunchKernelGGL(k_axpy_dev<R>, dim3(grid_blocks(G.Nab)), dim3(256), 0, e.stream, (const R *)dab,
                           (const R *)ab, (R *)ab, (const double *)alpha_dev + 1, (long)G.Nab);
        GLX_LAUNCH_CHECK("k_axpy_dev");
    }
    return GLX_OK;
}

template <typename R>
int L<R>::cg_direction_step(const Env &e, glx_grid g, double kappa2, double eps, double H, const uint8_t *mt,
                            const void *psi, const void *abei, const void *ab, const void *dpsi, const void *jac, void *dpsi_out,
                            const void *dab, const void *jacA, void *dab_out, const double *beta_dev, double *c,
                            double *alpha_dev) {
    const bool withA = dab != nullptr;
    if (!grid_ok(g) || !psi || !dpsi || !jac || !dpsi_out || !beta_dev || (!c && !alpha_dev))
        return set_error(GLX_E_ARG, "glx_cg_direction_step: invalid arguments");
    if (withA && (!ab || !jacA || !dab_out))
        return set_error(GLX_E_ARG, "glx_cg_direction_step: a vector potential direction needs ab, jacA and dab_out");
    if (dpsi_out == dpsi || (withA && dab_out == dab))
        return set_error(GLX_E_ARG, "glx_cg_direction_step: the new direction needs a buffer of its own");
    if (kappa2 > 0.0 && !(abei || ab)) return set_error(GLX_E_ARG, "glx_cg_direction_step: finite kappa needs ab");
    const Geo G = make_geo(g.Nx, g.Ny, g.dx, g.dy);
    const long W = fit_row_waves(withA ? (const void *)k_coef_dir_rows<R, 17> : (const void *)k_coef_dir_rows<R, 5>,
                                 cg_row_waves(g.Nx, g.Ny), e.num_cus);
    const int nb = (int)((W + 3) / 4);
    const double dxdy = double(R(g.dx)) * double(R(g.dy));
    DirUpdate<R> U{(const cplx<R> *)dpsi, (const cplx<R> *)jac, (const R *)dab, (const R *)jacA,
                   (cplx<R> *)dpsi_out, (R *)dab_out, R(0), R(0)};
    const int nc = withA ? 17 : 5;
    // the reference's coefficient kernels read only the scalar epsilon
    if (withA)
        hipLaunchKernelGGL((k_coef_dir_rows<R, 17>), dim3(nb), dim3(RED_BLOCK), 0, e.stream, G, R(kappa2), R(eps),
                           (const R *)nullptr, R(H), mt, (const cplx<R> *)psi, (const R *)abei, (const R *)ab, U,
                           beta_dev, e.partials, W);
    else
        hipLaunchKernelGGL((k_coef_dir_rows<R, 5>), dim3(nb), dim3(RED_BLOCK), 0, e.stream, G, R(kappa2), R(eps),
                           (const R *)nullptr, R(H), mt, (const cplx<R> *)psi, (const R *)abei, (const R *)ab, U,
                           beta_dev, e.partials, W);
    GLX_LAUNCH_CHECK("k_coef_dir_rows");
    if (c) {
        hipLaunchKernelGGL(k_finish_partials, dim3(1), dim3(RED_BLOCK), 0, e.stream, e.partials, nb, nc, dxdy, e.result);
        GLX_LAUNCH_CHECK("k_finish_partials");
        return fetch_result(e, c, nc);
    }
    if (withA)
        hipLaunchKernelGGL(k_line_search<17>, dim3(1), dim3(RED_BLOCK), 0, e.stream, e.partials, nb, dxdy, e.result,
                           alpha_dev);
    else
        hipLaunchKernelGGL(k_line_search<5>, dim3(1), dim3(RED_BLOCK), 0, e.stream, e.partials, nb, dxdy, e.result,
                           alpha_dev);
    GLX_LAUNCH_CHECK("k_line_search");
    return GLX_OK;
}

template <typename R>
int L<R>::axpy(const Env &e, const void *x, const void *y, void *z, double alpha, int64_t n, int minus) {
    if (!x || !y || !z || n < 0) return set_error(GLX_E_ARG, "glx_axpy: invalid arguments");
    if (n == 0) return GLX_OK;
    hipLaunchKernelGGL(k_axpy<R>, dim3(grid_blocks(n)), dim3(256), 0, e.stream, (const R *)x, (const R *)y, (R *)z, R(alpha), (long)n, minus);
    GLX_LAUNCH_CHECK("k_axpy");
    return GLX_OK;
}

}  // namespace glx
