/*
 * glx: MI355X (gfx950) kernels for svirl's 2D Ginzburg-Landau solver.
 *
 * A C ABI over hand-written HIP kernels. Plain device pointers and sizes
 * only; every entry point returns 0 on success or a nonzero GLX_E* / HIP
 * error code, with a message from glx_last_error().
 *
 * Device buffers use the reference's flat layouts (so the Python side can
 * hand over the same buffers svirl's kernels received):
 *   psi : complex (re, im interleaved), node n = i + Nx*j
 *   ab  : real, a at i + (Nx-1)*j, then b at Na + i + Nx*j, Na = (Nx-1)*Ny
 *   mt  : uint8 (0/1) per cell, i + (Nx-1)*j; NULL = no material tiling
 *   eps_map : real per node; NULL = scalar epsilon
 *   abei : real, same layout as ab; NULL = absent
 * "real" is double for GLX_F64 and float for GLX_F32 (the FP32-mixed form:
 * fields and stencil arithmetic in float, all reductions, residuals and
 * line-search coefficients in double).
 *
 * Reference interfaces replaced (paths relative to the svirl repository):
 * the kernels are svirl/cuda/{td,cg,observables,utils,reduction,
 * block_reduction}.h, launched through pycuda by
 * svirl/parallel/startup.py:74-82 (get_function), with geometry baked in
 * at JIT time (startup.py:54-63); here the geometry is a glx_grid
 * argument. All calls run on one library-owned HIP stream; calls that return
 * host values synchronise it.
 */
#ifndef GLX_H
#define GLX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { GLX_F64 = 0, GLX_F32 = 1 };

enum {
    GLX_OK = 0,
    GLX_E_ARG = -1,      /* invalid argument (size, NULL, precision) */
    GLX_E_HIP = -2,      /* a HIP runtime call failed */
    GLX_E_NOINIT = -3,   /* glx_init not called */
    GLX_E_TIMEOUT = -4,  /* a grid barrier timed out inside a persistent kernel */
    GLX_E_CAPACITY = -5  /* an output list was too small */
};

/* Grid geometry; replaces the %(Nx)s, %(Ny)s, %(dx)s, %(dy)s template
 * constants of svirl/parallel/startup.py:54-60. */
typedef struct glx_grid {
    int32_t Nx, Ny;
    double dx, dy;
} glx_grid;

/* ---------------------------------------------------------------- runtime */

/* Select the device and create the library stream
 * (replaces Startup.__init_device, svirl/parallel/startup.py:87-93). */
int glx_init(int device);
const char *glx_last_error(void);
int glx_synchronize(void);
/* Number of compute units of the current device (for launch sizing). */
int glx_device_info(int *num_cus, int *max_wg_per_cu_hint, char *name, int name_len);

/* Device memory owned by the caller (replaces pycuda gpuarray allocation,
 * svirl/storage/arrays.py:574-591, and Utils.copy_dtod,
 * svirl/parallel/utils.py:22-27). */
int glx_malloc(void **ptr, size_t bytes);
int glx_free(void *ptr);
int glx_memcpy_h2d(void *dst, const void *src, size_t bytes);
int glx_memcpy_d2h(void *dst, const void *src, size_t bytes);
int glx_memcpy_d2d(void *dst, const void *src, size_t bytes);
int glx_memset(void *dst, int value, size_t bytes);

/* Stream-ordered HIP event timing of library work (milliseconds). */
int glx_timer_start(void);
int glx_timer_stop(float *ms);

/* ------------------------------------------------------------------- TDGL */

/* One Jacobi sweep of backward-Euler psi, with the reference's residual:
 * atomicMax(r2max, int(min(1e4*max|dpsi|/stop_eps, 1e8))). r2max is a device
 * int32 the caller zeroes. With langevin_c > 1e-32 and jstep == 0 the noisy
 * right-hand side is written back to psi_rhs.
 * Replaces iterate_order_parameter_jacobi_step, svirl/cuda/td.h:4-133,
 * launched at svirl/solvers/td.py:168-187. */
int glx_td_psi_sweep(int prec, glx_grid g, double dt, double eps, const void *eps_map,
                     const uint8_t *mt, const void *ab, void *psi_rhs, const void *psi,
                     void *psi_next, double langevin_c, uint32_t jstep, uint32_t rand_t,
                     double stop_eps, int32_t *r2max);

/* One Jacobi sweep of the vector potential. abi_ab_rhs is the supercurrent
 * source buffer and may alias ab_next (each thread reads its own source
 * links before it writes them, as the reference's ping-pong relies on).
 * Replaces iterate_vector_potential_jacobi_step, svirl/cuda/td.h:310-463,
 * launched at svirl/solvers/td.py:277-300. */
int glx_td_a_sweep(int prec, glx_grid g, double dt, double kappa2, double rho, double H,
                   const uint8_t *mt, const void *psi, const void *abi_ab_rhs, void *ab_rhs,
                   const void *ab, void *ab_next, double langevin_c, uint32_t jstep,
                   uint32_t rand_t, double stop_eps, int32_t *r2max);

/* psi[n] <- |psi[n]| at lock_ns[0..lock_N) (svirl/cuda/td.h:295-307). */
int glx_td_phase_lock(int prec, void *psi, int32_t lock_N, const int32_t *lock_ns);

typedef struct glx_td_params {
    double dt;
    double eps;          /* scalar epsilon (ignored where eps_map is given) */
    double kappa2;       /* kappa^2; only read when solveA */
    double rho;          /* 1/sigma */
    double H;
    double stop_psi;     /* stop criteria after the td.py:51-66 clamps */
    double stop_A;
    double langevin_psi;
    double langevin_A;
    uint32_t rand_t;     /* TD._random_t at entry */
    int32_t solveA;      /* kappa < inf */
    int32_t max_sweeps;  /* Jacobi cap per solve; the reference uses 1024 */
    int32_t eqn;         /* 0: psi then A, 1: psi only, 2: A only */
    const void *a_src;   /* NULL: the A sweeps take their supercurrent source
                          * from A^(2*floor(j/2)) (td.py:264-303); else a
                          * buffer of Nab reals (not one of ab_bufs) that
                          * every A sweep reads instead: svirl's fixed-vortex
                          * path binds vpi + A^n there (td.py:256-266).
                          * Not supported by glx_batch_td_run. */
} glx_td_params;

/* nsteps TDGL steps in one persistent launch: per step up to max_sweeps psi
 * sweeps and (if solveA) up to max_sweeps A sweeps, each sweep followed by a
 * grid barrier and the reference's stop rule, with no host round trip.
 * psi_bufs / ab_bufs are three same-sized buffers rotated in place of the
 * reference's ψ_rhs copy and pointer swaps; *psi_cur / *ab_cur index the one
 * holding the state on entry and are updated on return. rhs_psi / rhs_ab are
 * only used with Langevin noise (may be NULL otherwise). sweeps (host,
 * 2*nsteps int32) receives (S_psi, S_A) per step.
 * Replaces the host loops TD.__iterate / __iterate_order_parameter_gpu /
 * __iterate_vector_potential_gpu, svirl/solvers/td.py:157-366, with the
 * A-source aliasing of td.py:264-303 kept. wg / wg_size select the
 * persistent grid of both solves (wg 0: the library default, two
 * workgroups per CU for the psi solve and three for the A solve, within
 * the co-resident limit). */
int glx_td_run(int prec, glx_grid g, const glx_td_params *p, const void *eps_map,
               const uint8_t *mt, void *const psi_bufs[3], int32_t *psi_cur,
               void *const ab_bufs[3], int32_t *ab_cur, void *rhs_psi, void *rhs_ab,
               int32_t nsteps, int32_t *sweeps, int32_t wg, int32_t wg_size);

/* ------------------------------------------------------------ batched TD */

/* One instance of a batched TD run (no counterpart in the reference, which
 * runs one GLSolver per process; this is the batched form of glx_td_run).
 * Each instance has its own grid, parameters, epsilon map, mask and buffer
 * triplets; psi[psi_cur] / ab[ab_cur] hold its state on entry and exit.
 * Langevin noise and eqn modes other than 0 are not supported here. */
typedef struct glx_batch_instance {
    glx_grid g;
    glx_td_params p;
    const void *eps_map;
    const uint8_t *mt;
    void *psi[3];
    void *ab[3];
    int32_t psi_cur;
    int32_t ab_cur;
} glx_batch_instance;

/* Advance n instances by nsteps TD steps. Every step is ONE persistent
 * launch in which each instance is swept by its own team of co-resident
 * workgroups with team-local barriers, so instances stop their Jacobi
 * solves independently (ragged sweep counts). Teams are sized in
 * proportion to each instance's wave-rows times weight[i] (its relative
 * sweeps per step, e.g. from an earlier run; NULL = 1). With use_graph != 0,
 * the step launches are captured once into a HIP graph and replayed.
 * wg_per_cu (co-resident workgroups per CU, 0 = as many as fit),
 * min_rows (fewest rows per wave when sizing teams, 0 = 16) and psi_depth
 * (sweeps per psi pass, 2..4, i.e. strips of 64 - 2*psi_depth columns;
 * 0 = 3) are the tunable tile shape; the results do not depend on them.
 * sweeps (host, nsteps*n*2 int32, may be NULL) receives (S_psi, S_A) per
 * step and instance. Instances must not share buffers; n must not exceed
 * the number of co-resident workgroups. */
int glx_batch_td_run(int prec, glx_batch_instance *inst, int32_t n, int32_t nsteps, int32_t *sweeps,
                     int32_t use_graph, int32_t wg_per_cu, int32_t min_rows, const double *weight,
                     int32_t psi_depth);

/* Bytes of the fixed device scratch array behind glx_batch_td_run_noalloc. */
#define GLX_NOALLOC_SCRATCH_BYTES (8u << 20)

/* glx_batch_td_run without any device allocation. No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc). Same arguments and results as glx_batch_td_run with
 * use_graph = 0: the run's descriptor table, team map, barrier words, state
 * and sweep log live in a file-scope device array of
 * GLX_NOALLOC_SCRATCH_BYTES (8 MiB, held by each GPU the library runs on
 * until the process ends) instead of a per-call hipMalloc. A run that needs
 * more (about 0.6 KB per instance plus 8 bytes per instance and step) fails
 * with GLX_E_ARG before anything is launched; split nsteps over several
 * calls. Host-side planning uses ordinary host memory. Every call on one
 * GPU shares the array, and the array keeps the last call's contents: each
 * call rewrites all it reads before its first step. Calls that overlap on
 * one GPU (from several threads) race on it. */
int glx_batch_td_run_noalloc(int prec, glx_batch_instance *inst, int32_t n, int32_t nsteps, int32_t *sweeps,
                             int32_t wg_per_cu, int32_t min_rows, const double *weight, int32_t psi_depth);

/* ------------------------------------------------------------ batched CG */

/* One instance of a batched CG minimisation (no counterpart in the
 * reference; the batched form of cg.py:452-551 / 238-322). kappa2 <= 0
 * means kappa = inf (psi only, ab may be NULL); all instances of one call
 * must agree on that. eps_map (may be NULL) is used by the gradient and by
 * the line-search coefficients. iters, converged, grad_rel, grad0, stalled
 * and fp64_iters are outputs. */
typedef struct glx_batch_cg_instance {
    glx_grid g;
    double kappa2;
    double eps;
    double H;
    const void *eps_map;
    const uint8_t *mt;
    void *psi;
    void *ab;
    int32_t iters;
    int32_t converged;
    double grad_rel;
    double grad0;        /* out: ||grad F||_0 of the stop rule */
    int32_t stalled;     /* out: 1 = stopped by the stall rule, not converged */
    int32_t fp64_iters;  /* out: iterations of the FP64 finish (included in iters) */
} glx_batch_cg_instance;

/* Minimise n instances by Polak-Ribiere CG with an exact polynomial line
 * search: each instance is owned by one workgroup that runs whole
 * iterations (gradient, FP64 PR sums and gradient norm, direction, the 17
 * (or 5) FP64 line-search coefficients, an FP64 2-D quartic minimisation on
 * the device, update) and stops on its own when ||grad F|| <= rtol *
 * ||grad F||_0 or after max_iters (a ragged batch). iters_per_launch
 * iterations run per kernel launch (0 = 50).
 * stall_iters > 0 adds a stall rule: an instance whose relative gradient
 * norm has not halved within stall_iters iterations stops (stalled = 1).
 * fp64_finish (FP32-mixed batches with a stall rule only): the stalled
 * instances are copied into FP64 and continue there from their iteration
 * count, with a steepest-descent step and the FP32 run's ||grad F||_0, to
 * rtol or max_iters in all; the FP64 result is rounded back into the FP32
 * psi and ab (so an FP32 re-evaluation of the gradient sits at FP32
 * rounding, not at rtol). */
int glx_batch_cg_run(int prec, glx_batch_cg_instance *inst, int32_t n, int32_t max_iters, double rtol,
                     int32_t iters_per_launch, int32_t stall_iters, int32_t fp64_finish);

/* One psi Jacobi solve to the reference's stop rule, on the device
 * (TD.__iterate_order_parameter_gpu, td.py:157-218: the sweep loop, its
 * r2max read-back per sweep replaced by the device barrier). ab is read
 * only; psi_bufs / *psi_cur as in glx_td_run; *sweeps receives the sweep
 * count. p->eqn is ignored. */
int glx_td_psi_solve(int prec, glx_grid g, const glx_td_params *p, const void *eps_map, const uint8_t *mt,
                     const void *ab, void *const psi_bufs[3], int32_t *psi_cur, void *rhs_psi, int32_t *sweeps);

/* One A Jacobi solve (TD.__iterate_vector_potential_gpu, td.py:252-325),
 * with the reference's source aliasing A^(2*floor(j/2)); psi is read only. */
int glx_td_a_solve(int prec, glx_grid g, const glx_td_params *p, const uint8_t *mt, const void *psi,
                   void *const ab_bufs[3], int32_t *ab_cur, void *rhs_ab, int32_t *sweeps);

/* One TDGL step, psi then (if p->solveA) A (TD.__iterate, td.py:356-361);
 * sweeps[0..1] = (S_psi, S_A). Equivalent to glx_td_run with nsteps = 1. */
int glx_td_step(int prec, glx_grid g, const glx_td_params *p, const void *eps_map, const uint8_t *mt,
                void *const psi_bufs[3], int32_t *psi_cur, void *const ab_bufs[3], int32_t *ab_cur,
                void *rhs_psi, void *rhs_ab, int32_t *sweeps);

/* Duration (ms, HIP events on the library stream) of the persistent kernel
 * of the last glx_td_run call, excluding its reset and read-back. */
int glx_td_last_kernel_ms(float *ms);

/* Per-launch timing of glx_td_run's solve kernels (no counterpart in the
 * reference; a measurement hook). While enabled, glx_td_run brackets every
 * psi-solve and A-solve launch with HIP events on the library stream;
 * glx_td_solve_ms returns the summed durations (ms) and launch counts of the
 * last glx_td_run call (zeros when profiling was off). */
int glx_td_profile(int32_t enable);
int glx_td_solve_ms(double *psi_ms, int32_t *psi_launches, double *a_ms, int32_t *a_launches);

/* Memory passes of the last glx_td_run call (no counterpart in the
 * reference; a measurement hook, like glx_td_solve_ms): the psi solves'
 * passes over the fields (a fused pass of two or three sweeps counts once)
 * and the redone ones among them (a pass repeated up to an earlier sweep
 * that met the stop rule), the same two for the A solves, and the A passes
 * that move one more A buffer (a three-sweep pass stores A^2; an odd sweep
 * j >= 3, or any sweep with a fixed source, reads a separate source). The
 * compulsory bytes of a psi pass are (8s + e + m) per node, of an A pass
 * (8s + m), plus 2s for each such extra buffer (s: bytes per real; e = s
 * with an epsilon map; m = 1 with a mask). */
int glx_td_passes(int32_t *psi_passes, int32_t *psi_redo, int32_t *a_passes, int32_t *a_redo, int32_t *a_xbuf);

/* Bandwidth of a plain streaming kernel on this device (no counterpart in
 * the reference; a measurement hook giving the practical ceiling next to
 * the HBM peak): kind 0 reads n float4, kind 1 copies n float4, kind 2
 * streams the traffic of one FP32 psi pass over n nodes (psi, right-hand
 * side, a, b, eps and a mask byte read, psi written: 37 bytes per node).
 * grid workgroups of 256 threads, reps timed launches after one warm-up;
 * *gbs receives bytes moved / time in GB/s. Allocates and frees its own
 * buffers. */
int glx_stream_probe(int32_t kind, int64_t n, int32_t grid, int32_t reps, double *gbs);

/* Failure-handling hook (no counterpart in the reference; for tests of the
 * timeout path). The next glx_td_run or glx_batch_td_run call (including
 * one made by glx_td_step / glx_td_psi_solve / glx_td_a_solve) starts with
 * the grid barrier's error flag raised, as a barrier poll that ran past its
 * 5 s bound would leave it: every multi-workgroup team leaves its solve at
 * its first barrier (a one-workgroup team of a batch never polls the flag
 * and runs on), the kernels end normally, and the call returns
 * GLX_E_TIMEOUT without updating *psi_cur / *ab_cur or the instances'
 * current indices. Like a real timeout it leaves the run's other buffers
 * undefined (after one step, only buffers that were not current have been
 * written). The run after it is normal. */
int glx_td_force_barrier_abort(void);

/* Largest co-resident grid for glx_td_run at a workgroup size. */
int glx_td_max_grid(int prec, int32_t wg_size, int32_t *max_wg);

/* ------------------------------------------------------------ observables */

/* B per cell (svirl/cuda/observables.h:4-37; observables.py:36-58). */
int glx_magnetic_field(int prec, glx_grid g, const void *abei, const void *ab, void *B);
/* kappa^2 (curl B - H) on links, ghost-cell boundary
 * (observables.h:40-156; observables.py:86-110). */
int glx_current_density(int prec, glx_grid g, double kappa2, double H, const void *abei,
                        const void *ab, void *jxjy);
/* r_Du id Im(conj psi0 U psi1) on links (observables.h:159-235). */
int glx_supercurrent_density(int prec, glx_grid g, const uint8_t *mt, const void *psi,
                             const void *abei, const void *ab, void *jsxjsy);
/* B (cells), J and J_s (links) in one pass over the nodes: the results of
 * glx_magnetic_field, glx_current_density and glx_supercurrent_density, to
 * the bit, with each link value the three share read once (the three
 * kernels of svirl/cuda/observables.h:4-235 behind observables.py:36-121,
 * fused). Any of B, J, Js may be NULL to skip it; psi may be NULL when Js
 * is. No counterpart in the reference: it serves J_n = J - J_s
 * (normalcurrent_density, observables.py:113-121) and callers that read
 * all three. */
int glx_bulk_observables(int prec, glx_grid g, double kappa2, double H, const uint8_t *mt,
                         const void *psi, const void *abei, const void *ab, void *B, void *J,
                         void *Js);

/* Total GL free energy, FP64 accumulation, returned on the host
 * (free_energy_pseudodensity, observables.h:250-362, + Reduction.gsum,
 * svirl/parallel/reduction.py:40-110; observables.py:124-149). */
int glx_free_energy(int prec, glx_grid g, double kappa2, double eps, const void *eps_map,
                    double H, const uint8_t *mt, const void *psi, const void *abei,
                    const void *ab, double *F);

/* Vortex cells: winding number of every cell as in
 * svirl/observables/vortex_detector.py:55-72 (np.angle, np.mod, np.round
 * semantics), computed in FP64, with H first rounded to the precision
 * `prec` as svirl's Params stores it (svirl/vars/params.py:131-132). Flagged cells are written to `rec` as
 * records of 14 doubles: (cell n, v, psi00.re, .im, psip0.re, .im,
 * psipp.re, .im, psi0p.re, .im, a00, a0p, b00, bp0) where the link values
 * are ab + abi (abi may be NULL) for the host triangulation
 * (vortex_detector.py:103-147). Records come in ascending cell id n, the
 * reference's loop order; *count is the total number of flagged cells (may
 * exceed capacity: then GLX_E_CAPACITY and only the first `capacity`
 * records are written). The library keeps a device buffer of 12 bytes per
 * 63 cells plus 4 bytes per cell row between calls. */
int glx_vortex_cells(int prec, glx_grid g, const void *psi, const void *ab, const void *abi,
                     double H, double *rec, int32_t capacity, int32_t *count);

/* Sub-cell vortex positions of `count` records of glx_vortex_cells (in
 * device memory, any order), the triangulation loop of
 * svirl/observables/vortex_detector.py:103-147 with its helpers
 * find_zero, find_zero_line, find_intersection (:75-101), in FP64 with the
 * reference's operation order. Writes 4 doubles per record to `xyv`:
 * (x, y, round(v), keep), keep = 1.0 where the reference appends
 * [x, y, round(v)] to its list (two zero points on each of the Re and Im
 * lines, non-parallel lines, point within the cell's 3x3 neighbourhood),
 * else 0.0. The reference's output is the kept rows in ascending cell id.
 * Asynchronous on the library stream. */
int glx_vortex_triangulate(glx_grid g, const double *rec, int32_t count, double *xyv);

/* ------------------------------------------------------------- reductions */

/* Sum of n reals, FP64 accumulation, deterministic order (Reduction.gsum,
 * svirl/parallel/reduction.py:40-110; block_reduction.h:80-92).
 * block_size 0 = default; any multiple of 64 up to 1024 is accepted. */
int glx_sum(int prec, const void *x, int64_t n, int32_t block_size, double *out);
/* Column sums of an nv x ne row-major real array (Reduction.gsum_v,
 * reduction.py:115-173; reduction.h:56-82). out: ne doubles on the host. */
int glx_sum_v(int prec, const void *x, int64_t nv, int32_t ne, int32_t block_size, double *out);
/* Polak-Ribiere sums over n reals: out[0] = sum x*(x-y), out[1] = sum y*y
 * (xmyx_*_sum + x_mag2_*_sum + gsum, svirl/cuda/utils.h:12-70,
 * cg.py:422-449). A complex array is passed as 2N reals. */
int glx_pr_sums(int prec, const void *x, const void *y, int64_t n, double *out);

/* ----------------------------------------------------------------- CG */

/* G += dxdy dF/dpsi (free_energy_jacobian_psi, svirl/cuda/cg.h:15-121). */
int glx_cg_jacobian_psi(int prec, glx_grid g, double kappa2, double eps, const void *eps_map,
                        double H, const uint8_t *mt, const void *psi, const void *abei,
                        const void *ab, void *G);
/* G += 2 dxdy [kappa^2 curlcurl - r id js] (free_energy_jacobian_A,
 * cg.h:124-301). */
int glx_cg_jacobian_A(int prec, glx_grid g, double kappa2, double H, const uint8_t *mt,
                      const void *psi, const void *abei, const void *ab, void *G);
/* c0..c4 of F(psi + alpha dpsi) (free_energy_conjgrad_coef_psi,
 * cg.h:314-474, + gsum_v; cg.py:184-224). eps is the scalar the reference
 * passes (0 when an epsilon map is set; the kernel never reads the map). */
int glx_cg_coef_psi(int prec, glx_grid g, double kappa2, double eps, double H,
                    const uint8_t *mt, const void *psi, const void *dpsi, const void *abei,
                    const void *ab, double *c5);
/* The 17 coefficients of free_energy_conjgrad_coef (cg.h:477-731;
 * cg.py:325-375), in kernel order c00..c04, c10..c14, c20..c24, c30, c40. */
int glx_cg_coef(int prec, glx_grid g, double kappa2, double eps, double H, const uint8_t *mt,
                const void *psi, const void *dpsi, const void *abei, const void *ab,
                const void *dab, double *c17);
/* z = alpha*x + y and z = alpha*x - y over n reals (axpy_r/axpy_c and
 * axmy_r/axmy_c, svirl/cuda/utils.h:73-114; a complex array is 2N reals). */
int glx_axpy(int prec, const void *x, const void *y, void *z, double alpha, int64_t n);
int glx_axmy(int prec, const void *x, const void *y, void *z, double alpha, int64_t n);
/* The gradient pass of one CG iteration, fused: both Jacobians
 * (free_energy_jacobian_psi / _A, cg.h:15-301; written, which equals the
 * reference's += into a zeroed buffer), the free energy of the state
 * (observables.py:124-149, returned in *F, FP64) and the Polak-Ribiere sums
 * against jac_prev / jacA_prev, whose betas max(num/den, 0) go to
 * beta_next_dev[0..1] (device, FP64 rounded to the working precision,
 * cg.py:422-449). kappa2 <= 0: psi only (jacA, jacA_prev unused; the
 * psi Jacobian still reads the fixed vector potential). One
 * pass over the state replaces the reference's gradient, beta and energy
 * launches (cg.py:458-549). */
int glx_cg_gradient_pass(int prec, glx_grid g, double kappa2, double eps, const void *eps_map, double H,
                         const uint8_t *mt, const void *psi, const void *abei, const void *ab, void *jac,
                         const void *jac_prev, void *jacA, const void *jacA_prev, double *beta_next_dev, double *F);
/* Polak-Ribiere beta left in device memory, as the reference keeps it
 * (__compute_beta_psi/_A, cg.py:422-449; divide_scalars_positive,
 * utils.h:139-146): *beta_dev = max(sum x*(x-y) / sum y*y, 0), sums in FP64,
 * the quotient rounded to the working precision and stored as a double. */
int glx_cg_pr_beta(int prec, const void *jac, const void *jac_prev, int64_t n, double *beta_dev);
/* z = alpha*x - y with alpha read from device memory (the reference's
 * axmy_c/axmy_r called with the device beta, cg.py:283-290). */
int glx_axmy_dev(int prec, const void *x, const void *y, void *z, const double *alpha_dev, int64_t n);
/* The direction update of one CG iteration fused into its line-search
 * coefficients: dpsi_out = beta_dev[0]*dpsi - jac and (dab not NULL)
 * dab_out = beta_dev[1]*dab - jacA (axmy_c/axmy_r with the device betas,
 * cg.py:283-290, 491-530), formed where the coefficient pass reads the
 * direction (free_energy_conjgrad_coef_psi / _coef, cg.h:314-731) and
 * written on the way; the outputs must be buffers other than dpsi / dab.
 * Then, with c != NULL: the 5 (dab NULL) or 17 coefficients, reduced in
 * FP64, returned in c (the reference's host line search follows); with
 * c == NULL: the device line search of glx_cg_line_step on them
 * (alpha_dev written; the state is not updated: see
 * glx_cg_update_gradient_pass). No counterpart in the reference as one
 * call: it replaces the axmy launches and the coefficient launch of an
 * iteration. */
int glx_cg_direction_step(int prec, glx_grid g, double kappa2, double eps, double H,
                          const uint8_t *mt, const void *psi, const void *abei, const void *ab,
                          const void *dpsi, const void *jac, void *dpsi_out,
                          const void *dab, const void *jacA, void *dab_out,
                          const double *beta_dev, double *c, double *alpha_dev);
/* The update of one CG iteration fused into the next gradient pass:
 * glx_cg_gradient_pass evaluated at psi + alpha_dev[0]*dpsi and (dab not
 * NULL) ab + alpha_dev[1]*dab (axpy_c/axpy_r, cg.py:296, 535-536, with the
 * alphas read from device memory as doubles and rounded to the working
 * precision), and the updated state written to psi_out / ab_out, buffers
 * other than psi, dpsi, ab and dab, which then hold the state. F NULL:
 * nothing waits; the energy is copied to library-owned pinned memory and
 * glx_cg_energy_wait returns it (work enqueued after this call, such as
 * the next iteration's glx_cg_direction_step, may run meanwhile). No
 * counterpart in the reference as one call. */
int glx_cg_update_gradient_pass(int prec, glx_grid g, double kappa2, double eps,
                                const void *eps_map, double H, const uint8_t *mt, const void *psi,
                                const void *abei, const void *ab, const void *dpsi, const void *dab,
                                const double *alpha_dev, void *psi_out, void *ab_out, void *jac,
                                const void *jac_prev, void *jacA, const void *jacA_prev,
                                double *beta_next_dev, double *F);
/* The energy of the last glx_cg_update_gradient_pass called with F NULL,
 * once its copy has arrived (waits for that copy only). */
int glx_cg_energy_wait(double *F);
/* The line search and update of one CG iteration on the device, for the
 * FP32-mixed form: the line-search coefficients (as glx_cg_coef, or
 * glx_cg_coef_psi when dab is NULL; cg.py:184-224, 325-375), their
 * minimum from (0, 0) found in FP64 on the device (the role of the
 * reference's numpy polyroots / scipy BFGS, cg.py:227-235, 378-419; a
 * damped Newton descent on the quartic), and psi += alpha_psi * dpsi,
 * ab += alpha_A * dab (cg.py:296, 535-536). alpha_dev: two device doubles
 * (alpha_psi, alpha_A) written; the coefficients are left in the library's
 * result buffer. No host synchronisation. */
int glx_cg_line_step(int prec, glx_grid g, double kappa2, double eps, double H, const uint8_t *mt, void *psi,
                     const void *dpsi, const void *abei, void *ab, const void *dab, double *alpha_dev);

#ifdef __cplusplus
}
#endif

#endif /* GLX_H */
