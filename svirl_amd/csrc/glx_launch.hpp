// Host-side launchers shared between the precision translation units
// (glx_f64.hip, glx_f32.hip) and the C API (glx_api.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/glx.h"
#include "glx_limits.hpp"

namespace glx {

// Library-owned device scratch handed to every launcher.
struct Env {
    hipStream_t stream;
    double *partials;        // RED_MAX_BLOCKS * 32 doubles
    double *result;          // 64 doubles
    unsigned long long *arrive; // grid barrier arrival words (MAX_GRID_BLOCKS)
    unsigned long long *rel;    // grid barrier release words (NREL, REL_STRIDE apart)
    int *err;                   // grid barrier error flag
    int *state;              // persistent-kernel state
    int *sweeps;             // per-step sweep counts (capacity sweeps_cap steps)
    int sweeps_cap;
    int num_cus;
    hipEvent_t kev0, kev1;   // around the persistent TD kernel
    float *last_kernel_ms;   // host slot written by td_run
    // per-launch timing of the TD solves (glx_td_profile): an event before
    // and after every solve launch; td_run sums them per kind into solve_ms
    // (psi, A) and counts the launches in solve_n
    bool td_profile;
    hipEvent_t *tev;
    int tev_cap;
    double *solve_ms;
    int *solve_n;
    // memory passes of the last td_run (psi passes, psi redone, A passes, A
    // redone, A passes with one more A buffer), or NULL
    int *passes;
    // asynchronous energy readback of glx_cg_update_gradient_pass: the copy
    // goes to pinned host memory and ev_energy is recorded after it
    double *energy_host;
    hipEvent_t ev_energy;
    // the next TD run (single or batched) starts with the barrier error flag
    // raised, so every team leaves at its first barrier as after a timeout
    // (glx_td_force_barrier_abort; cleared by that run)
    int abort_next;
    // a device buffer kept between calls and grown on demand (the vortex
    // detector's per-row masks); grow_workspace() below
    void **grow;
    size_t *grow_bytes;
};

// Device memory a launcher uses in place of allocating its own per-call
// block: ptr NULL = allocate as usual. `what` names the entry point in
// error messages.
struct Workspace {
    void *ptr;
    size_t bytes;
    const char *what;
};

template <typename R> struct L {
    static int psi_sweep(const Env &e, glx_grid g, double dt, double eps, const void *eps_map, const uint8_t *mt,
                         const void *ab, void *psi_rhs, const void *psi, void *psi_next, double langevin_c,
                         uint32_t jstep, uint32_t rand_t, double stop_eps, int32_t *r2max);
    static int a_sweep(const Env &e, glx_grid g, double dt, double kappa2, double rho, double H, const uint8_t *mt,
                       const void *psi, const void *src, void *ab_rhs, const void *ab, void *ab_next,
                       double langevin_c, uint32_t jstep, uint32_t rand_t, double stop_eps, int32_t *r2max);
    static int td_run(const Env &e, glx_grid g, const glx_td_params *p, const void *eps_map, const uint8_t *mt,
                      void *const P[3], int32_t *pc, void *const Q[3], int32_t *qc, void *rhs_psi, void *rhs_ab,
                      int32_t nsteps, int32_t *sweeps, int32_t wg, int32_t wg_size);
    static int batch_td_run(const Env &e, glx_batch_instance *inst, int32_t n, int32_t nsteps, int32_t *sweeps,
                            int32_t use_graph, int32_t wg_per_cu, int32_t min_rows, const double *weight,
                            int32_t psi_depth, const Workspace &ws);
    static int batch_cg_run(const Env &e, glx_batch_cg_instance *inst, int32_t n, int32_t max_iters, double rtol,
                            int32_t iters_per_launch, int32_t stall_iters, int32_t fp64_finish);
    static int td_max_grid(const Env &e, int32_t wg_size, int32_t *max_wg);
    static int phase_lock(const Env &e, void *psi, int32_t n, const int32_t *ns);
    static int magnetic_field(const Env &e, glx_grid g, const void *abei, const void *ab, void *B);
    static int current_density(const Env &e, glx_grid g, double kappa2, double H, const void *abei, const void *ab,
                               void *J);
    static int supercurrent_density(const Env &e, glx_grid g, const uint8_t *mt, const void *psi, const void *abei,
                                    const void *ab, void *Js);
    static int bulk_observables(const Env &e, glx_grid g, double kappa2, double H, const uint8_t *mt, const void *psi,
                                const void *abei, const void *ab, void *B, void *J, void *Js);
    static int free_energy(const Env &e, glx_grid g, double kappa2, double eps, const void *eps_map, double H,
                           const uint8_t *mt, const void *psi, const void *abei, const void *ab, double *F);
    static int vortex_cells(const Env &e, glx_grid g, const void *psi, const void *ab, const void *abi, double H,
                            double *rec, int32_t capacity, int32_t *count);
    static int vortex_triangulate(const Env &e, glx_grid g, const double *rec, int32_t count, double *xyv);
    static int sum(const Env &e, const void *x, int64_t n, int32_t bs, double *out);
    static int sum_v(const Env &e, const void *x, int64_t nv, int32_t ne, int32_t bs, double *out);
    static int pr_sums(const Env &e, const void *x, const void *y, int64_t n, double *out);
    static int jacobian_psi(const Env &e, glx_grid g, double kappa2, double eps, const void *eps_map, double H,
                            const uint8_t *mt, const void *psi, const void *abei, const void *ab, void *G);
    static int jacobian_A(const Env &e, glx_grid g, double kappa2, double H, const uint8_t *mt, const void *psi,
                          const void *abei, const void *ab, void *G);
    static int coef_psi(const Env &e, glx_grid g, double kappa2, double eps, double H, const uint8_t *mt,
                        const void *psi, const void *dpsi, const void *abei, const void *ab, double *c5);
    static int coef(const Env &e, glx_grid g, double kappa2, double eps, double H, const uint8_t *mt, const void *psi,
                    const void *dpsi, const void *abei, const void *ab, const void *dab, double *c17);
    static int pr_beta(const Env &e, const void *x, const void *y, int64_t n, double *beta_dev);
    static int axmy_dev(const Env &e, const void *x, const void *y, void *z, const double *alpha_dev, int64_t n);
    static int cg_line_step(const Env &e, glx_grid g, double kappa2, double eps, double H, const uint8_t *mt, void *psi,
                            const void *dpsi, const void *abei, void *ab, const void *dab, double *alpha_dev);
    static int cg_direction_step(const Env &e, glx_grid g, double kappa2, double eps, double H, const uint8_t *mt,
                                 const void *psi, const void *abei, const void *ab, const void *dpsi, const void *jac,
                                 void *dpsi_out, const void *dab, const void *jacA, void *dab_out,
                                 const double *beta_dev, double *c, double *alpha_dev);
    static int cg_update_gradient_pass(const Env &e, glx_grid g, double kappa2, double eps, const void *eps_map,
                                       double H, const uint8_t *mt, const void *psi, const void *abei, const void *ab,
                                       const void *dpsi, const void *dab, const double *alpha_dev, void *psi_out,
                                       void *ab_out, void *jac, const void *jac_prev, void *jacA,
                                       const void *jacA_prev, double *beta_next_dev, double *F);
    static int cg_gradient_pass(const Env &e, glx_grid g, double kappa2, double eps, const void *eps_map, double H,
                                const uint8_t *mt, const void *psi, const void *abei, const void *ab, void *jac,
                                const void *jac_prev, void *jacA, const void *jacA_prev, double *beta_next_dev,
                                double *F);
    static int axpy(const Env &e, const void *x, const void *y, void *z, double alpha, int64_t n, int minus);
};

// Error reporting implemented in glx_api.hip.
int set_error(int code, const char *fmt, ...);
int check_hip(hipError_t st, const char *what);

}  // namespace glx
