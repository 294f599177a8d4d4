// extern "C" entry points of libglx: argument dispatch by precision, the
// library context (device, stream, scratch) and error reporting.
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <mutex>
#include <vector>

#include "glx_launch.hpp"

namespace glx {

namespace {

struct Ctx {
    bool ready = false;
    int device = -1;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    float last_kernel_ms = 0.f;
    double solve_ms[2] = {0.0, 0.0};
    int solve_n[2] = {0, 0};
    int passes[5] = {0, 0, 0, 0, 0};
    std::vector<hipEvent_t> tev;
    hipDeviceProp_t prop{};
    Env env{};
    void *lib_scratch = nullptr;  // partials, barrier words, run state (hipMalloc at glx_init)
    void *grow = nullptr;         // Env::grow
    size_t grow_bytes = 0;
};

Ctx g_ctx;
std::mutex g_mu;

// Device addresses of `scratch` (below), one per device, taken at glx_init.
constexpr int MAX_DEVICES = 64;
void *g_scratch_dev[MAX_DEVICES] = {};
thread_local char g_err[512] = "";

constexpr int SWEEPS_CAP = 1 << 16;  // steps per persistent launch

}  // namespace

int set_error(int code, const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return code;
}

int check_hip(hipError_t st, const char *what) {
    if (st == hipSuccess) return GLX_OK;
    return set_error(GLX_E_HIP, "%s: %s (%d)", what, hipGetErrorString(st), (int)st);
}

static int ensure_ready() {
    if (!g_ctx.ready) return set_error(GLX_E_NOINIT, "glx_init has not been called");
    return GLX_OK;
}

}  // namespace glx

using namespace glx;

// Workspace of glx_batch_td_run_noalloc: a fixed device array, so that entry
// point needs no device allocation. Its host-side name is not a device
// address; glx_init stores the address hipGetSymbolAddress gives for the
// current device in g_scratch_dev, and only that address is handed to
// kernels and copies. Every call on one device shares the array.
static __device__ __align__(256) unsigned char scratch[GLX_NOALLOC_SCRATCH_BYTES];

// Streaming kernels of known traffic for glx_stream_probe (grid-stride
// loops bounded by the element count).
static __global__ void k_probe_read(const float4 *__restrict__ a, float *out, long n) {
    float acc = 0.f;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const float4 v = a[i];
        acc += (v.x + v.y) + (v.z + v.w);
    }
    if (acc == 1.0e30f) out[0] = acc;  // keeps the loads; never true for the zeroed input
}
static __global__ void k_probe_copy(const float4 *__restrict__ a, float4 *__restrict__ b, long n) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) b[i] = a[i];
}
// The traffic of one FP32 psi pass per node: psi and the right-hand side
// (8 B each), a, b and eps (4 B each) and a mask byte read, psi out (8 B).
static __global__ void k_probe_psi_mix(const float2 *__restrict__ p, const float2 *__restrict__ r,
                                       const float *__restrict__ a, const float *__restrict__ b,
                                       const float *__restrict__ e, const unsigned char *__restrict__ m,
                                       float2 *__restrict__ o, long n) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const float2 pv = p[i], rv = r[i];
        const float s = a[i] + b[i] + e[i] + (float)m[i];
        o[i] = make_float2(pv.x * s + rv.x, pv.y * s + rv.y);
    }
}

#define GLX_REQUIRE_INIT() do { int _r = ensure_ready(); if (_r) return _r; } while (0)
#define GLX_DISPATCH(prec, call)                                                   \
    do {                                                                           \
        GLX_REQUIRE_INIT();                                                        \
        if ((prec) == GLX_F64) return L<double>::call;                             \
        if ((prec) == GLX_F32) return L<float>::call;                              \
        return set_error(GLX_E_ARG, "unknown precision %d", (int)(prec));          \
    } while (0)

extern "C" {

const char *glx_last_error(void) { return g_err; }

int glx_init(int device) {
    std::lock_guard<std::mutex> lock(g_mu);
    if (g_ctx.ready) {
        if (device == g_ctx.device) return GLX_OK;
        return set_error(GLX_E_ARG, "glx already initialised on device %d", g_ctx.device);
    }
    int n = 0;
    hipError_t st = hipGetDeviceCount(&n);
    if (st != hipSuccess) return check_hip(st, "hipGetDeviceCount");
    if (device < 0 || device >= n) return set_error(GLX_E_ARG, "device %d not present (%d devices)", device, n);
    if (device >= MAX_DEVICES) return set_error(GLX_E_ARG, "device %d above the supported %d", device, MAX_DEVICES);
    if ((st = hipSetDevice(device)) != hipSuccess) return check_hip(st, "hipSetDevice");
    hipDeviceProp_t prop;
    if ((st = hipGetDeviceProperties(&prop, device)) != hipSuccess) return check_hip(st, "hipGetDeviceProperties");
    if ((st = hipStreamCreateWithFlags(&g_ctx.stream, hipStreamNonBlocking)) != hipSuccess)
        return check_hip(st, "hipStreamCreate");
    if ((st = hipEventCreate(&g_ctx.ev0)) != hipSuccess) return check_hip(st, "hipEventCreate");
    if ((st = hipEventCreate(&g_ctx.ev1)) != hipSuccess) return check_hip(st, "hipEventCreate");
    if ((st = hipEventCreate(&g_ctx.env.kev0)) != hipSuccess) return check_hip(st, "hipEventCreate");
    if ((st = hipEventCreate(&g_ctx.env.kev1)) != hipSuccess) return check_hip(st, "hipEventCreate");
    g_ctx.env.last_kernel_ms = &g_ctx.last_kernel_ms;
    g_ctx.env.solve_ms = g_ctx.solve_ms;
    g_ctx.env.solve_n = g_ctx.solve_n;
    g_ctx.env.passes = g_ctx.passes;
    g_ctx.env.grow = &g_ctx.grow;
    g_ctx.env.grow_bytes = &g_ctx.grow_bytes;
    // library scratch layout: partials | result | arrive | rel | err | state | sweeps
    const size_t partials = sizeof(double) * 1024 * 32;
    const size_t result = sizeof(double) * 64;
    const size_t arrive = sizeof(unsigned long long) * MAX_GRID_BLOCKS;
    const size_t rel = sizeof(unsigned long long) * NREL * REL_STRIDE;
    const size_t err = 512;
    const size_t state = 512;
    const size_t sweeps = sizeof(int) * 2 * SWEEPS_CAP;
    const size_t total = partials + result + arrive + rel + err + state + sweeps;
    if ((st = hipMalloc(&g_ctx.lib_scratch, total)) != hipSuccess) return check_hip(st, "hipMalloc(library scratch)");
    if ((st = hipMemset(g_ctx.lib_scratch, 0, total)) != hipSuccess) return check_hip(st, "hipMemset(library scratch)");
    char *p = (char *)g_ctx.lib_scratch;
    Env &e = g_ctx.env;
    e.stream = g_ctx.stream;
    e.partials = (double *)p; p += partials;
    e.result = (double *)p; p += result;
    e.arrive = (unsigned long long *)p; p += arrive;
    e.rel = (unsigned long long *)p; p += rel;
    e.err = (int *)p; p += err;
    e.state = (int *)p; p += state;
    e.sweeps = (int *)p; p += sweeps;
    e.sweeps_cap = SWEEPS_CAP;
    e.num_cus = prop.multiProcessorCount;
    if ((st = hipHostMalloc((void **)&e.energy_host, 64, hipHostMallocDefault)) != hipSuccess)
        return check_hip(st, "hipHostMalloc(energy)");
    e.energy_host[0] = 0.0;
    if ((st = hipEventCreateWithFlags(&e.ev_energy, hipEventDisableTiming)) != hipSuccess)
        return check_hip(st, "hipEventCreate");
    if (!g_scratch_dev[device]) {  // with `device` current (hipSetDevice above)
        void *sp = nullptr;
        if ((st = hipGetSymbolAddress(&sp, HIP_SYMBOL(scratch))) != hipSuccess)
            return check_hip(st, "hipGetSymbolAddress(scratch)");
        g_scratch_dev[device] = sp;
    }
    g_ctx.prop = prop;
    g_ctx.device = device;
    g_ctx.ready = true;
    return GLX_OK;
}

int glx_synchronize(void) {
    GLX_REQUIRE_INIT();
    return check_hip(hipStreamSynchronize(g_ctx.stream), "hipStreamSynchronize");
}

int glx_device_info(int *num_cus, int *max_wg_per_cu_hint, char *name, int name_len) {
    GLX_REQUIRE_INIT();
    const hipDeviceProp_t &prop = g_ctx.prop;  // queried once by glx_init
    if (num_cus) *num_cus = prop.multiProcessorCount;
    if (max_wg_per_cu_hint) *max_wg_per_cu_hint = prop.maxThreadsPerMultiProcessor / 256;
    if (name && name_len > 0) {
        snprintf(name, (size_t)name_len, "%s %s", prop.name, prop.gcnArchName);
    }
    return GLX_OK;
}

int glx_malloc(void **ptr, size_t bytes) {
    GLX_REQUIRE_INIT();
    if (!ptr) return set_error(GLX_E_ARG, "glx_malloc: NULL out pointer");
    *ptr = nullptr;
    if (bytes == 0) bytes = 1;
    return check_hip(hipMalloc(ptr, bytes), "hipMalloc");
}

int glx_free(void *ptr) {
    GLX_REQUIRE_INIT();
    if (!ptr) return GLX_OK;
    hipError_t st = hipStreamSynchronize(g_ctx.stream);
    if (st != hipSuccess) return check_hip(st, "hipStreamSynchronize");
    return check_hip(hipFree(ptr), "hipFree");
}

static int copy(void *dst, const void *src, size_t bytes, hipMemcpyKind kind, const char *what) {
    GLX_REQUIRE_INIT();
    if (bytes == 0) return GLX_OK;
    if (!dst || !src) return set_error(GLX_E_ARG, "%s: NULL pointer", what);
    hipError_t st = hipMemcpyAsync(dst, src, bytes, kind, g_ctx.stream);
    if (st != hipSuccess) return check_hip(st, what);
    if (kind == hipMemcpyDeviceToDevice) return GLX_OK;
    return check_hip(hipStreamSynchronize(g_ctx.stream), what);
}

int glx_memcpy_h2d(void *dst, const void *src, size_t bytes) { return copy(dst, src, bytes, hipMemcpyHostToDevice, "glx_memcpy_h2d"); }
int glx_memcpy_d2h(void *dst, const void *src, size_t bytes) { return copy(dst, src, bytes, hipMemcpyDeviceToHost, "glx_memcpy_d2h"); }
int glx_memcpy_d2d(void *dst, const void *src, size_t bytes) { return copy(dst, src, bytes, hipMemcpyDeviceToDevice, "glx_memcpy_d2d"); }

int glx_memset(void *dst, int value, size_t bytes) {
    GLX_REQUIRE_INIT();
    if (bytes == 0) return GLX_OK;
    if (!dst) return set_error(GLX_E_ARG, "glx_memset: NULL pointer");
    return check_hip(hipMemsetAsync(dst, value, bytes, g_ctx.stream), "hipMemsetAsync");
}

int glx_timer_start(void) {
    GLX_REQUIRE_INIT();
    return check_hip(hipEventRecord(g_ctx.ev0, g_ctx.stream), "hipEventRecord");
}

int glx_timer_stop(float *ms) {
    GLX_REQUIRE_INIT();
    if (!ms) return set_error(GLX_E_ARG, "glx_timer_stop: NULL");
    hipError_t st = hipEventRecord(g_ctx.ev1, g_ctx.stream);
    if (st == hipSuccess) st = hipEventSynchronize(g_ctx.ev1);
    if (st == hipSuccess) st = hipEventElapsedTime(ms, g_ctx.ev0, g_ctx.ev1);
    return check_hip(st, "glx_timer_stop");
}

int glx_td_psi_sweep(int prec, glx_grid g, double dt, double eps, const void *eps_map, const uint8_t *mt,
                     const void *ab, void *psi_rhs, const void *psi, void *psi_next, double langevin_c,
                     uint32_t jstep, uint32_t rand_t, double stop_eps, int32_t *r2max) {
    GLX_DISPATCH(prec, psi_sweep(g_ctx.env, g, dt, eps, eps_map, mt, ab, psi_rhs, psi, psi_next, langevin_c, jstep,
                                 rand_t, stop_eps, r2max));
}

int glx_td_a_sweep(int prec, glx_grid g, double dt, double kappa2, double rho, double H, const uint8_t *mt,
                   const void *psi, const void *abi_ab_rhs, void *ab_rhs, const void *ab, void *ab_next,
                   double langevin_c, uint32_t jstep, uint32_t rand_t, double stop_eps, int32_t *r2max) {
    GLX_DISPATCH(prec, a_sweep(g_ctx.env, g, dt, kappa2, rho, H, mt, psi, abi_ab_rhs, ab_rhs, ab, ab_next,
                               langevin_c, jstep, rand_t, stop_eps, r2max));
}

int glx_td_phase_lock(int prec, void *psi, int32_t lock_N, const int32_t *lock_ns) {
    GLX_DISPATCH(prec, phase_lock(g_ctx.env, psi, lock_N, lock_ns));
}

int glx_td_run(int prec, glx_grid g, const glx_td_params *p, const void *eps_map, const uint8_t *mt,
               void *const psi_bufs[3], int32_t *psi_cur, void *const ab_bufs[3], int32_t *ab_cur, void *rhs_psi,
               void *rhs_ab, int32_t nsteps, int32_t *sweeps, int32_t wg, int32_t wg_size) {
    if (g_ctx.ready && g_ctx.env.td_profile && nsteps > 0 && nsteps <= SWEEPS_CAP) {
        // grow the event pool to two events per solve launch
        std::lock_guard<std::mutex> lk(g_mu);
        const size_t need = 4 * (size_t)nsteps;
        while (g_ctx.tev.size() < need) {
            hipEvent_t ev;
            hipError_t st = hipEventCreate(&ev);
            if (st != hipSuccess) return check_hip(st, "hipEventCreate");
            g_ctx.tev.push_back(ev);
        }
        g_ctx.env.tev = g_ctx.tev.data();
        g_ctx.env.tev_cap = (int)g_ctx.tev.size();
    }
    struct ClearAbort { ~ClearAbort() { g_ctx.env.abort_next = 0; } } clear_abort;  // one run only
    GLX_DISPATCH(prec, td_run(g_ctx.env, g, p, eps_map, mt, psi_bufs, psi_cur, ab_bufs, ab_cur, rhs_psi, rhs_ab,
                              nsteps, sweeps, wg, wg_size));
}

int glx_td_psi_solve(int prec, glx_grid g, const glx_td_params *p, const void *eps_map, const uint8_t *mt,
                     const void *ab, void *const psi_bufs[3], int32_t *psi_cur, void *rhs_psi, int32_t *sweeps) {
    if (!p || !ab || !sweeps) return set_error(GLX_E_ARG, "glx_td_psi_solve: NULL argument");
    glx_td_params q = *p;
    q.eqn = 1;
    q.langevin_A = 0.0;
    void *const Q[3] = {const_cast<void *>(ab), const_cast<void *>(ab), const_cast<void *>(ab)};
    int32_t qc = 0, sw[2] = {0, 0};
    const int rc = glx_td_run(prec, g, &q, eps_map, mt, psi_bufs, psi_cur, Q, &qc, rhs_psi, nullptr, 1, sw, 0, 256);
    if (rc == GLX_OK) *sweeps = sw[0];
    return rc;
}

int glx_td_a_solve(int prec, glx_grid g, const glx_td_params *p, const uint8_t *mt, const void *psi,
                   void *const ab_bufs[3], int32_t *ab_cur, void *rhs_ab, int32_t *sweeps) {
    if (!p || !psi || !sweeps) return set_error(GLX_E_ARG, "glx_td_a_solve: NULL argument");
    glx_td_params q = *p;
    q.eqn = 2;
    q.solveA = 1;
    q.langevin_psi = 0.0;
    void *const P[3] = {const_cast<void *>(psi), const_cast<void *>(psi), const_cast<void *>(psi)};
    int32_t pc = 0, sw[2] = {0, 0};
    const int rc = glx_td_run(prec, g, &q, nullptr, mt, P, &pc, ab_bufs, ab_cur, nullptr, rhs_ab, 1, sw, 0, 256);
    if (rc == GLX_OK) *sweeps = sw[1];
    return rc;
}

int glx_td_step(int prec, glx_grid g, const glx_td_params *p, const void *eps_map, const uint8_t *mt,
                void *const psi_bufs[3], int32_t *psi_cur, void *const ab_bufs[3], int32_t *ab_cur,
                void *rhs_psi, void *rhs_ab, int32_t *sweeps) {
    if (!p) return set_error(GLX_E_ARG, "glx_td_step: NULL parameters");
    glx_td_params q = *p;
    q.eqn = 0;
    return glx_td_run(prec, g, &q, eps_map, mt, psi_bufs, psi_cur, ab_bufs, ab_cur, rhs_psi, rhs_ab, 1, sweeps, 0, 256);
}

int glx_td_last_kernel_ms(float *ms) {
    GLX_REQUIRE_INIT();
    if (!ms) return set_error(GLX_E_ARG, "glx_td_last_kernel_ms: NULL");
    *ms = g_ctx.last_kernel_ms;
    return GLX_OK;
}

int glx_batch_td_run(int prec, glx_batch_instance *inst, int32_t n, int32_t nsteps, int32_t *sweeps,
                     int32_t use_graph, int32_t wg_per_cu, int32_t min_rows, const double *weight,
                     int32_t psi_depth) {
    struct ClearAbort { ~ClearAbort() { g_ctx.env.abort_next = 0; } } clear_abort;  // one run only
    const Workspace ws{nullptr, 0, "glx_batch_td_run"};
    GLX_DISPATCH(prec, batch_td_run(g_ctx.env, inst, n, nsteps, sweeps, use_graph, wg_per_cu, min_rows, weight,
                                    psi_depth, ws));
}

int glx_batch_td_run_noalloc(int prec, glx_batch_instance *inst, int32_t n, int32_t nsteps, int32_t *sweeps,
                             int32_t wg_per_cu, int32_t min_rows, const double *weight, int32_t psi_depth) {
    GLX_REQUIRE_INIT();
    struct ClearAbort { ~ClearAbort() { g_ctx.env.abort_next = 0; } } clear_abort;  // one run only
    const Workspace ws{g_scratch_dev[g_ctx.device], (size_t)GLX_NOALLOC_SCRATCH_BYTES, "glx_batch_td_run_noalloc"};
    if (!ws.ptr) return set_error(GLX_E_NOINIT, "glx_batch_td_run_noalloc: no scratch address for device %d", g_ctx.device);
    GLX_DISPATCH(prec, batch_td_run(g_ctx.env, inst, n, nsteps, sweeps, 0, wg_per_cu, min_rows, weight, psi_depth,
                                    ws));
}

int glx_batch_cg_run(int prec, glx_batch_cg_instance *inst, int32_t n, int32_t max_iters, double rtol,
                     int32_t iters_per_launch, int32_t stall_iters, int32_t fp64_finish) {
    GLX_DISPATCH(prec, batch_cg_run(g_ctx.env, inst, n, max_iters, rtol, iters_per_launch, stall_iters, fp64_finish));
}

int glx_td_force_barrier_abort(void) {
    GLX_REQUIRE_INIT();
    g_ctx.env.abort_next = 1;
    return GLX_OK;
}

int glx_td_profile(int32_t enable) {
    GLX_REQUIRE_INIT();
    g_ctx.env.td_profile = enable != 0;
    return GLX_OK;
}

int glx_td_solve_ms(double *psi_ms, int32_t *psi_launches, double *a_ms, int32_t *a_launches) {
    GLX_REQUIRE_INIT();
    if (!psi_ms || !psi_launches || !a_ms || !a_launches) return set_error(GLX_E_ARG, "glx_td_solve_ms: NULL");
    *psi_ms = g_ctx.solve_ms[0];
    *psi_launches = g_ctx.solve_n[0];
    *a_ms = g_ctx.solve_ms[1];
    *a_launches = g_ctx.solve_n[1];
    return GLX_OK;
}

int glx_td_passes(int32_t *psi_passes, int32_t *psi_redo, int32_t *a_passes, int32_t *a_redo, int32_t *a_xbuf) {
    GLX_REQUIRE_INIT();
    if (!psi_passes || !psi_redo || !a_passes || !a_redo || !a_xbuf) return set_error(GLX_E_ARG, "glx_td_passes: NULL");
    *psi_passes = g_ctx.passes[0];
    *psi_redo = g_ctx.passes[1];
    *a_passes = g_ctx.passes[2];
    *a_redo = g_ctx.passes[3];
    *a_xbuf = g_ctx.passes[4];
    return GLX_OK;
}

int glx_stream_probe(int32_t kind, int64_t n, int32_t grid, int32_t reps, double *gbs) {
    GLX_REQUIRE_INIT();
    if (kind < 0 || kind > 2 || n < 1024 || grid < 1 || reps < 1 || !gbs)
        return set_error(GLX_E_ARG, "glx_stream_probe: invalid arguments");
    // kind 0: read n float4; 1: copy n float4; 2: the psi-pass mix over n nodes
    const size_t f4 = (size_t)n * sizeof(float4);
    const size_t sizes[7] = {kind == 2 ? (size_t)n * 8 : f4, kind == 1 ? f4 : (kind == 2 ? (size_t)n * 8 : 64),
                             kind == 2 ? (size_t)n * 4 : 0, kind == 2 ? (size_t)n * 4 : 0,
                             kind == 2 ? (size_t)n * 4 : 0, kind == 2 ? (size_t)n : 0,
                             kind == 2 ? (size_t)n * 8 : 0};
    void *buf[7] = {};
    hipError_t st = hipSuccess;
    for (int k = 0; k < 7 && st == hipSuccess; ++k)
        if (sizes[k]) {
            st = hipMalloc(&buf[k], sizes[k]);
            if (st == hipSuccess) st = hipMemsetAsync(buf[k], 0, sizes[k], g_ctx.stream);
        }
    double bytes = 0.0;
    auto launch = [&]() {
        if (kind == 0) {
            hipLaunchKernelGGL(k_probe_read, dim3(grid), dim3(256), 0, g_ctx.stream, (const float4 *)buf[0], (float *)buf[1], (long)n);
            bytes = double(f4);
        } else if (kind == 1) {
            hipLaunchKernelGGL(k_probe_copy, dim3(grid), dim3(256), 0, g_ctx.stream, (const float4 *)buf[0], (float4 *)buf[1], (long)n);
            bytes = 2.0 * double(f4);
        } else {
            hipLaunchKernelGGL(k_probe_psi_mix, dim3(grid), dim3(256), 0, g_ctx.stream, (const float2 *)buf[0],
                               (const float2 *)buf[1], (const float *)buf[2], (const float *)buf[3], (const float *)buf[4],
                               (const unsigned char *)buf[5], (float2 *)buf[6], (long)n);
            bytes = 37.0 * double(n);
        }
    };
    float ms = 0.f;
    if (st == hipSuccess) {
        launch();  // warm-up
        st = hipEventRecord(g_ctx.ev0, g_ctx.stream);
        for (int k = 0; k < reps && st == hipSuccess; ++k) {
            launch();
            st = hipGetLastError();
        }
        if (st == hipSuccess) st = hipEventRecord(g_ctx.ev1, g_ctx.stream);
        if (st == hipSuccess) st = hipEventSynchronize(g_ctx.ev1);
        if (st == hipSuccess) st = hipEventElapsedTime(&ms, g_ctx.ev0, g_ctx.ev1);
    }
    for (int k = 0; k < 7; ++k)
        if (buf[k]) (void)hipFree(buf[k]);
    if (st != hipSuccess) return check_hip(st, "glx_stream_probe");
    *gbs = bytes * reps / (double(ms) * 1e-3) / 1e9;
    return GLX_OK;
}

int glx_td_max_grid(int prec, int32_t wg_size, int32_t *max_wg) {
    GLX_DISPATCH(prec, td_max_grid(g_ctx.env, wg_size, max_wg));
}

int glx_magnetic_field(int prec, glx_grid g, const void *abei, const void *ab, void *B) {
    GLX_DISPATCH(prec, magnetic_field(g_ctx.env, g, abei, ab, B));
}

int glx_current_density(int prec, glx_grid g, double kappa2, double H, const void *abei, const void *ab, void *jxjy) {
    GLX_DISPATCH(prec, current_density(g_ctx.env, g, kappa2, H, abei, ab, jxjy));
}

int glx_supercurrent_density(int prec, glx_grid g, const uint8_t *mt, const void *psi, const void *abei,
                             const void *ab, void *jsxjsy) {
    GLX_DISPATCH(prec, supercurrent_density(g_ctx.env, g, mt, psi, abei, ab, jsxjsy));
}

int glx_bulk_observables(int prec, glx_grid g, double kappa2, double H, const uint8_t *mt, const void *psi,
                         const void *abei, const void *ab, void *B, void *J, void *Js) {
    GLX_DISPATCH(prec, bulk_observables(g_ctx.env, g, kappa2, H, mt, psi, abei, ab, B, J, Js));
}

int glx_free_energy(int prec, glx_grid g, double kappa2, double eps, const void *eps_map, double H,
                    const uint8_t *mt, const void *psi, const void *abei, const void *ab, double *F) {
    GLX_DISPATCH(prec, free_energy(g_ctx.env, g, kappa2, eps, eps_map, H, mt, psi, abei, ab, F));
}

int glx_vortex_cells(int prec, glx_grid g, const void *psi, const void *ab, const void *abi, double H, double *rec,
                     int32_t capacity, int32_t *count) {
    GLX_DISPATCH(prec, vortex_cells(g_ctx.env, g, psi, ab, abi, H, rec, capacity, count));
}

int glx_cg_direction_step(int prec, glx_grid g, double kappa2, double eps, double H, const uint8_t *mt,
                          const void *psi, const void *abei, const void *ab, const void *dpsi, const void *jac, void *dpsi_out,
                          const void *dab, const void *jacA, void *dab_out, const double *beta_dev, double *c,
                          double *alpha_dev) {
    GLX_DISPATCH(prec, cg_direction_step(g_ctx.env, g, kappa2, eps, H, mt, psi, abei, ab, dpsi, jac, dpsi_out, dab,
                                         jacA, dab_out, beta_dev, c, alpha_dev));
}

int glx_cg_update_gradient_pass(int prec, glx_grid g, double kappa2, double eps, const void *eps_map, double H,
                                const uint8_t *mt, const void *psi, const void *abei, const void *ab,
                                const void *dpsi, const void *dab, const double *alpha_dev, void *psi_out,
                                void *ab_out, void *jac, const void *jac_prev, void *jacA, const void *jacA_prev,
                                double *beta_next_dev, double *F) {
    GLX_DISPATCH(prec, cg_update_gradient_pass(g_ctx.env, g, kappa2, eps, eps_map, H, mt, psi, abei, ab, dpsi, dab,
                                               alpha_dev, psi_out, ab_out, jac, jac_prev, jacA, jacA_prev,
                                               beta_next_dev, F));
}

int glx_cg_energy_wait(double *F) {
    GLX_REQUIRE_INIT();
    if (!F) return set_error(GLX_E_ARG, "glx_cg_energy_wait: NULL pointer");
    hipError_t st = hipEventSynchronize(g_ctx.env.ev_energy);
    if (st != hipSuccess) return check_hip(st, "hipEventSynchronize(energy)");
    *F = g_ctx.env.energy_host[0];
    return GLX_OK;
}

int glx_vortex_triangulate(glx_grid g, const double *rec, int32_t count, double *xyv) {
    GLX_REQUIRE_INIT();
    return L<double>::vortex_triangulate(g_ctx.env, g, rec, count, xyv);
}

int glx_sum(int prec, const void *x, int64_t n, int32_t block_size, double *out) {
    GLX_DISPATCH(prec, sum(g_ctx.env, x, n, block_size, out));
}

int glx_sum_v(int prec, const void *x, int64_t nv, int32_t ne, int32_t block_size, double *out) {
    GLX_DISPATCH(prec, sum_v(g_ctx.env, x, nv, ne, block_size, out));
}

int glx_pr_sums(int prec, const void *x, const void *y, int64_t n, double *out) {
    GLX_DISPATCH(prec, pr_sums(g_ctx.env, x, y, n, out));
}

int glx_cg_jacobian_psi(int prec, glx_grid g, double kappa2, double eps, const void *eps_map, double H,
                        const uint8_t *mt, const void *psi, const void *abei, const void *ab, void *G) {
    GLX_DISPATCH(prec, jacobian_psi(g_ctx.env, g, kappa2, eps, eps_map, H, mt, psi, abei, ab, G));
}

int glx_cg_jacobian_A(int prec, glx_grid g, double kappa2, double H, const uint8_t *mt, const void *psi,
                      const void *abei, const void *ab, void *G) {
    GLX_DISPATCH(prec, jacobian_A(g_ctx.env, g, kappa2, H, mt, psi, abei, ab, G));
}

int glx_cg_coef_psi(int prec, glx_grid g, double kappa2, double eps, double H, const uint8_t *mt, const void *psi,
                    const void *dpsi, const void *abei, const void *ab, double *c5) {
    GLX_DISPATCH(prec, coef_psi(g_ctx.env, g, kappa2, eps, H, mt, psi, dpsi, abei, ab, c5));
}

int glx_cg_coef(int prec, glx_grid g, double kappa2, double eps, double H, const uint8_t *mt, const void *psi,
                const void *dpsi, const void *abei, const void *ab, const void *dab, double *c17) {
    GLX_DISPATCH(prec, coef(g_ctx.env, g, kappa2, eps, H, mt, psi, dpsi, abei, ab, dab, c17));
}

int glx_cg_gradient_pass(int prec, glx_grid g, double kappa2, double eps, const void *eps_map, double H,
                         const uint8_t *mt, const void *psi, const void *abei, const void *ab, void *jac,
                         const void *jac_prev, void *jacA, const void *jacA_prev, double *beta_next_dev, double *F) {
    GLX_DISPATCH(prec, cg_gradient_pass(g_ctx.env, g, kappa2, eps, eps_map, H, mt, psi, abei, ab, jac, jac_prev, jacA,
                                        jacA_prev, beta_next_dev, F));
}

int glx_cg_pr_beta(int prec, const void *jac, const void *jac_prev, int64_t n, double *beta_dev) {
    GLX_DISPATCH(prec, pr_beta(g_ctx.env, jac, jac_prev, n, beta_dev));
}

int glx_axmy_dev(int prec, const void *x, const void *y, void *z, const double *alpha_dev, int64_t n) {
    GLX_DISPATCH(prec, axmy_dev(g_ctx.env, x, y, z, alpha_dev, n));
}

int glx_cg_line_step(int prec, glx_grid g, double kappa2, double eps, double H, const uint8_t *mt, void *psi,
                     const void *dpsi, const void *abei, void *ab, const void *dab, double *alpha_dev) {
    GLX_DISPATCH(prec, cg_line_step(g_ctx.env, g, kappa2, eps, H, mt, psi, dpsi, abei, ab, dab, alpha_dev));
}

int glx_axpy(int prec, const void *x, const void *y, void *z, double alpha, int64_t n) {
    GLX_DISPATCH(prec, axpy(g_ctx.env, x, y, z, alpha, n, 0));
}

int glx_axmy(int prec, const void *x, const void *y, void *z, double alpha, int64_t n) {
    GLX_DISPATCH(prec, axpy(g_ctx.env, x, y, z, alpha, n, 1));
}

}  // extern "C"
