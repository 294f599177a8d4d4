"""Benchmark: TDGL node-steps/s and CG iterations/s at 4097^2 nodes (Config 3),
FP32-mixed and FP64, with the roofline of the dominant kernel and a CPU
baseline.

Workload (Config 3): Lx = Ly = 2048, dx = 0.5 (4097^2 nodes), kappa = 5,
sigma = 200, H = 0.05, dt = 0.05, seeded random psi, a mask of random
circular holes (5% of the area) and a Gaussian random pinning field
eps(r) (std 0.2). A TD "step" is one TDGL time step: the psi Jacobi solve
and the A Jacobi solve, each swept to the reference's stop rule. A CG
iteration is one Polak-Ribiere step with its exact polynomial line search.

value = node-steps/s of the FP32-mixed TD (whole job, one GPU), timed over
K steps after W warm-up steps with the fields resident in HBM. The FP64 form
is timed the same way and reported beside it; CG iterations/s follow.

roofline: the dominant kernel is the persistent psi solve (k_td_psi_solve,
one launch per step). A second run of the same workload brackets every
solve launch with HIP events on the library stream (glx_td_profile);
achieved = algorithmic bytes of the psi launches / their summed duration.
Algorithmic bytes per node per psi sweep: 8s + e + m (s = 4 or 8, e = s
for the eps map, m = 1 for the mask) times the logged sweep counts (the
fused two-sweep passes move less; `traffic` is the measured HBM bytes per
launch from the committed rocprofv3 counter summary, when present).
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
PMC_SUMMARY = os.path.join(ROOT, "profiles", "td_psi_solve_4097_f32_pmc.json")


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def psi_bytes(sweeps, s, e, m, N):
    """Algorithmic HBM bytes of the psi solves of the given (S_psi, S_A) steps."""
    return sum(sp for sp, _ in sweeps) * (8 * s + e + m) * N


def a_bytes(sweeps, s, m, N):
    total = 0
    for _, sa in sweeps:
        for j in range(sa):
            total += (8 * s + m) if (j % 2 == 0 or j == 1) else (10 * s + m)
    return total * N


def run_td(gl, dt, warmup, steps, profile=False):
    """Warm up, then run `steps` TD steps in one glx_td_run call."""
    from svirl_amd import _native as nat
    gl.solve._init_td()
    td = gl.solve._td
    if warmup:
        td.steps_per_launch = max(1, min(warmup, 200))
        gl.solve.td(dt=dt, Nt=warmup)
    nat.call("glx_synchronize")
    n0 = len(td.sweeps)
    td.steps_per_launch = steps
    nat.call("glx_td_profile", 1 if profile else 0)
    t0 = time.perf_counter()
    gl.solve.td(dt=dt, Nt=steps)
    nat.call("glx_synchronize")
    wall = time.perf_counter() - t0
    nat.call("glx_td_profile", 0)
    kms = ctypes.c_float()
    nat.call("glx_td_last_kernel_ms", ctypes.byref(kms))
    pm, am = ctypes.c_double(), ctypes.c_double()
    pn, an = ctypes.c_int32(), ctypes.c_int32()
    nat.call("glx_td_solve_ms", ctypes.byref(pm), ctypes.byref(pn), ctypes.byref(am), ctypes.byref(an))
    return dict(sweeps=td.sweeps[n0:], wall=wall, kernel_s=kms.value / 1e3,
                psi_s=pm.value / 1e3, psi_n=pn.value, a_s=am.value / 1e3, a_n=an.value)


def run_cg(gl, iters):
    """`iters` CG iterations (the stop rule is disabled by a zero tolerance)."""
    from svirl_amd import _native as nat
    gl.solve._init_cg()
    cg = gl.solve._cg
    saved = cg._CG__convergence_rtol
    cg._CG__convergence_rtol = 0.0
    gl.solve.cg(n_iter=2)  # warm-up
    nat.call("glx_synchronize")
    t0 = time.perf_counter()
    gl.solve.cg(n_iter=iters)
    nat.call("glx_synchronize")
    t = time.perf_counter() - t0
    cg._CG__convergence_rtol = saved
    return dict(iters=len(cg.cg_energies), wall=t, F=float(cg.cg_energies[-1]))


def run_batch(args):
    """Config 4: n instances (FP32-mixed), batched steps vs the same
    instances run one after another (each through its own persistent
    launches), and the batched FP32 result against the batched FP64 form."""
    from svirl_amd import _native as nat
    from svirl_amd import workloads
    from svirl_amd.batch import CANDIDATES, BatchTD
    log(f"[batch] building {args.batch_n} Config 4 instances")
    sims = workloads.config4(n=args.batch_n, dtype=np.float32)
    nodes = sum(int(s.cfg.N) for s in sims)
    b = BatchTD(sims, use_graph=True, autotune=True)
    b.td(dt=0.1, Nt=b.tune_steps * (len(CANDIDATES) + 1) + 1)  # autotune on the first steps
    nat.call("glx_synchronize")
    t0 = time.perf_counter()
    b.td(dt=0.1, Nt=args.batch_steps)
    nat.call("glx_synchronize")
    tb = time.perf_counter() - t0
    rate_b = nodes * args.batch_steps / tb
    log(f"[batch] {len(sims)} instances, {nodes} nodes: {rate_b:.4g} node-steps/s batched "
        f"({1e3 * tb / args.batch_steps:.3f} ms/step, shape {b.shape})")
    for s in sims:
        s.solve.td(dt=0.1, Nt=1)  # warm each instance's own launches
    nat.call("glx_synchronize")
    t0 = time.perf_counter()
    for s in sims:
        s.solve.td(dt=0.1, Nt=args.batch_solo_steps)
    nat.call("glx_synchronize")
    ts = time.perf_counter() - t0
    rate_s = nodes * args.batch_solo_steps / ts
    log(f"[batch] one instance at a time: {rate_s:.4g} node-steps/s; speedup {rate_b / rate_s:.2f}x")
    sw = np.array([np.mean(np.array(s.solve._td.sweeps), axis=0) for s in sims])
    res = {"instances": len(sims), "nodes": nodes, "node_steps_per_s": rate_b, "ms_per_step": 1e3 * tb / args.batch_steps,
           "per_instance_node_steps_per_s": rate_s, "speedup_vs_per_instance": rate_b / rate_s,
           "tuned_shape": {"wg_per_cu": b.shape[0], "min_rows": b.shape[1]},
           "tuning_ms_per_step": {f"{k[0]},{k[1]}": v for k, v in b.timings.items()},
           "mean_sweeps_per_step": {"psi": float(sw[:, 0].mean()), "A": float(sw[:, 1].mean())},
           "dtype": "f32", "graph": True}
    del sims
    # precision: the first 32 instances, FP32-mixed vs FP64, batched, same steps
    k = min(32, args.batch_n)
    s32 = workloads.config4(n=args.batch_n, dtype=np.float32)[:k]
    s64 = workloads.config4(n=args.batch_n, dtype=np.float64)[:k]
    BatchTD(s32, autotune=False).td(dt=0.1, Nt=args.batch_steps)
    BatchTD(s64, autotune=False).td(dt=0.1, Nt=args.batch_steps)
    fe, pe, je = [], [], []
    for a, c in zip(s32, s64):
        fa, fc = a.observables.free_energy_fp64(), c.observables.free_energy_fp64()
        fe.append(abs(fa - fc) / abs(fc))
        pa = np.asarray(a.vars.order_parameter, np.complex128)
        pc = np.asarray(c.vars.order_parameter, np.complex128)
        pe.append(np.linalg.norm(pa - pc) / np.linalg.norm(pc))
        ja = np.concatenate([np.ravel(x) for x in a.observables.current_density]).astype(np.float64)
        jc = np.concatenate([np.ravel(x) for x in c.observables.current_density]).astype(np.float64)
        je.append(np.linalg.norm(ja - jc) / max(np.linalg.norm(jc), 1e-300))
    res["fp32_vs_fp64"] = {"instances": k, "after_steps": args.batch_steps, "F_rel_max": float(np.max(fe)),
                           "psi_L2_rel_max": float(np.max(pe)), "J_L2_rel_max": float(np.max(je))}
    log(f"[batch] FP32 vs FP64 over {k} instances: F {np.max(fe):.2e}, psi {np.max(pe):.2e}, J {np.max(je):.2e}")
    return res


def run_batch_cg(args):
    """Config 5: n instances of 257^2 (kappa = 5) minimised by the batched
    CG, FP32-mixed then FP64, up to --cg-batch-iters iterations with the
    ragged stop ||grad F|| <= 1e-8 ||grad F||_0."""
    from svirl_amd import workloads
    from svirl_amd.batch import BatchCG
    res = {"instances": args.cg_batch_n, "max_iters": args.cg_batch_iters, "rtol": 1e-8}
    F = {}
    for name, dtype in (("fp32", np.float32), ("fp64", np.float64)):
        log(f"[batch-cg] building {args.cg_batch_n} Config 5 instances ({name})")
        sims = workloads.config5(n=args.cg_batch_n, dtype=dtype)
        nodes = sum(int(s.cfg.N) for s in sims)
        b = BatchCG(sims)
        b.cg(max_iter=2, rtol=0.0)  # warm-up (two iterations from the random state)
        t0 = time.perf_counter()
        iters, conv, grel = b.cg(max_iter=args.cg_batch_iters, rtol=1e-8)
        t = time.perf_counter() - t0
        total = int(iters.sum())
        res[name] = {"instance_iters_per_s": total / t, "node_iters_per_s": float(np.sum(iters * (nodes / len(sims)))) / t,
                     "wall_s": t, "converged": int(conv.sum()), "iters_min": int(iters.min()),
                     "iters_max": int(iters.max()), "grad_rel_median": float(np.median(grel))}
        F[name] = np.array([s.observables.free_energy_fp64() for s in sims[:64]])
        log(f"[batch-cg] {name}: {total / t:.4g} instance-iterations/s, {res[name]['node_iters_per_s']:.4g} "
            f"node-iterations/s, {int(conv.sum())} converged, median grad_rel {np.median(grel):.2e}")
        del sims, b
    res["fp32_vs_fp64_F_rel_max"] = float(np.max(np.abs(F["fp32"] - F["fp64"]) / np.abs(F["fp64"])))
    res["fp32_vs_fp64_F_rel_median"] = float(np.median(np.abs(F["fp32"] - F["fp64"]) / np.abs(F["fp64"])))
    log(f"[batch-cg] FP32 vs FP64 free energy (first 64 instances): max rel {res['fp32_vs_fp64_F_rel_max']:.2e}")
    return res


def calibrate_copy(n=1 << 26, reps=3):
    """FP32 axpy over n reals (reads 8n bytes, writes 4n), for scaling the
    rocprofv3 FETCH_SIZE / WRITE_SIZE counters."""
    from svirl_amd import _native as nat
    x = nat.DeviceArray.zeros(n, np.float32)
    y = nat.DeviceArray.zeros(n, np.float32)
    for _ in range(reps):
        nat.call("glx_axpy", nat.GLX_F32, x.ptr, y.ptr, y.ptr, 0.5, n)
    nat.call("glx_synchronize")
    x.free()
    y.free()


def cpu_baseline(L, dt, steps):
    """The FP64 oracle on a smaller instance of the same configuration."""
    os.environ.setdefault("OMP_NUM_THREADS", "1")
    from oracle import gl_oracle as O
    from svirl_amd import workloads
    dx = 0.5
    Nx = int(round(L / dx)) + 1
    g = O.Geom(Nx, Nx, dx, dx)
    mt = workloads.holes_mask(g.Nxc, g.Nyc, dx, dx)
    em = workloads.pinning_field(Nx, Nx, dx)
    psi = O.random_psi(g.N, 1.0, seed=3)
    mt_flat = np.ascontiguousarray(mt.T).ravel()
    psi = np.where(O.mask_at_nodes(g, mt_flat), psi, 0)
    st = O.OracleState(g, psi, O.symmetric_gauge(g, 0.05), kappa=5.0, sigma=200.0, H=0.05,
                       eps_map=np.ascontiguousarray(em.T).ravel(), mt=mt_flat)
    st.td(dt=dt, Nt=1)  # warm-up step
    t0 = time.perf_counter()
    st.td(dt=dt, Nt=steps)
    t = time.perf_counter() - t0
    return g.N * steps / t, t, g.N, st.sweeps[1:]


def _cpu_batch_task(spec):
    """One oracle task of the batched baselines (runs in a worker or
    inline): ('td', Nx, Ny, H, kappa, sigma, seed, n) -> n TD steps of a
    Config 4 instance; ('cg', ...) -> n CG iterations of a Config 5
    instance. Returns (nodes * n, seconds)."""
    from oracle import gl_oracle as O
    from svirl_amd import workloads
    kind, Nx, Ny, H, kappa, sigma, seed, n = spec
    g = O.Geom(Nx, Ny, 0.5, 0.5)
    em = workloads.pinning_field(Nx, Ny, 0.5, seed=seed)
    st = O.OracleState(g, O.random_psi(g.N, 1.0, seed=seed), O.symmetric_gauge(g, H), kappa=kappa, sigma=sigma, H=H,
                       eps_map=np.ascontiguousarray(em.T).ravel())
    t0 = time.perf_counter()
    if kind == 'td':
        st.td(dt=0.1, Nt=n)
    else:
        st.cg(n_iter=n)
    return g.N * n, time.perf_counter() - t0


def cpu_batch_baselines(workers, n_td=16, n_cg=8, cg_iters=3):
    """The FP64 oracle on samples of Config 4 (one TD step of each of its
    first n_td instances, drawn as workloads.config4 draws them) and Config
    5 (cg_iters CG iterations of n_cg instances), on one core and on a pool
    of `workers` processes. Runs before the GPU is touched, so the pool is
    forked from a process without GPU state."""
    import multiprocessing as mp
    os.environ.setdefault("OMP_NUM_THREADS", "1")
    rng = np.random.default_rng(40)  # workloads.config4's draws
    td = []
    for k in range(n_td):
        Lx, Ly = float(rng.choice((32, 64, 128, 256, 512))), float(rng.choice((32, 64, 128, 256, 512)))
        H = float(rng.uniform(0.02, 0.5))
        kappa = float(rng.choice([2.0, 5.0, np.inf]))
        td.append(('td', int(round(Lx / 0.5)) + 1, int(round(Ly / 0.5)) + 1, H, kappa, 200.0, 40 * 1000 + k, 1))
    rng = np.random.default_rng(50)  # workloads.config5's draws
    cg = [('cg', 257, 257, float(rng.uniform(0.05, 0.3)), 5.0, 1.0, 50 * 10000 + k, cg_iters) for k in range(n_cg)]
    out = {}
    with mp.get_context("fork").Pool(workers) as pool:
        for name, specs, unit in (("config4", td, "node-steps/s"), ("config5", cg, "node-iterations/s")):
            t0 = time.perf_counter()
            r1 = [_cpu_batch_task(s) for s in specs]
            t1 = time.perf_counter() - t0
            t0 = time.perf_counter()
            rp = pool.map(_cpu_batch_task, specs, chunksize=1)
            tp = time.perf_counter() - t0
            work = sum(w for w, _ in r1)
            sample = (f"FP64 NumPy oracle, one TD step of each of the first {len(specs)} Config 4 instances "
                      f"({work} nodes)" if name == "config4" else
                      f"FP64 NumPy oracle, {cg_iters} CG iterations (reference line search) of {len(specs)} "
                      f"Config 5 instances of 257^2")
            out[name] = {"value": work / t1, "unit": unit, "cores": 1, "kind": "port",
                         "pool_value": sum(w for w, _ in rp) / tp, "pool_cores": workers, "sample": sample}
            log(f"[cpu] {name}: {work / t1:.4g} {unit} on 1 core, {sum(w for w, _ in rp) / tp:.4g} on {workers} "
                f"processes ({sample})")
    return out


def pmc_traffic():
    """Measured HBM bytes per node per psi sweep of k_td_psi_solve<float>
    from the committed counter summary (FETCH_SIZE + WRITE_SIZE, scaled by
    the copy calibration), or None."""
    try:
        with open(PMC_SUMMARY) as f:
            return json.load(f).get("bytes_per_node_sweep")
    except (OSError, ValueError):
        return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--L", type=float, default=2048.0, help="box size (2048 = 4097^2 nodes)")
    ap.add_argument("--profile-steps", type=int, default=100)
    ap.add_argument("--cg-iters", type=int, default=20)
    ap.add_argument("--skip-fp64", action="store_true")
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--skip-cg", action="store_true")
    ap.add_argument("--skip-batch", action="store_true")
    ap.add_argument("--batch-n", type=int, default=256)
    ap.add_argument("--batch-steps", type=int, default=60)
    ap.add_argument("--batch-solo-steps", type=int, default=10)
    ap.add_argument("--skip-batch-cg", action="store_true")
    ap.add_argument("--cg-batch-n", type=int, default=1024)
    ap.add_argument("--cg-batch-iters", type=int, default=300)
    ap.add_argument("--calibrate", action="store_true",
                    help="run FP32 axpy dispatches of known size first (counter calibration)")
    ap.add_argument("--cpu-L", type=float, default=512.0)
    ap.add_argument("--cpu-steps", type=int, default=3)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (each instance runs on a "
              "single GPU; a multi-GPU run would split the 4097^2 grid into row slabs, one per GPU, or "
              "the batched instances among GPUs).", file=sys.stderr)
        sys.exit(2)

    cpu_batch = None
    if not args.skip_cpu:
        # before any GPU call: the pool's workers are forked from this process
        cpu_batch = cpu_batch_baselines(max(1, min(16, os.cpu_count() or 1)))

    from svirl_amd import _native as nat
    from svirl_amd import workloads
    nat.init(0)
    info = nat.device_info()
    log(f"device: {info}")
    if args.calibrate:
        calibrate_copy()
        log("[calib] 3 FP32 axpy dispatches over 2^26 reals")
    dt = 0.05
    res = {}
    for prec in (["fp32"] if args.skip_fp64 else ["fp32", "fp64"]):
        dtype = np.float32 if prec == "fp32" else np.float64
        s = 4 if prec == "fp32" else 8
        log(f"[{prec}] building Config 3 state, L={args.L}")
        gl = workloads.config3(L=args.L, dtype=dtype)
        N = int(gl.cfg.N)
        log(f"[{prec}] N = {N}; warmup {args.warmup} steps, timing {args.steps}")
        r = run_td(gl, dt, args.warmup, args.steps)
        sp = np.array(r["sweeps"])
        out = dict(N=N, wall=r["wall"], kernel_s=r["kernel_s"], sweeps_psi=float(sp[:, 0].mean()),
                   sweeps_A=float(sp[:, 1].mean()), node_steps_per_s=N * args.steps / r["wall"])
        out["all_gbs"] = (psi_bytes(r["sweeps"], s, s, 1, N) + a_bytes(r["sweeps"], s, 1, N)) / r["kernel_s"] / 1e9
        log(f"[{prec}] {out['node_steps_per_s']:.4g} node-steps/s ({1e3 * r['wall'] / args.steps:.3f} ms/step wall, "
            f"{1e3 * r['kernel_s'] / args.steps:.3f} ms/step on the GPU), "
            f"sweeps/step psi {out['sweeps_psi']:.2f} A {out['sweeps_A']:.2f}")
        if args.steps <= 10:
            log(f"[{prec}] sweeps per step (psi, A): {r['sweeps']}")
        if args.profile_steps:
            p = run_td(gl, dt, 0, args.profile_steps, profile=True)
            pb = psi_bytes(p["sweeps"], s, s, 1, N)
            ab = a_bytes(p["sweeps"], s, 1, N)
            out["psi_launch_us"] = 1e6 * p["psi_s"] / max(1, p["psi_n"])
            out["a_launch_us"] = 1e6 * p["a_s"] / max(1, p["a_n"])
            out["psi_gbs"] = pb / p["psi_s"] / 1e9
            out["a_gbs"] = ab / p["a_s"] / 1e9 if p["a_s"] > 0 else None
            out["psi_bytes_per_launch"] = pb / max(1, p["psi_n"])
            log(f"[{prec}] psi solve {out['psi_launch_us']:.1f} us/launch, {out['psi_gbs']:.1f} GB/s algorithmic; "
                f"A solve {out['a_launch_us']:.1f} us/launch")
        out["F_td"] = gl.observables.free_energy_fp64()
        # vortex detection on the final state (winding + triangulation on the
        # GPU, ordering on the host); the second call is timed
        for _ in range(2):
            gl.solve.vortices_detected = False
            t0 = time.perf_counter()
            vx = gl.vortex_detector.vortices[0]
            out["vortex_ms"] = 1e3 * (time.perf_counter() - t0)
        out["vortices"] = int(len(vx))
        out["vortex_cells"] = int(len(gl.vortex_detector.vortex_cells[0]))
        log(f"[{prec}] vortex detection {out['vortex_ms']:.1f} ms: {out['vortex_cells']} cells flagged, "
            f"{out['vortices']} vortices")
        if not args.skip_fp64:
            jx, jy = gl.observables.current_density
            out["psi_h"] = np.asarray(gl.vars.order_parameter, dtype=np.complex128).ravel()
            out["J_h"] = np.concatenate([np.asarray(jx, dtype=np.float64).ravel(), np.asarray(jy, dtype=np.float64).ravel()])
        if not args.skip_cg:
            c = run_cg(gl, args.cg_iters)
            out["cg_iters_per_s"] = c["iters"] / c["wall"]
            out["cg_ms_per_iter"] = 1e3 * c["wall"] / c["iters"]
            out["F_cg"] = c["F"]
            log(f"[{prec}] CG {out['cg_iters_per_s']:.2f} iters/s ({out['cg_ms_per_iter']:.2f} ms/iter)")
        res[prec] = out
        del gl

    batch = None
    if not args.skip_batch:
        batch = run_batch(args)
    batch_cg = None
    if not args.skip_batch_cg:
        batch_cg = run_batch_cg(args)

    cpu = None
    if not args.skip_cpu:
        log(f"[cpu] oracle on a {int(round(args.cpu_L / 0.5)) + 1}^2 instance of the same config, {args.cpu_steps} steps")
        rate, t, Nc, csw = cpu_baseline(args.cpu_L, dt, args.cpu_steps)
        cpu = {"value": rate, "unit": "node-steps/s", "cores": 1, "kind": "port",
               "sample": f"FP64 NumPy oracle, Config 3 physics on {int(round(args.cpu_L / 0.5)) + 1}^2 nodes "
                         f"(L={args.cpu_L}), {args.cpu_steps} TD steps after 1 warm-up step, {t:.1f} s, "
                         f"sweeps/step {np.mean(csw, axis=0).round(2).tolist()}"}
        log(f"[cpu] {rate:.4g} node-steps/s")

    r32 = res["fp32"]
    per_node_sweep = pmc_traffic()
    traffic = None
    if per_node_sweep and r32.get("psi_bytes_per_launch"):
        # measured bytes per node-sweep x this run's node-sweeps per launch
        traffic = per_node_sweep * r32["psi_bytes_per_launch"] / 37.0
    roof = {"bound": "hbm", "achieved": r32.get("psi_gbs"), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": (r32["psi_gbs"] / HBM_PEAK_GBS) if r32.get("psi_gbs") else None,
            "traffic": traffic,
            "traffic_bytes_per_node_sweep": per_node_sweep,
            "algorithmic_bytes_per_node_sweep": 37,
            "physical_GBs": (traffic / (r32["psi_launch_us"] * 1e-6) / 1e9) if traffic else None,
            "kernel": "k_td_psi_solve<float, gauge-shifted A>",
            "launch_us": r32.get("psi_launch_us"),
            "algorithmic_bytes_per_launch": r32.get("psi_bytes_per_launch"),
            "traffic_source": os.path.relpath(PMC_SUMMARY, ROOT) if traffic else None}
    out = {
        "metric": "TDGL node-steps/s at 4097^2 (FP32-mixed; FP64 and CG iters/s beside it)",
        "value": r32["node_steps_per_s"],
        "unit": "node-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * r32["wall"] / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded Config 3 inputs)",
        "config": {"workload": "Config 3: 4097^2 nodes, kappa=5, sigma=200, H=0.05, dt=0.05, 5% circular holes, "
                               "Gaussian pinning eps std 0.2", "nodes": r32["N"], "L": args.L},
        "roofline": roof,
        "sweeps_per_step": {"psi": r32["sweeps_psi"], "A": r32["sweeps_A"]},
        "a_solve": {"launch_us": r32.get("a_launch_us"), "achieved_GBs": r32.get("a_gbs")},
        "td_all_kernels_GBs": r32["all_gbs"],
        "vortex_detector": {"fp32_ms": r32["vortex_ms"], "fp32_vortices": r32["vortices"],
                            "fp32_cells_flagged": r32["vortex_cells"]},
        "cpu_baseline": cpu,
    }
    if cpu_batch:
        out["cpu_baseline_batched"] = cpu_batch
    if batch:
        out["batch"] = batch
    if batch_cg:
        out["batch_cg"] = batch_cg
    if "cg_iters_per_s" in r32:
        out["cg"] = {"fp32_iters_per_s": r32["cg_iters_per_s"], "fp32_ms_per_iter": r32["cg_ms_per_iter"]}
    if "fp64" in res:
        r64 = res["fp64"]
        out["fp64"] = {"node_steps_per_s": r64["node_steps_per_s"], "ms_per_step": 1e3 * r64["wall"] / args.steps,
                       "psi_solve_GBs": r64.get("psi_gbs"),
                       "psi_solve_frac": (r64["psi_gbs"] / HBM_PEAK_GBS) if r64.get("psi_gbs") else None,
                       "psi_launch_us": r64.get("psi_launch_us"),
                       "sweeps_per_step": {"psi": r64["sweeps_psi"], "A": r64["sweeps_A"]}, "F": r64["F_td"]}
        out["fp32_over_fp64"] = r32["node_steps_per_s"] / r64["node_steps_per_s"]
        out["vortex_detector"].update({"fp64_ms": r64["vortex_ms"], "fp64_vortices": r64["vortices"],
                                       "fp64_cells_flagged": r64["vortex_cells"]})
        out["fp32_F_rel_err_vs_fp64"] = abs(r32["F_td"] - r64["F_td"]) / abs(r64["F_td"])
        out["fp32_vs_fp64"] = {
            "after_steps": args.warmup + args.steps + args.profile_steps,
            "F_rel": abs(r32["F_td"] - r64["F_td"]) / abs(r64["F_td"]),
            "psi_L2_rel": float(np.linalg.norm(r32["psi_h"] - r64["psi_h"]) / np.linalg.norm(r64["psi_h"])),
            "J_L2_rel": float(np.linalg.norm(r32["J_h"] - r64["J_h"]) / np.linalg.norm(r64["J_h"])),
            "reference": "FP64 GPU form on the same inputs (validated against the FP64 oracle by the tests)"}
        if "cg_iters_per_s" in r64:
            out["cg"]["fp64_iters_per_s"] = r64["cg_iters_per_s"]
            out["cg"]["fp64_ms_per_iter"] = r64["cg_ms_per_iter"]
            out["cg"]["fp32_F_rel_err_vs_fp64"] = abs(r32["F_cg"] - r64["F_cg"]) / abs(r64["F_cg"])
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
