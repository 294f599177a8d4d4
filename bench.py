"""Benchmark: TDGL node-steps/s at 4097^2 nodes (Config 3), FP32-mixed and
FP64, with the roofline of the persistent TD kernel and a CPU baseline.

Workload (Config 3): Lx = Ly = 2048, dx = 0.5 (4097^2 nodes), kappa = 5,
sigma = 200, H = 0.05, dt = 0.05, seeded random psi, a mask of random
circular holes (5% of the area) and a Gaussian random pinning field
eps(r) (std 0.2). A "step" is one TDGL time step: the psi Jacobi solve and
the A Jacobi solve, each swept to the reference's stop rule.

value = node-steps/s of the FP32-mixed form (whole job, one GPU), timed
over K steps after W warm-up steps with the fields resident in HBM. The FP64
form is timed the same way and reported beside it. roofline.achieved =
algorithmic HBM bytes of the timed launch / its kernel duration (HIP events
on the library stream); bytes per node per sweep: psi 8s + e + m, A 8s + m
(even sweeps and sweep 1) or 10s + m (odd sweeps >= 3), s = 4 or 8,
e = s (eps map), m = 1 (mask).
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


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def bytes_per_step(sweeps, s, e, m, N):
    """Algorithmic HBM bytes of the given (S_psi, S_A) steps."""
    total = 0
    for sp, sa in sweeps:
        total += sp * (8 * s + e + m)
        for j in range(sa):
            total += (8 * s + m) if (j % 2 == 0 or j == 1) else (10 * s + m)
    return total * N


def run_td(gl, dt, warmup, steps):
    """Warm up, then time `steps` TD steps in one persistent launch."""
    from svirl_amd import _native as nat
    gl.solve._init_td()
    td = gl.solve._td
    if warmup:
        td.steps_per_launch = max(1, min(warmup, 200))
        gl.solve.td(dt=dt, Nt=warmup)
    nat.call("glx_synchronize")
    n0 = len(td.sweeps)
    td.steps_per_launch = steps
    t0 = time.perf_counter()
    gl.solve.td(dt=dt, Nt=steps)
    nat.call("glx_synchronize")
    wall = time.perf_counter() - t0
    kms = ctypes.c_float()
    nat.call("glx_td_last_kernel_ms", ctypes.byref(kms))
    return td.sweeps[n0:], wall, kms.value / 1e3


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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--L", type=float, default=2048.0, help="box size (2048 = 4097^2 nodes)")
    ap.add_argument("--skip-fp64", action="store_true")
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--cpu-L", type=float, default=512.0)
    ap.add_argument("--cpu-steps", type=int, default=3)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (each instance runs on a "
              "single GPU; see DESIGN.md for how the path would shard).", file=sys.stderr)
        sys.exit(2)

    from svirl_amd import _native as nat
    from svirl_amd import workloads
    nat.init(0)
    info = nat.device_info()
    log(f"device: {info}")
    dt = 0.05
    res = {}
    for prec in (["fp32"] if args.skip_fp64 else ["fp32", "fp64"]):
        dtype = np.float32 if prec == "fp32" else np.float64
        log(f"[{prec}] building Config 3 state, L={args.L}")
        gl = workloads.config3(L=args.L, dtype=dtype)
        N = int(gl.cfg.N)
        log(f"[{prec}] N = {N}; warmup {args.warmup} steps, timing {args.steps}")
        sweeps, wall, ksec = run_td(gl, dt, args.warmup, args.steps)
        s = 4 if prec == "fp32" else 8
        B = bytes_per_step(sweeps, s, s, 1, N)
        sp = np.array(sweeps)
        res[prec] = dict(N=N, wall=wall, kernel_s=ksec, bytes=B, sweeps_psi=float(sp[:, 0].mean()),
                         sweeps_A=float(sp[:, 1].mean()), node_steps_per_s=N * args.steps / wall,
                         gbs=B / ksec / 1e9, F=gl.observables.free_energy_fp64())
        log(f"[{prec}] {res[prec]['node_steps_per_s']:.4g} node-steps/s, {res[prec]['gbs']:.1f} GB/s, "
            f"sweeps/step psi {res[prec]['sweeps_psi']:.2f} A {res[prec]['sweeps_A']:.2f}")
        del gl

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
    out = {
        "metric": "TDGL node-steps/s at 4097^2 (FP32-mixed; FP64 beside it)",
        "value": r32["node_steps_per_s"],
        "unit": "node-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * r32["wall"] / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 fields, f64 reductions/residuals",
        "data": "synthetic (seeded Config 3 inputs)",
        "config": {"workload": "Config 3: 4097^2 nodes, kappa=5, sigma=200, H=0.05, dt=0.05, 5% circular holes, "
                               "Gaussian pinning eps std 0.2", "nodes": r32["N"], "L": args.L},
        "roofline": {"bound": "hbm", "achieved": r32["gbs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": r32["gbs"] / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "k_td_run<float>", "kernel_s": r32["kernel_s"], "algorithmic_bytes": r32["bytes"]},
        "sweeps_per_step": {"psi": r32["sweeps_psi"], "A": r32["sweeps_A"]},
        "cpu_baseline": cpu,
    }
    if "fp64" in res:
        r64 = res["fp64"]
        out["fp64"] = {"node_steps_per_s": r64["node_steps_per_s"], "ms_per_step": 1e3 * r64["wall"] / args.steps,
                       "achieved_GBs": r64["gbs"], "frac": r64["gbs"] / HBM_PEAK_GBS,
                       "sweeps_per_step": {"psi": r64["sweeps_psi"], "A": r64["sweeps_A"]}, "F": r64["F"]}
        out["fp32_over_fp64"] = r32["node_steps_per_s"] / r64["node_steps_per_s"]
        out["fp32_F_rel_err_vs_fp64"] = abs(r32["F"] - r64["F"]) / abs(r64["F"])
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
