"""Benchmark: TDGL node-steps/s and CG iterations/s at 4097^2 nodes (Config 3),
FP32-mixed and FP64, with the roofline of the dominant kernel and a CPU
baseline; Configs 1, 2, 4 and 5 beside it.

Workload of `value` (Config 3): Lx = Ly = 2048, dx = 0.5 (4097^2 nodes),
kappa = 5, sigma = 200, H = 0.05, dt = 0.05, seeded random psi, a mask of
random circular holes (5% of the area) and a Gaussian random pinning field
eps(r) (std 0.2). A TD "step" is one TDGL time step: the psi Jacobi solve and
the A Jacobi solve, each swept to the reference's stop rule. A CG iteration
is one Polak-Ribiere step with its exact polynomial line search.

value = node-steps/s of the FP32-mixed TD (whole job, one GPU), timed over
K steps after W warm-up steps with the fields resident in HBM. The FP64 form
is timed the same way and reported beside it; CG iterations/s follow.

roofline: the dominant kernel is the persistent psi solve (k_td_psi_solve,
one launch per step). A second run of the same workload brackets every
solve launch with HIP events on the library stream (glx_td_profile) and
reads the memory passes the solves made (glx_td_passes). achieved =
compulsory bytes of the psi launches / their summed duration, where a pass
over the fields (one, two or three fused Jacobi sweeps) must move
(8s + e + m) bytes per node: psi in, right-hand side, a, b, eps, mask, psi
out (s = 4 or 8, e = s with the eps map, m = 1 with the mask). traffic =
the HBM bytes per launch measured by rocprofv3 counters (FETCH_SIZE +
WRITE_SIZE, calibrated by a copy kernel; committed summary under
profiles/), per node-pass, times this run's node-passes per launch.

Precision: every configuration's FP32-mixed result (and the FP64 one) is
compared with the FP64 oracle's stored results (tests/golden/*.npz, written
by tests/golden/make_golden.py) at the horizon the oracle reaches, and
beyond it with the FP64 GPU form.
"""
import argparse
import ctypes
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
PMC_SUMMARY = os.path.join(ROOT, "profiles", "td_psi_solve_4097_f32_pmc.json")
GOLD = os.path.join(ROOT, "tests", "golden")


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def gold(name):
    with np.load(os.path.join(GOLD, name + ".npz")) as z:
        return {k: z[k] for k in z.files}


def rel(a, b):
    return float(abs(a - b) / abs(b))


def l2rel(a, b):
    return float(np.linalg.norm(np.ravel(a - b)) / max(np.linalg.norm(np.ravel(b)), 1e-300))


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


# ---------------------------------------------------------------------------
# runs on the GPU
# ---------------------------------------------------------------------------

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
    ps = [ctypes.c_int32() for _ in range(5)]
    nat.call("glx_td_passes", *[ctypes.byref(x) for x in ps])
    return dict(sweeps=td.sweeps[n0:], wall=wall, kernel_s=kms.value / 1e3, psi_s=pm.value / 1e3, psi_n=pn.value,
                a_s=am.value / 1e3, a_n=an.value, passes=[x.value for x in ps])


def run_cg(gl, iters, stop_rule=False):
    """CG iterations/s: a 2-iteration warm-up, then `iters` iterations
    (with stop_rule False the energy stop rule is disabled by a zero
    tolerance; True keeps the reference's rule, which may end it earlier)."""
    from svirl_amd import _native as nat
    gl.solve._init_cg()
    cg = gl.solve._cg
    saved = cg._CG__convergence_rtol
    if not stop_rule:
        cg._CG__convergence_rtol = 0.0
        gl.solve.cg(n_iter=2)  # warm-up
    nat.call("glx_synchronize")
    t0 = time.perf_counter()
    gl.solve.cg(n_iter=iters)
    nat.call("glx_synchronize")
    t = time.perf_counter() - t0
    cg._CG__convergence_rtol = saved
    return dict(iters=len(cg.cg_energies), wall=t, F=float(cg.cg_energies[-1]), energies=list(cg.cg_energies))


def td_pass_bytes(s, e, m, N, r):
    """Compulsory bytes of the logged memory passes: (8s + e + m) per node
    per psi pass; (8s + m) per node per A pass plus 2s per node for each A
    pass that moves one more A buffer."""
    psi_passes, _, a_passes, _, a_xbuf = r["passes"]
    return psi_passes * (8 * s + e + m) * N, (a_passes * (8 * s + m) + a_xbuf * 2 * s) * N


def rows_state(gl, rows, rows_b):
    psi = np.asarray(gl.vars.order_parameter, np.complex128).T
    a, b = gl.vars.vector_potential
    jx, jy = gl.observables.current_density
    return dict(psi_rows=psi[rows], a_rows=np.asarray(a, np.float64).T[rows], b_rows=np.asarray(b, np.float64).T[rows_b],
                Ja_rows=np.asarray(jx, np.float64).T[rows], Jb_rows=np.asarray(jy, np.float64).T[rows_b])


def rows_errors(gl, z, sfx):
    """F, psi, A and J errors of a GPU state against the oracle's sampled
    rows (relative L2 over the rows)."""
    g = rows_state(gl, z["rows" + sfx], z["rows_b" + sfx])
    out = {"psi_L2_rel": l2rel(g["psi_rows"], z["psi_rows" + sfx]),
           "A_L2_rel": l2rel(np.concatenate([g["a_rows"].ravel(), g["b_rows"].ravel()]),
                             np.concatenate([z["a_rows" + sfx].ravel(), z["b_rows" + sfx].ravel()])),
           "J_L2_rel": l2rel(np.concatenate([g["Ja_rows"].ravel(), g["Jb_rows"].ravel()]),
                             np.concatenate([z["Ja_rows" + sfx].ravel(), z["Jb_rows" + sfx].ravel()])),
           "rows": int(len(z["rows" + sfx]))}
    return out


def bench_c3(args, prec):
    """Config 3 in one precision: timed TD, profiled TD (roofline), vortex
    detection, CG iterations/s; the state afterwards for the FP32-vs-FP64
    comparison."""
    from svirl_amd import workloads
    dtype = np.float32 if prec == "fp32" else np.float64
    s = 4 if prec == "fp32" else 8
    dt = 0.05
    log(f"[c3 {prec}] building Config 3 state, L={args.L}")
    gl = workloads.config3(L=args.L, dtype=dtype)
    N = int(gl.cfg.N)
    log(f"[c3 {prec}] N = {N}; warmup {args.warmup} steps, timing {args.steps}")
    r = run_td(gl, dt, args.warmup, args.steps)
    sp = np.array(r["sweeps"])
    out = dict(N=N, wall=r["wall"], kernel_s=r["kernel_s"], sweeps_psi=float(sp[:, 0].mean()),
               sweeps_A=float(sp[:, 1].mean()), node_steps_per_s=N * args.steps / r["wall"])
    pb, ab = td_pass_bytes(s, s, 1, N, r)
    out["all_gbs"] = (pb + ab) / r["kernel_s"] / 1e9
    log(f"[c3 {prec}] {out['node_steps_per_s']:.4g} node-steps/s ({1e3 * r['wall'] / args.steps:.3f} ms/step wall, "
        f"{1e3 * r['kernel_s'] / args.steps:.3f} ms/step on the GPU), sweeps/step psi {out['sweeps_psi']:.2f} "
        f"A {out['sweeps_A']:.2f}; passes {r['passes']}")
    if args.profile_steps:
        p = run_td(gl, dt, 0, args.profile_steps, profile=True)
        pb, ab = td_pass_bytes(s, s, 1, N, p)
        nps = p["passes"]
        out["psi_launch_us"] = 1e6 * p["psi_s"] / max(1, p["psi_n"])
        out["a_launch_us"] = 1e6 * p["a_s"] / max(1, p["a_n"])
        out["psi_gbs"] = pb / p["psi_s"] / 1e9
        out["a_gbs"] = ab / p["a_s"] / 1e9 if p["a_s"] > 0 else None
        out["psi_bytes_per_launch"] = pb / max(1, p["psi_n"])
        out["psi_passes_per_launch"] = nps[0] / max(1, p["psi_n"])
        out["psi_redone_per_launch"] = nps[1] / max(1, p["psi_n"])
        out["a_passes_per_launch"] = nps[2] / max(1, p["a_n"])
        spp = np.array(p["sweeps"])
        out["psi_sweeps_per_launch"] = float(spp[:, 0].mean())
        out["psi_per_sweep_GBs"] = float(spp[:, 0].sum()) * (8 * s + s + 1) * N / p["psi_s"] / 1e9
        log(f"[c3 {prec}] psi solve {out['psi_launch_us']:.1f} us/launch, {out['psi_passes_per_launch']:.2f} passes "
            f"({out['psi_sweeps_per_launch']:.2f} sweeps) per launch, {out['psi_gbs']:.1f} GB/s compulsory; "
            f"A solve {out['a_launch_us']:.1f} us/launch")
    out["F_td"] = gl.observables.free_energy_fp64()
    for _ in range(2):  # vortex detection on the final state; the second call is timed
        gl.solve.vortices_detected = False
        t0 = time.perf_counter()
        vx = gl.vortex_detector.vortices[0]
        out["vortex_ms"] = 1e3 * (time.perf_counter() - t0)
    out["vortices"] = int(len(vx))
    out["vortex_cells"] = int(len(gl.vortex_detector.vortex_cells[0]))
    log(f"[c3 {prec}] vortex detection {out['vortex_ms']:.1f} ms: {out['vortex_cells']} cells flagged, "
        f"{out['vortices']} vortices")
    jx, jy = gl.observables.current_density
    out["psi_h"] = np.asarray(gl.vars.order_parameter, dtype=np.complex128).ravel()
    out["J_h"] = np.concatenate([np.asarray(jx, dtype=np.float64).ravel(), np.asarray(jy, dtype=np.float64).ravel()])
    if not args.skip_cg:
        c = run_cg(gl, args.cg_iters)
        out["cg_iters_per_s"] = c["iters"] / c["wall"]
        out["cg_ms_per_iter"] = 1e3 * c["wall"] / c["iters"]
        out["F_cg"] = c["F"]
        log(f"[c3 {prec}] CG {out['cg_iters_per_s']:.2f} iters/s ({out['cg_ms_per_iter']:.2f} ms/iter)")
    del gl
    # against the oracle: a fresh state, the oracle's two steps
    z = gold("c3_4097")
    if args.L == 2048.0:
        g2 = workloads.config3(L=args.L, dtype=dtype)
        g2.solve.td(dt=dt, Nt=int(z["Nt"]))
        e = rows_errors(g2, z, "_td")
        e["F_rel"] = rel(g2.observables.free_energy_fp64(), float(z["F_td"]))
        e["sweeps_max_diff"] = int(np.max(np.abs(np.array(g2.solve._td.sweeps) - z["sweeps"])))
        e["horizon"] = f"{int(z['Nt'])} TD steps"
        out["vs_oracle"] = e
        log(f"[c3 {prec}] vs oracle after {int(z['Nt'])} steps: F {e['F_rel']:.2e}, psi {e['psi_L2_rel']:.2e}, "
            f"A {e['A_L2_rel']:.2e}, J {e['J_L2_rel']:.2e}")
        del g2
    return out


def bench_c1(args):
    """Config 1 (129^2, kappa = inf): TD Nt = 1000 then CG to the
    reference's energy stop rule, FP64 and FP32-mixed; rates, and errors
    against the oracle's whole run."""
    from svirl_amd import workloads
    z = gold("c1_129")
    res = {}
    for prec, dtype in (("fp64", np.float64), ("fp32", np.float32)):
        gl = workloads.config1(dtype)
        N = int(gl.cfg.N)
        gl.solve._init_td()
        gl.solve._td.steps_per_launch = 1000
        t0 = time.perf_counter()
        gl.solve.td(dt=0.1, Nt=1000)
        from svirl_amd import _native as nat
        nat.call("glx_synchronize")
        t_td = time.perf_counter() - t0
        F_td = gl.observables.free_energy_fp64()
        psi = gl.flatten_array(gl.vars.order_parameter).astype(np.complex128)
        jx, jy = gl.observables.current_density
        J = np.concatenate([gl.flatten_a_array(jx), gl.flatten_b_array(jy)]).astype(np.float64)
        gl.vortex_detector.vortices
        cells, _ = gl.vortex_detector.vortex_cells
        c = run_cg(gl, 1000, stop_rule=True)
        o = {"td_node_steps_per_s": N * 1000 / t_td, "td_s": t_td,
             "sweeps_psi_total": int(np.sum(np.array(gl.solve._td.sweeps)[:, 0])),
             "cg_iters": c["iters"], "cg_iters_per_s": c["iters"] / c["wall"],
             "vs_oracle": {"horizon": "the whole run (TD 1000 steps, then CG to the stop rule)",
                           "F_rel_after_td": rel(F_td, float(z["F_td"])),
                           "psi_L2_rel_after_td": l2rel(psi, z["psi_td"]),
                           "J_L2_rel_after_td": l2rel(J, z["J_td"]),
                           "vortex_cells_equal_after_td": bool(np.array_equal(cells, z["cells_td"])),
                           "sweeps_equal": bool(np.array_equal(np.array(gl.solve._td.sweeps), z["sweeps"])),
                           "cg_iters_oracle": int(len(z["cg_energies"])),
                           "F_rel_after_cg": rel(c["F"], float(z["cg_energies"][-1]))}}
        res[prec] = o
        v = o["vs_oracle"]
        log(f"[c1 {prec}] TD {o['td_node_steps_per_s']:.4g} node-steps/s ({1e3 * t_td:.1f} ms for 1000 steps), "
            f"CG {o['cg_iters_per_s']:.1f} iters/s ({c['iters']} iterations); vs oracle: F {v['F_rel_after_td']:.2e}, "
            f"psi {v['psi_L2_rel_after_td']:.2e}, J {v['J_L2_rel_after_td']:.2e}, cells equal "
            f"{v['vortex_cells_equal_after_td']}, CG F {v['F_rel_after_cg']:.2e}")
        del gl
    return res


def bench_c2(args):
    """Config 2 (1025^2, kappa = 5): TD and CG rates in both forms, and
    errors against the oracle after 3 TD steps and 5 CG iterations."""
    from svirl_amd import workloads
    z = gold("c2_1025")
    res = {}
    for prec, dtype in (("fp64", np.float64), ("fp32", np.float32)):
        s = 8 if prec == "fp64" else 4
        gl = workloads.config2(dtype)
        N = int(gl.cfg.N)
        gl.solve.td(dt=0.1, Nt=int(z["Nt"]))
        e = rows_errors(gl, z, "_td")
        e["F_rel"] = rel(gl.observables.free_energy_fp64(), float(z["F_td"]))
        e["sweeps_max_diff"] = int(np.max(np.abs(np.array(gl.solve._td.sweeps) - z["sweeps"])))
        gl.solve.cg(n_iter=len(z["cg_energies"]))
        e2 = rows_errors(gl, z, "_cg")
        e["cg_F_rel_max"] = float(np.max(np.abs(np.array(gl.solve._cg.cg_energies) - z["cg_energies"])
                                         / np.abs(z["cg_energies"])))
        e["cg_psi_L2_rel"] = e2["psi_L2_rel"]
        e["cg_J_L2_rel"] = e2["J_L2_rel"]
        e["horizon"] = f"{int(z['Nt'])} TD steps, then {len(z['cg_energies'])} CG iterations"
        del gl
        gl = workloads.config2(dtype)
        r = run_td(gl, 0.1, args.c2_warmup, args.c2_steps, profile=True)
        pb, ab = td_pass_bytes(s, 0, 0, N, r)
        c = run_cg(gl, args.c2_cg_iters)
        o = {"td_node_steps_per_s": N * args.c2_steps / r["wall"], "ms_per_step": 1e3 * r["wall"] / args.c2_steps,
             "steps": args.c2_steps, "sweeps_per_step": np.array(r["sweeps"]).mean(axis=0).tolist(),
             "psi_solve_GBs": pb / r["psi_s"] / 1e9, "cg_iters_per_s": c["iters"] / c["wall"],
             "cg_ms_per_iter": 1e3 * c["wall"] / c["iters"], "vs_oracle": e}
        res[prec] = o
        log(f"[c2 {prec}] TD {o['td_node_steps_per_s']:.4g} node-steps/s ({o['ms_per_step']:.3f} ms/step), CG "
            f"{o['cg_iters_per_s']:.1f} iters/s; vs oracle: F {e['F_rel']:.2e}, psi {e['psi_L2_rel']:.2e}, "
            f"A {e['A_L2_rel']:.2e}, J {e['J_L2_rel']:.2e}, CG F {e['cg_F_rel_max']:.2e}")
        del gl
    res["fp32_over_fp64_td"] = res["fp32"]["td_node_steps_per_s"] / res["fp64"]["td_node_steps_per_s"]
    return res


def run_batch(args):
    """Config 4: n instances (FP32-mixed), batched steps vs the same
    instances run one after another (each through its own persistent
    launches); the batched FP32 result against the oracle's sample and
    against the batched FP64 form."""
    from svirl_amd import _native as nat
    from svirl_amd import workloads
    from svirl_amd.batch import BatchTD
    log(f"[c4] building {args.batch_n} Config 4 instances")
    sims = workloads.config4(n=args.batch_n, dtype=np.float32)
    nodes = sum(int(s.cfg.N) for s in sims)
    b = BatchTD(sims, use_graph=True, autotune=True)
    b.td(dt=0.1, Nt=b.tune_need())  # autotune on the first steps
    nat.call("glx_synchronize")
    t0 = time.perf_counter()
    b.td(dt=0.1, Nt=args.batch_steps)
    nat.call("glx_synchronize")
    tb = time.perf_counter() - t0
    rate_b = nodes * args.batch_steps / tb
    log(f"[c4] {len(sims)} instances, {nodes} nodes: {rate_b:.4g} node-steps/s batched "
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
    log(f"[c4] one instance at a time: {rate_s:.4g} node-steps/s; speedup {rate_b / rate_s:.2f}x")
    sw = np.array([np.mean(np.array(s.solve._td.sweeps), axis=0) for s in sims])
    res = {"instances": len(sims), "nodes": nodes, "node_steps_per_s": rate_b, "ms_per_step": 1e3 * tb / args.batch_steps,
           "per_instance_node_steps_per_s": rate_s, "speedup_vs_per_instance": rate_b / rate_s,
           "tuned_shape": {"wg_per_cu": b.shape[0], "min_rows": b.shape[1], "psi_depth": b.shape[2],
                           "strip_columns": 64 - 2 * b.shape[2]},
           "tuning_ms_per_step": {f"wg{k[0]},rows{k[1]},depth{k[2]}": v for k, v in b.timings.items()},
           "mean_sweeps_per_step": {"psi": float(sw[:, 0].mean()), "A": float(sw[:, 1].mean())},
           "dtype": "f32", "graph": True}
    del sims
    # against the oracle: the whole batch from its seeded start, the oracle's steps
    z = gold("c4_sample")
    if args.batch_n == 256:
        sims = workloads.config4(n=args.batch_n, dtype=np.float32)
        BatchTD(sims, autotune=False).td(dt=0.1, Nt=int(z["Nt"]))
        fe, pe, je = [], [], []
        for n, k in enumerate(z["index"]):
            gl = sims[int(k)]
            fe.append(rel(gl.observables.free_energy_fp64(), float(z[f"F_{n}"])))
            pe.append(l2rel(gl.flatten_array(gl.vars.order_parameter).astype(np.complex128), z[f"psi_{n}"]))
            jx, jy = gl.observables.current_density
            je.append(l2rel(np.concatenate([gl.flatten_a_array(jx), gl.flatten_b_array(jy)]).astype(np.float64),
                            z[f"J_{n}"]))
        res["vs_oracle"] = {"horizon": f"{int(z['Nt'])} TD steps", "instances": [int(k) for k in z["index"]],
                            "F_rel_max": float(np.max(fe)), "psi_L2_rel_max": float(np.max(pe)),
                            "J_L2_rel_max": float(np.max(je))}
        log(f"[c4] FP32 vs oracle ({len(fe)} instances, {int(z['Nt'])} steps): F {np.max(fe):.2e}, psi {np.max(pe):.2e}, "
            f"J {np.max(je):.2e}")
        del sims
    # beyond the oracle's horizon: the first 32 instances, FP32-mixed vs FP64, batched, same steps
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
    log(f"[c4] FP32 vs FP64 over {k} instances: F {np.max(fe):.2e}, psi {np.max(pe):.2e}, J {np.max(je):.2e}")
    return res


def run_batch_cg(args):
    """Config 5: n instances of 257^2 (kappa = 5) minimised by the batched
    CG, FP32-mixed then FP64, up to --cg-batch-iters iterations with the
    ragged stop ||grad F|| <= 1e-8 ||grad F||_0; the first iteration of
    sampled instances against the oracle's exact line minimum. The
    FP32-mixed run stops an instance whose gradient norm has not halved in
    --cg-batch-stall iterations (its FP32 floor) and finishes it in FP64."""
    from svirl_amd import workloads
    from svirl_amd.batch import BatchCG
    z = gold("c5_sample")
    res = {"instances": args.cg_batch_n, "max_iters": args.cg_batch_iters, "rtol": 1e-8}
    F = {}
    for name, dtype in (("fp32", np.float32), ("fp64", np.float64)):
        specs = workloads.config5_specs(n=args.cg_batch_n)
        vo = {}
        if args.cg_batch_n == 1024:
            # the first iteration of the sampled instances, against the oracle
            # (instances of a batch are independent: a batch of the sampled
            # ones computes what the full batch computes for them)
            sample = [workloads.instance(specs[int(k)], dtype) for k in z["index"]]
            BatchCG(sample).cg(max_iter=1, rtol=0.0)
            fr = [(gl.observables.free_energy_fp64() - float(z[f"F1_{n}"])) / abs(float(z[f"F1_{n}"]))
                  for n, gl in enumerate(sample)]
            jr = []
            for n, gl in enumerate(sample):
                jx, jy = gl.observables.current_density
                jr.append(l2rel(np.concatenate([gl.flatten_a_array(jx), gl.flatten_b_array(jy)]).astype(np.float64),
                                z[f"J1_{n}"]))
            vo = {"horizon": "1 iteration (exact line minimum along -grad F)", "instances": [int(k) for k in z["index"]],
                  "F_rel_max": float(np.max(np.abs(fr))), "J_L2_rel_max": float(np.max(jr))}
            del sample
        # the timed run starts from the initial state, so ||grad F||_0 is the
        # initial state's gradient norm
        log(f"[c5] building {args.cg_batch_n} Config 5 instances ({name})")
        sims = [workloads.instance(sp, dtype) for sp in specs]
        nodes = sum(int(s.cfg.N) for s in sims)
        b = BatchCG(sims)
        mixed = dtype == np.float32
        t0 = time.perf_counter()
        iters, conv, grel = b.cg(max_iter=args.cg_batch_iters, rtol=1e-8,
                                 stall_iters=args.cg_batch_stall if mixed else 0, fp64_finish=mixed)
        t = time.perf_counter() - t0
        total = int(iters.sum())
        res[name] = {"instance_iters_per_s": total / t, "node_iters_per_s": float(np.sum(iters * (nodes / len(sims)))) / t,
                     "wall_s": t, "converged": int(conv.sum()), "iters_min": int(iters.min()),
                     "iters_max": int(iters.max()), "grad_rel_median": float(np.median(grel)), "vs_oracle": vo}
        if mixed:
            fin = b.fp64_iters > 0
            res[name].update({"stall_iters": args.cg_batch_stall, "finished_in_fp64": int(fin.sum()),
                              "converged_in_fp32": int((conv & ~fin).sum()),
                              "fp32_iters_total": int((iters - b.fp64_iters).sum()),
                              "fp64_finish_iters_total": int(b.fp64_iters.sum())})
        F[name] = np.array([s.observables.free_energy_fp64() for s in sims[:64]])
        log(f"[c5] {name}: {total / t:.4g} instance-iterations/s, {res[name]['node_iters_per_s']:.4g} "
            f"node-iterations/s, {int(conv.sum())} converged, median grad_rel {np.median(grel):.2e}; first step vs "
            f"oracle F {vo.get('F_rel_max', float('nan')):.2e}, J {vo.get('J_L2_rel_max', float('nan')):.2e}; "
            f"wall {t:.1f} s"
            + (f"; {res[name]['finished_in_fp64']} finished in FP64 after stalling" if mixed else ""))
        del sims, b
    res["fp32_vs_fp64_F_rel_max"] = float(np.max(np.abs(F["fp32"] - F["fp64"]) / np.abs(F["fp64"])))
    res["fp32_vs_fp64_F_rel_median"] = float(np.median(np.abs(F["fp32"] - F["fp64"]) / np.abs(F["fp64"])))
    log(f"[c5] FP32 vs FP64 free energy (first 64 instances): max rel {res['fp32_vs_fp64_F_rel_max']:.2e}")
    return res


def stream_ceiling(n=1 << 26):
    """The practical bandwidth ceiling of this device for the psi pass's
    traffic mix (glx_stream_probe kind 2: 37 bytes per node, 6 reads and one
    write, as a plain grid-stride stream), best over a few grid sizes, and
    the plain read and copy rates beside it."""
    from svirl_amd import _native as nat
    out = {}
    for name, kind, m in (("psi_mix", 2, n), ("read", 0, n // 2), ("copy", 1, n // 2)):
        best = 0.0
        for grid in (1024, 2048, 8192, 65536):
            g = ctypes.c_double()
            nat.call("glx_stream_probe", kind, m, grid, 5, ctypes.byref(g))
            best = max(best, g.value)
        out[name + "_GBs"] = best
    log(f"[stream] psi-pass mix {out['psi_mix_GBs']:.0f} GB/s, read {out['read_GBs']:.0f} GB/s, "
        f"copy {out['copy_GBs']:.0f} GB/s (best over grids)")
    return out


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


# ---------------------------------------------------------------------------
# CPU baselines (the FP64 oracle on the host cores)
# ---------------------------------------------------------------------------

def cpu_baseline(L, psi_sweeps, a_sweeps):
    """The FP64 oracle on the Config 3 workload itself (4097^2 at L = 2048):
    the timed sample is `psi_sweeps` psi Jacobi sweeps and `a_sweeps` A
    sweeps of the first step from the seeded start; node-steps/s follows
    from the time per sweep and the sweeps per step the GPU run logged."""
    os.environ.setdefault("OMP_NUM_THREADS", "1")
    from oracle import gl_oracle as O
    from svirl_amd import workloads
    dx = 0.5
    Nx = int(round(L / dx)) + 1
    g = O.Geom(Nx, Nx, dx, dx)
    mt = np.ascontiguousarray(workloads.holes_mask(g.Nxc, g.Nyc, dx, dx).T).ravel()
    em = np.ascontiguousarray(workloads.pinning_field(Nx, Nx, dx).T).ravel()
    psi = np.where(O.mask_at_nodes(g, mt), O.random_psi(g.N, 1.0, seed=3), 0)
    ab = O.symmetric_gauge(g, 0.05)
    st = O.OracleState(g, psi, ab, kappa=5.0, sigma=200.0, H=0.05, eps_map=em, mt=mt)
    t0 = time.perf_counter()
    cur, rhs = st.psi, st.psi.copy()
    for j in range(psi_sweeps):
        cur, rhs, _ = O.psi_jacobi_sweep(g, 0.05, st.eps_scalar, em, mt, ab, rhs, cur, jstep=j)
    t_psi = (time.perf_counter() - t0) / psi_sweeps
    t0 = time.perf_counter()
    A, rhs_a = ab.copy(), ab.copy()
    for j in range(a_sweeps):
        A, rhs_a, _ = O.a_jacobi_sweep(g, 0.05, st.kappa2, st.rho, st.H, mt, cur, ab, rhs_a, A, jstep=j)
    t_a = (time.perf_counter() - t0) / a_sweeps
    return g.N, t_psi, t_a


def _cpu_batch_task(spec):
    """One oracle task of the batched baselines (runs in a worker or
    inline): ('td', spec, n) -> n TD steps of a Config 4 instance; ('cg',
    spec, n) -> n CG iterations of a Config 5 instance. Returns (nodes * n,
    seconds)."""
    from oracle import gl_oracle as O
    from svirl_amd import workloads
    kind, s, n = spec
    g = O.Geom(s["Nx"], s["Ny"], s["dx"], s["dx"])
    em = workloads.pinning_field(s["Nx"], s["Ny"], s["dx"], seed=s["seed"])
    st = O.OracleState(g, O.random_psi(g.N, 1.0, seed=s["seed"]), O.symmetric_gauge(g, s["H"]), kappa=s["kappa"],
                       sigma=s["sigma"], H=s["H"], eps_map=np.ascontiguousarray(em.T).ravel())
    t0 = time.perf_counter()
    if kind == 'td':
        st.td(dt=0.1, Nt=n)
    else:
        st.cg(n_iter=n)
    return g.N * n, time.perf_counter() - t0


def cpu_batch_baselines(workers, n_td=16, n_cg=8, cg_iters=3):
    """The FP64 oracle on samples of Config 4 (one TD step of each of its
    first n_td instances) and Config 5 (cg_iters CG iterations of n_cg
    instances), on one core and on a pool of `workers` processes. Runs before
    the GPU is touched, so the pool is forked from a process without GPU
    state."""
    import multiprocessing as mp
    from svirl_amd import workloads
    os.environ.setdefault("OMP_NUM_THREADS", "1")
    td = [('td', s, 1) for s in workloads.config4_specs()[:n_td]]
    cg = [('cg', s, cg_iters) for s in workloads.config5_specs()[:n_cg]]
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
                         "pool_value": sum(w for w, _ in rp) / tp, "pool_cores": workers, "sample": sample,
                         "cpu": cpu_model()}
            log(f"[cpu] {name}: {work / t1:.4g} {unit} on 1 core, {sum(w for w, _ in rp) / tp:.4g} on {workers} "
                f"processes ({sample})")
    return out


def pmc_traffic():
    """Measured HBM bytes per node per memory pass of k_td_psi_solve<float>
    at Config 3 from the committed counter summary (FETCH_SIZE + WRITE_SIZE,
    scaled by the copy calibration), or None."""
    try:
        with open(PMC_SUMMARY) as f:
            return json.load(f).get("bytes_per_node_pass")
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
    ap.add_argument("--skip-c1", action="store_true")
    ap.add_argument("--skip-c2", action="store_true")
    ap.add_argument("--c2-steps", type=int, default=200)
    ap.add_argument("--c2-warmup", type=int, default=20)
    ap.add_argument("--c2-cg-iters", type=int, default=20)
    ap.add_argument("--skip-batch", action="store_true")
    ap.add_argument("--batch-n", type=int, default=256)
    ap.add_argument("--batch-steps", type=int, default=60)
    ap.add_argument("--batch-solo-steps", type=int, default=10)
    ap.add_argument("--skip-batch-cg", action="store_true")
    ap.add_argument("--cg-batch-n", type=int, default=1024)
    ap.add_argument("--cg-batch-iters", type=int, default=3000)
    ap.add_argument("--cg-batch-stall", type=int, default=300)
    ap.add_argument("--calibrate", action="store_true",
                    help="run FP32 axpy dispatches of known size first (counter calibration)")
    ap.add_argument("--cpu-psi-sweeps", type=int, default=3)
    ap.add_argument("--cpu-a-sweeps", type=int, default=2)
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
    nat.init(0)
    info = nat.device_info()
    log(f"device: {info}")
    if args.calibrate:
        calibrate_copy()
        log("[calib] 3 FP32 axpy dispatches over 2^26 reals")
    ceiling = stream_ceiling()
    res = {}
    for prec in (["fp32"] if args.skip_fp64 else ["fp32", "fp64"]):
        res[prec] = bench_c3(args, prec)
    c1 = None if args.skip_c1 else bench_c1(args)
    c2 = None if args.skip_c2 else bench_c2(args)
    batch = None if args.skip_batch else run_batch(args)
    batch_cg = None if args.skip_batch_cg else run_batch_cg(args)

    r32 = res["fp32"]
    cpu = None
    if not args.skip_cpu:
        log(f"[cpu] oracle on the Config 3 workload (L={args.L}): {args.cpu_psi_sweeps} psi and {args.cpu_a_sweeps} "
            f"A sweeps")
        N, t_psi, t_a = cpu_baseline(args.L, args.cpu_psi_sweeps, args.cpu_a_sweeps)
        sps, spa = r32["sweeps_psi"], r32["sweeps_A"]
        rate = N / (sps * t_psi + spa * t_a)
        cpu = {"value": rate, "unit": "node-steps/s", "cores": 1, "kind": "port", "cpu": cpu_model(),
               "host_cores": os.cpu_count(),
               "sample": f"FP64 NumPy oracle on the Config 3 workload ({N} nodes, seeded holes, pinning map): "
                         f"{args.cpu_psi_sweeps} psi sweeps ({t_psi:.2f} s each) and {args.cpu_a_sweeps} A sweeps "
                         f"({t_a:.2f} s each) of the first step, rated at the GPU run's {sps:.2f} psi and "
                         f"{spa:.2f} A sweeps per step"}
        log(f"[cpu] {rate:.4g} node-steps/s ({cpu['sample']})")

    per_node_pass = pmc_traffic()
    traffic = None
    if per_node_pass and r32.get("psi_passes_per_launch"):
        traffic = per_node_pass * r32["psi_passes_per_launch"] * r32["N"]
    roof = {"bound": "hbm", "achieved": r32.get("psi_gbs"), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": (r32["psi_gbs"] / HBM_PEAK_GBS) if r32.get("psi_gbs") else None,
            "traffic": traffic,
            "kernel": "k_td_psi_solve<float, gauge-shifted A>",
            "launch_us": r32.get("psi_launch_us"),
            "compulsory_bytes_per_node_pass": 37,
            "passes_per_launch": r32.get("psi_passes_per_launch"),
            "redone_passes_per_launch": r32.get("psi_redone_per_launch"),
            "sweeps_per_launch": r32.get("psi_sweeps_per_launch"),
            "algorithmic_bytes_per_launch": r32.get("psi_bytes_per_launch"),
            "traffic_bytes_per_node_pass": per_node_pass,
            "traffic_over_compulsory": (per_node_pass / 37.0) if per_node_pass else None,
            "physical_GBs": (traffic / (r32["psi_launch_us"] * 1e-6) / 1e9) if traffic else None,
            "physical_frac": (traffic / (r32["psi_launch_us"] * 1e-6) / 1e9 / HBM_PEAK_GBS) if traffic else None,
            "per_sweep_GBs": r32.get("psi_per_sweep_GBs"),
            "stream_ceiling": ceiling,
            "frac_of_stream_ceiling": (r32["psi_gbs"] / ceiling["psi_mix_GBs"]) if r32.get("psi_gbs") else None,
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
    precision = {}
    if "vs_oracle" in r32:
        precision["c3_fp32"] = r32["vs_oracle"]
    if c1:
        out["c1"] = c1
        precision["c1_fp32"] = c1["fp32"]["vs_oracle"]
    if c2:
        out["c2"] = c2
        precision["c2_fp32"] = c2["fp32"]["vs_oracle"]
    if batch:
        out["batch"] = batch
        if "vs_oracle" in batch:
            precision["c4_fp32"] = batch["vs_oracle"]
    if batch_cg:
        out["batch_cg"] = batch_cg
        if batch_cg["fp32"].get("vs_oracle"):
            precision["c5_fp32"] = batch_cg["fp32"]["vs_oracle"]
    if "cg_iters_per_s" in r32:
        out["cg"] = {"fp32_iters_per_s": r32["cg_iters_per_s"], "fp32_ms_per_iter": r32["cg_ms_per_iter"]}
    if "fp64" in res:
        r64 = res["fp64"]
        if "vs_oracle" in r64:
            precision["c3_fp64"] = r64["vs_oracle"]
        out["fp64"] = {"node_steps_per_s": r64["node_steps_per_s"], "ms_per_step": 1e3 * r64["wall"] / args.steps,
                       "psi_solve_GBs": r64.get("psi_gbs"),
                       "psi_solve_frac": (r64["psi_gbs"] / HBM_PEAK_GBS) if r64.get("psi_gbs") else None,
                       "psi_launch_us": r64.get("psi_launch_us"),
                       "psi_passes_per_launch": r64.get("psi_passes_per_launch"),
                       "sweeps_per_step": {"psi": r64["sweeps_psi"], "A": r64["sweeps_A"]}, "F": r64["F_td"]}
        out["fp32_over_fp64"] = r32["node_steps_per_s"] / r64["node_steps_per_s"]
        out["vortex_detector"].update({"fp64_ms": r64["vortex_ms"], "fp64_vortices": r64["vortices"],
                                       "fp64_cells_flagged": r64["vortex_cells"]})
        out["fp32_vs_fp64"] = {
            "after_steps": args.warmup + args.steps + args.profile_steps,
            "F_rel": rel(r32["F_td"], r64["F_td"]),
            "psi_L2_rel": l2rel(r32["psi_h"], r64["psi_h"]),
            "J_L2_rel": l2rel(r32["J_h"], r64["J_h"]),
            "reference": "FP64 GPU form on the same inputs, past the oracle's horizon"}
        if "cg_iters_per_s" in r64:
            out["cg"]["fp64_iters_per_s"] = r64["cg_iters_per_s"]
            out["cg"]["fp64_ms_per_iter"] = r64["cg_ms_per_iter"]
            out["cg"]["fp32_F_rel_err_vs_fp64"] = rel(r32["F_cg"], r64["F_cg"])
    out["fp32_vs_oracle"] = precision
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
