"""Ahead-of-time build of libglx.so for gfx950 with hipcc.

The FP64 kernels are compiled without FMA contraction (separate multiply
and add roundings, as the reference's arithmetic is written); the FP32
kernels allow contraction. All HIP code is built with packed-FP32 ops
disabled. Output: svirl_amd/_lib/libglx.so (in-tree, so it travels with
the source tree).
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "_lib")
LIB = os.path.join(LIBDIR, "libglx.so")
ARCH = os.environ.get("GLX_OFFLOAD_ARCH", "gfx950")

COMMON = [
    "-std=c++17", "-O3", "-fPIC", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]
UNITS = [
    ("glx_f64.hip", ["-ffp-contract=off"]),
    ("glx_f32.hip", ["-ffp-contract=fast"]),
    ("glx_api.hip", []),
]


def hipcc():
    exe = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(exe):
        raise RuntimeError("hipcc not found")
    return exe


def build(verbose=False, extra_flags=()):
    """Compile every HIP unit and link libglx.so. Returns the library path."""
    os.makedirs(LIBDIR, exist_ok=True)
    objdir = os.path.join(LIBDIR, "obj")
    os.makedirs(objdir, exist_ok=True)
    objs = []
    procs = []
    for src, flags in UNITS:
        obj = os.path.join(objdir, src.replace(".hip", ".o"))
        cmd = [hipcc()] + COMMON + list(flags) + list(extra_flags) + ["-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
        objs.append(obj)
    failed = []
    for src, cmd, p in procs:
        out, _ = p.communicate()
        text = out.decode(errors="replace")
        # hipcc's host pass does not know the packed-fp32 feature; that warning is expected
        lines = [ln for ln in text.splitlines() if "packed-fp32-ops" not in ln]
        if verbose and lines:
            print("\n".join(lines), flush=True)
        if p.returncode != 0:
            failed.append((src, "\n".join(lines)))
    if failed:
        raise RuntimeError("hipcc failed:\n" + "\n\n".join(f"[{s}]\n{t}" for s, t in failed))
    tmp = LIB + ".tmp"
    cmd = [hipcc(), "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + objs
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print(LIB)
