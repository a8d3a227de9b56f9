"""Builds librmh.so (HIP kernels + C ABI) for gfx950 in-tree with hipcc."""
import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "librmh.so")
SOURCES = ["rmh.cpp", "intra_subset.cpp", "rdsearch.cpp", "cabac.cpp", "loopfilter.cpp", "residual.hip", "mc.hip", "deblock.hip", "sao.hip",
           "cnn.hip", "loopfilter.hip", "forward.hip", "mc_refine.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared", "-fvisibility=hidden",
    "-Wall", "-Wno-unused-function", "-Wno-unused-value", "-Wno-unused-result", "-Wno-missing-braces",
    # packed FP32 instructions stay off (no floating point on this path anyway)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd))
    r = subprocess.run(cmd, capture_output=True, text=True)
    out = "\n".join(l for l in (r.stdout + r.stderr).splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + out)
    if verbose and out.strip():
        print(out)


def _stale(src, obj):
    """True when obj is missing or older than its source or any header in csrc/."""
    if not os.path.exists(obj):
        return True
    t = os.path.getmtime(obj)
    headers = [os.path.join(CSRC, h) for h in os.listdir(CSRC) if h.endswith(".h")]
    headers.append(os.path.join(os.path.dirname(HERE), "include", "rmh.h"))
    return any(os.path.getmtime(f) > t for f in [src] + headers)


def build(verbose=False, incremental=False):
    """Compiles every source to an object in parallel (one hipcc per file),
    then links librmh.so. With incremental, sources whose object is newer
    than the source and every header are not recompiled."""
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    compile_flags = [f for f in FLAGS if f != "-shared"]
    jobs = []
    for s in SOURCES:
        obj = os.path.join(objdir, s.replace(".", "_") + ".o")  # loopfilter.cpp and loopfilter.hip both build
        jobs.append(([HIPCC] + compile_flags + ["-c", os.path.join(CSRC, s), "-o", obj], obj))
    todo = [cmd for cmd, obj in jobs if not incremental or _stale(cmd[-3], obj)]
    workers = max(1, min(len(jobs), os.cpu_count() or 1, 16))
    with concurrent.futures.ThreadPoolExecutor(workers) as pool:
        for f in [pool.submit(_run, cmd, verbose) for cmd in todo]:
            f.result()
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC"] + [o for _, o in jobs] + ["-o", tmp], verbose)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True, incremental="--incremental" in sys.argv)
    print("built", LIB)
