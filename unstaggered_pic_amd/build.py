"""Build the HIP library libupic.so in-tree for gfx950.

Usage: python -m unstaggered_pic_amd.build  (or build() from Python)
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libupic.so")
SOURCES = ["particle_kernels.hip", "diag_kernels.hip", "outer_dev.hip", "mixed.hip", "mlp.hip", "field_kernels.hip", "api.hip", "sim.hip", "init.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "-std=c++17", "-O3", "-fPIC", "--offload-arch=gfx950", "-munsafe-fp-atomics",
    # keep the compiler from forming packed FP32 ops by itself
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result", "-Wno-unused-function",
]


def _compile(src, defines=(), tag=""):
    obj = os.path.join(CSRC, "_build" + tag, src + ".o")
    os.makedirs(os.path.dirname(obj), exist_ok=True)
    cmd = [HIPCC, *FLAGS, *[f"-D{d}" for d in defines], "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr}")
    return obj, r.stderr


def build(verbose=False, defines=(), out=None):
    """Compile and link the library.  `defines` (e.g. ["UPIC_MIXED_SCHED=0"]) and
    `out` build a tuning variant next to the default library; the loader
    picks a variant up through the UPIC_LIBRARY environment variable."""
    out = out or OUT
    tag = "" if out == OUT else "_" + os.path.basename(out).replace(".", "_")
    with ThreadPoolExecutor(max_workers=min(4, os.cpu_count() or 1)) as ex:
        results = list(ex.map(lambda s: _compile(s, defines, tag), SOURCES))
    for _, err in results:
        if verbose and err.strip():
            print(err, file=sys.stderr)
    objs = [o for o, _ in results]
    tmp = out + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", "--offload-arch=gfx950", *objs, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, out)
    if out == OUT:
        build_driver()
    return out


def build_driver():
    """The deck-driven driver (csrc/upic_run.cpp), a host C++17 program on
    the C ABI, next to the library (found through its rpath)."""
    exe = os.path.join(HERE, "upic_run")
    cmd = [os.environ.get("CXX", "g++"), "-std=c++17", "-O2", "-Wall", os.path.join(CSRC, "upic_run.cpp"),
           "-L" + HERE, "-lupic", "-Wl,-rpath,$ORIGIN", "-o", exe + ".tmp"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"driver build failed:\n{r.stderr}")
    os.replace(exe + ".tmp", exe)
    return exe


if __name__ == "__main__":
    import argparse
    ap = argparse.ArgumentParser()
    ap.add_argument("-D", dest="defines", action="append", default=[])
    ap.add_argument("-o", dest="out", default=None)
    a = ap.parse_args()
    print(build(verbose=True, defines=a.defines, out=a.out and os.path.abspath(a.out)))
