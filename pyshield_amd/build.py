"""Build the HIP library libshield_mp.so in-tree for gfx950.

Every .hip source under csrc/ is compiled to an object with hipcc (in
parallel) and linked into pyshield_amd/libshield_mp.so.  Packed FP32 VALU
instructions are disabled for all device code (``-packed-fp32-ops``).
"""

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libshield_mp.so")
ARCH = os.environ.get("SHIELD_OFFLOAD_ARCH", "gfx950")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
COMMON = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-fno-hip-fp32-correctly-rounded-divide-sqrt",
    # keep loop-invariant constants of the sub-step loop from being hoisted
    # into registers that stay live across the whole loop body
    "-mllvm", "-disable-machine-licm",
    "-I", os.path.join(HERE, os.pardir, "include"),
    "-Wall", "-Wno-unused-function",
]


# The single-sub-step forward kernel (mp_forward_s<S>.hip) is scheduled for
# instruction-level parallelism: at three wavefronts per SIMD its dependent
# FP64 chains are not all hidden by the other wavefronts, and the max-ilp
# strategy interleaves independent work without more registers (166 VGPRs, no
# spills).  Measured on the C384 state: mixed 10.24 -> 9.95 ms, FP64 17.06 ->
# 16.81 ms.  The looping
# instances (mp_forward_multi_s<S>.hip) keep the default scheduler: with
# max-ilp the two-sub-step FP64 kernel ran 3.05 ms against 2.61 ms at C96.
# The adjoint and the update kernels do not gain from it.
PER_FILE = {
    "mp_forward_s": ["-mllvm", "--amdgpu-sched-strategy=max-ilp"],
}


def flags_for(src):
    name = os.path.basename(src)
    return [f for prefix, fl in PER_FILE.items() if name.startswith(prefix) for f in fl]


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))


def _compile(src, verbose):
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    cmd = [HIPCC, *COMMON, *flags_for(src), "-c", src, "-o", obj]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{r.stdout}\n{r.stderr}")
    if verbose and r.stderr.strip():
        print(r.stderr, file=sys.stderr)
    return obj


def build(verbose=False, jobs=None):
    """Compile every HIP source and link libshield_mp.so.  Returns its path.
    The mixed kernel's saturation table (csrc/qs_table.h) is regenerated
    first from the constants it is built from."""
    r = subprocess.run([sys.executable, os.path.join(CSRC, "gen_qs_table.py")], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"gen_qs_table.py failed:\n{r.stdout}\n{r.stderr}")
    os.makedirs(BUILD, exist_ok=True)
    srcs = sources()
    jobs = jobs or min(len(srcs), os.cpu_count() or 1)
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(lambda s: _compile(s, verbose), srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
