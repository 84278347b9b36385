"""Build recipe for libotg.so (HIP kernels + C ABI) for gfx950.

`python -m opentimer_amd.build` compiles every source under csrc/ with
hipcc and links opentimer_amd/libotg.so in place.  The library has no
PyTorch dependency; Python reaches it through ctypes (opentimer_amd.native).
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
import tempfile
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")
LIB = os.path.join(HERE, "libotg.so")
ARCH = os.environ.get("OTG_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    # packed FP32 VALU forms stay off (hipcc would otherwise emit them from
    # float2/float4 arithmetic)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics",
    # keep every atomic as written: the wave-aggregating rewrite consumes the
    # returned value at once, which forces a wait for all outstanding memory
    # operations where the kernels issue an atomic ahead of its use
    "-mllvm", "-amdgpu-atomic-optimizer-strategy=None",
    "-Wall", "-Wno-unused-function", "-Wno-unused-result",
    f"-I{INCLUDE}",
]
CXX_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-Wall", f"-I{INCLUDE}", "-pthread"]


def _hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def sources():
    hip = sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))
    cpp = sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".cpp"))
    return hip, cpp


def build(verbose: bool = False, jobs: int | None = None) -> str:
    hipcc = _hipcc()
    hip, cpp = sources()
    tmp = tempfile.mkdtemp(prefix="otg_build_")
    try:
        cmds = []
        objs = []
        for src in hip:
            obj = os.path.join(tmp, os.path.basename(src) + ".o")
            cmds.append([hipcc, *HIP_FLAGS, "-c", src, "-o", obj])
            objs.append(obj)
        for src in cpp:
            obj = os.path.join(tmp, os.path.basename(src) + ".o")
            cmds.append(["g++", *CXX_FLAGS, "-c", src, "-o", obj])
            objs.append(obj)

        def run(cmd):
            if verbose:
                print(" ".join(cmd), flush=True)
            r = subprocess.run(cmd, capture_output=True, text=True)
            if r.returncode != 0:
                raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
            if verbose and r.stderr.strip():
                print(r.stderr, file=sys.stderr)
            return r

        with ThreadPoolExecutor(max_workers=jobs or min(8, os.cpu_count() or 1)) as ex:
            list(ex.map(run, cmds))
        out_tmp = LIB + ".tmp"
        run([hipcc, "-shared", "-fPIC", f"--offload-arch={ARCH}", *objs, "-o", out_tmp,
             "-pthread", "-Wl,-soname,libotg.so"])
        os.replace(out_tmp, LIB)
        build_bridge(run)
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    return LIB


BRIDGE_SRC = os.path.join(HERE, "bridge", "ot_cuda_bridge.cpp")
BRIDGE_LIB = os.path.join(HERE, "libot_cuda_bridge.so")
HARNESS_SRC = os.path.join(os.path.dirname(HERE), "tests", "bridge", "harness.cpp")
HARNESS_LIB = os.path.join(os.path.dirname(HERE), "tests", "bridge", "libbridge_harness.so")


def build_bridge(run) -> None:
    """The OpenTimer bridge (the reference's GPU entry points with their C++
    signatures, forwarding to libotg) and, when the test tree is present, the
    test harness that calls it the way the reference's timer does."""
    bridge_inc = os.path.join(HERE, "bridge")
    run(["g++", *CXX_FLAGS, f"-I{bridge_inc}", "-shared", BRIDGE_SRC, f"-L{HERE}", "-lotg",
         "-Wl,-rpath,$ORIGIN", "-Wl,-soname,libot_cuda_bridge.so", "-o", BRIDGE_LIB + ".tmp"])
    os.replace(BRIDGE_LIB + ".tmp", BRIDGE_LIB)
    if os.path.exists(HARNESS_SRC):
        run(["g++", *CXX_FLAGS, f"-I{bridge_inc}", "-shared", HARNESS_SRC, f"-L{HERE}",
             "-lot_cuda_bridge", "-Wl,-rpath,$ORIGIN/../../opentimer_amd",
             "-o", HARNESS_LIB + ".tmp"])
        os.replace(HARNESS_LIB + ".tmp", HARNESS_LIB)


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
