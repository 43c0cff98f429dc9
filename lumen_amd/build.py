"""Build the engine's shared library in-tree: ``lumen_amd/liblumen_hip.so`` for gfx950 (hipcc).

    python -m lumen_amd.build        # or lumen_amd.build.build()

Every HIP source is compiled with packed FP32 instructions disabled (``-target-feature -packed-fp32-ops``),
hardware FP64 atomics, and FP32 division / square root by the hardware's approximations with denormals flushed; the
library exports the C ABI of ``include/lumen_hip.h`` only.
"""

from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "liblumen_hip.so")
SOURCES = [os.path.join(HERE, "csrc", f) for f in ("engine.hip", "unified_tables.cpp", "bvh_build.cpp")]
DEPS = SOURCES + [os.path.join(HERE, "csrc", f) for f in ("nav.h", "optics.h", "philox.h", "adjoint.h", "unified.h")] + [
    os.path.join(ROOT, "include", "lumen_hip.h")]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("LUMEN_ARCH", "gfx950")

# FP32 division and square root by the hardware's reciprocal and root (<= 2.5 ulp, not correctly rounded) with
# float32 denormals flushed: about 10 instructions fewer per division on a VALU-bound loop; FP64 is untouched
FLAGS = ["-O3", "-std=c++17", f"--offload-arch={ARCH}", "-fPIC", "-shared", "-munsafe-fp-atomics",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-fvisibility=hidden",
         "-fno-hip-fp32-correctly-rounded-divide-sqrt", "-fgpu-flush-denormals-to-zero", "-Wno-unused-result"]


def build(verbose: bool = False, extra=()) -> str:
    """Compile the library (always) and return its path."""
    cmd = [HIPCC, *FLAGS, *extra, "-I", os.path.join(ROOT, "include"), *SOURCES, "-o", LIB + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed ({r.returncode}):\n{r.stderr[-6000:]}")
    os.replace(LIB + ".tmp", LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True, extra=sys.argv[1:]))
