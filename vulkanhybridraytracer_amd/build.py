"""Builds libhr.so (HIP for gfx950) in-tree with hipcc.

Flags: -ffp-contract=off keeps every float expression as written (no implicit
FMA), correctly rounded fp32 division and sqrt, packed-FP32 instructions
disabled."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "libhr.so")
SOURCES = ["csrc/lbvh.hip", "csrc/trace.hip", "csrc/wavefront.hip", "csrc/denoise.hip", "csrc/temporal.hip",
           "csrc/gbuffer_io.hip", "csrc/rt.hip", "csrc/hr_api.hip", "csrc/dn_api.hip",
           "csrc/image_io.cpp", "csrc/assets.cpp"]  # the .cpp files are host-only (asset ingest)
HEADERS = ["csrc/device_math.h", "csrc/internal.h", "csrc/trace_common.h", "csrc/api_internal.h", "csrc/image_io.h"]
LIBS = ["-lz"]
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off",
         "-fhip-fp32-correctly-rounded-divide-sqrt", "-Xclang", "-target-feature", "-Xclang",
         "-packed-fp32-ops", "-Wall", "-Wno-unused-function",
         # MFMA accumulators in VGPRs (gfx950 reads and writes them there): no
         # copies between VGPRs and AGPRs around the denoiser's MFMAs
         "-mllvm", "-amdgpu-mfma-vgpr-form"]


def hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if c and (os.path.sep not in c or os.path.exists(c)):
            return c
    return "hipcc"


def build(out=LIB, extra=(), verbose=False):
    cmd = [hipcc()] + FLAGS + list(extra) + ["-I", os.path.join(ROOT, "include")] + \
        [os.path.join(HERE, s) for s in SOURCES] + LIBS + ["-o", out + ".tmp"]
    if verbose:
        print(" ".join(cmd))
    subprocess.check_call(cmd)
    os.replace(out + ".tmp", out)
    return out


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
