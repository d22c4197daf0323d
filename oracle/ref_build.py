"""Compile the reference's host FL loops (src/compression/afl/afl_gpu.cu, the
only CPU implementation on its compression path) from the sources under
/root/reference into oracle/_ref/libddj_ref.so.  Test infrastructure only:
it cross-checks the oracle's FL restatement and times the reference's CPU
loop.  The file uses CUDA kernel syntax, so it is compiled host-side by
hipcc with the HIP runtime header pre-included (the role nvcc's implicit
cuda_runtime.h include plays); no header, library or generated file of the
reference is substituted."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REF_SRC = "/root/reference/src"
OUT = os.path.join(HERE, "_ref")


def build() -> str:
    afl = os.path.join(REF_SRC, "compression", "afl", "afl_gpu.cu")
    if not os.path.exists(afl):
        raise SystemExit("reference sources not present; nothing to build")
    os.makedirs(OUT, exist_ok=True)
    lib = os.path.join(OUT, "libddj_ref.so")
    cmd = ["/opt/rocm/bin/hipcc", "-x", "hip", "--offload-arch=gfx950", "-O2", "-fPIC", "-shared",
           "-ffp-contract=off", "-w", "-include", "hip/hip_runtime.h", f"-I{REF_SRC}", afl, os.path.join(HERE, "ref_driver.cpp"), "-o", lib]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise SystemExit(f"reference build failed:\n{' '.join(cmd)}\n{r.stderr[-4000:]}")
    return lib


if __name__ == "__main__":
    print(build())
