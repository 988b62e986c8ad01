"""Build the hipng shared library for gfx950 in-tree with hipcc."""

import os
import subprocess

_HERE = os.path.dirname(os.path.abspath(__file__))
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
SOURCES = [os.path.join(_HERE, "csrc", "hipng.hip")]
OUTPUT = os.path.join(_HERE, "_hipng.so")

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    # packed FP32 VALU forms stay off for all device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics",
]


def build(verbose=False):
    """Compile every HIP source into rascil_amd/_hipng.so (always recompiles)."""
    hipcc = os.path.join(ROCM, "bin", "hipcc")
    tmp = OUTPUT + ".tmp"
    cmd = [hipcc, *HIPCC_FLAGS, *SOURCES, f"-L{ROCM}/lib", "-lrocfft",
           f"-Wl,-rpath,{ROCM}/lib", "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, OUTPUT)
    return OUTPUT


if __name__ == "__main__":
    print(build(verbose=True))
