"""ORACLE — test infrastructure only. Builds oracle/liboracle.so from oracle/*.c with gcc."""

import os
import subprocess

_HERE = os.path.dirname(os.path.abspath(__file__))
SOURCES = [os.path.join(_HERE, "wstack_ref.c")]
OUTPUT = os.path.join(_HERE, "liboracle.so")


# host-only sanitizer build of the same sources (gcc; the library holds no GPU code)
SANITIZE_FLAGS = ["-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
                  "-fno-sanitize-recover=all"]


def build(verbose=False, sanitize=False, output=None):
    """Compile oracle/*.c into `output` (default oracle/liboracle.so); with `sanitize` under
    AddressSanitizer + UndefinedBehaviorSanitizer (load it with libasan preloaded)."""
    output = output or OUTPUT
    tmp = output + ".tmp"
    opt = SANITIZE_FLAGS if sanitize else ["-O2"]
    cmd = ["gcc", *opt, "-fPIC", "-shared", "-fopenmp", "-ffp-contract=off", *SOURCES,
           "-lm", "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, output)
    return output


if __name__ == "__main__":
    print(build(verbose=True))
