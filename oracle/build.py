"""ORACLE — test infrastructure only. Builds oracle/liboracle.so from oracle/*.c with gcc."""

import os
import subprocess

_HERE = os.path.dirname(os.path.abspath(__file__))
SOURCES = [os.path.join(_HERE, "wstack_ref.c")]
OUTPUT = os.path.join(_HERE, "liboracle.so")


def build(verbose=False):
    tmp = OUTPUT + ".tmp"
    cmd = ["gcc", "-O2", "-fPIC", "-shared", "-fopenmp", "-ffp-contract=off", *SOURCES,
           "-lm", "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, OUTPUT)
    return OUTPUT


if __name__ == "__main__":
    print(build(verbose=True))
