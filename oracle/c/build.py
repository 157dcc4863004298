"""Build oracle/c/liboracle_c.so (the oracle's C restatement; test infrastructure only).

Plain gcc, generic x86-64 code (no -march=native: the library is built in one container and run on
another host)."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle_c.so")


def build():
    tmp = f"{LIB}.{os.getpid()}.tmp"     # concurrent builds (parallel test workers) never share a file
    subprocess.run(["gcc", "-O3", "-fPIC", "-shared", "-std=c11", "-D_POSIX_C_SOURCE=200809L", "-Wall",
                    "-o", tmp, os.path.join(HERE, "mpf_c.c"), "-lpthread"], check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build())
