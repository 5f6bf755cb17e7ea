"""Build the oracle's C++ CPU baselines (oracle/_build/libcg_host.so, libgn_host.so) with g++."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "_build")
LIB = os.path.join(OUT, "libcg_host.so")
GN_LIB = os.path.join(OUT, "libgn_host.so")


def _one(src, lib):
    cmd = ["g++", "-O3", "-march=x86-64-v3", "-fopenmp", "-fPIC", "-shared", "-std=c++17", src, "-o", lib + ".tmp"]
    subprocess.run(cmd, check=True)
    os.replace(lib + ".tmp", lib)


def build():
    src = os.path.join(HERE, "cg_host.cpp")
    if not os.path.exists(src):
        return None
    os.makedirs(OUT, exist_ok=True)
    _one(src, LIB)
    _one(os.path.join(HERE, "gn_host.cpp"), GN_LIB)
    return LIB


if __name__ == "__main__":
    print(build())
