"""Builds oracle/cpu/libcpfem_cpu.so (g++, OpenMP) in-tree.

Two code generations are built: one for x86-64-v3 (AVX2 + FMA), loaded where the host CPU
has those instructions, and a baseline x86-64 one for any other host."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "cpfem_cpu.cpp")
LIBS = {"v3": os.path.join(HERE, "libcpfem_cpu_v3.so"), "base": os.path.join(HERE, "libcpfem_cpu.so")}
FLAGS = {"v3": ["-march=x86-64-v3", "-mtune=generic"], "base": []}


def build():
    for k, out in LIBS.items():
        cmd = ["g++", "-O3", "-std=c++17", "-fopenmp", "-fPIC", "-shared", *FLAGS[k], "-o", out + ".tmp", SRC]
        r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
        if r.returncode != 0:
            raise RuntimeError("CPU baseline build failed:\n" + " ".join(cmd) + "\n" + r.stdout.decode(errors="replace"))
        os.replace(out + ".tmp", out)
    return LIBS


if __name__ == "__main__":
    print(build())
