"""Compile the HIP library libcpfem_hip.so for gfx950 in-tree (hipcc, no JIT cache).

The translation units are compiled in parallel and linked into the package
directory, where ``cpfem_amd._native`` loads it.
"""
import os
import shutil
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "csrc")
INCLUDE = os.path.join(os.path.dirname(os.path.dirname(HERE)), "include")
LIB = os.path.join(HERE, "libcpfem_hip.so")
SOURCES = ("cp_capi.hip", "cp_kernels.hip", "cp_krylov.hip", "cp_adjoint.hip", "cp_direct.hip")
ARCH = os.environ.get("CPFEM_OFFLOAD_ARCH", "gfx950")
FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-I", INCLUDE,
         # packed FP32 instructions stay off: some of their forms have given wrong results on
         # MI355X with no error
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         # multiply-adds are fused per source expression only (not by the back end across
         # statements), so an inline function compiles to the same arithmetic in every kernel
         # it is inlined into: a parked point finished by the second assembly kernel gets
         # bitwise the results the first kernel would have produced
         "-ffp-contract=on"]


def _hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found: the HIP library cannot be built")


def build(verbose=False):
    """Compile every HIP source for gfx950 and link libcpfem_hip.so (always recompiles)."""
    hipcc = _hipcc()
    objdir = os.path.join(HERE, "_obj")
    os.makedirs(objdir, exist_ok=True)
    procs = []
    for src in SOURCES:
        obj = os.path.join(objdir, src.replace(".hip", ".o"))
        cmd = [hipcc, *FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
        procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
    objs = []
    for cmd, p in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            raise RuntimeError("HIP compile failed:\n" + " ".join(cmd) + "\n" + out.decode(errors="replace"))
        if verbose:
            print(" ".join(cmd))
        objs.append(cmd[-1])
    tmp = LIB + ".tmp"
    cmd = [hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
    if r.returncode != 0:
        raise RuntimeError("HIP link failed:\n" + " ".join(cmd) + "\n" + r.stdout.decode(errors="replace"))
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
