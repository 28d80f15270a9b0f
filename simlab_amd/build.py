"""Builds libsimlab_cfd.so (gfx950 kernels + host library) in-tree.

hipcc compiles the kernels and the HIP host code; g++ compiles the OpenMP
voxelizer. The shared library lands next to this file; objects go to
simlab_amd/_build/. Every build recompiles (the caller decides when).
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libsimlab_cfd.so")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")
ARCH = "gfx950"

# packed-FP32 instructions are disabled for every HIP translation unit; FMA
# contraction is on except where a `#pragma clang fp contract(off)` says not
HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-ffp-contract=fast-honor-pragmas",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-result", "-Wno-unused-const-variable",
]
CXX_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-fopenmp", "-Wall", "-Wno-unused-function",
             f"-I{ROCM}/include", "-D__HIP_PLATFORM_AMD__"]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout)
        raise RuntimeError(f"build step failed: {' '.join(cmd[:3])} ...")
    elif verbose and r.stdout.strip():
        # hipcc's host pass warns about the packed-fp32 feature; keep it short
        lines = [l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l]
        if lines:
            print("\n".join(lines[-20:]))


# kernels.hip is compiled once per population storage (SIMLAB_KSTORE =
# simlab_storage value) and once for the storage-independent kernels and the
# dispatch (-1); all translation units compile in parallel
KSTORES = (-1, 0, 1, 2, 3)


def build(verbose=False, save_temps=False):
    os.makedirs(OUT, exist_ok=True)
    jobs = []
    for k in KSTORES:
        obj = os.path.join(OUT, f"kernels_{k if k >= 0 else 'common'}.o")
        cmd = [HIPCC, *HIP_FLAGS, f"-DSIMLAB_KSTORE={k}", "-x", "hip", "-c", os.path.join(CSRC, "kernels.hip"), "-o", obj]
        if save_temps:
            cmd.insert(1, "--save-temps")
        jobs.append((cmd, obj))
    obj = os.path.join(OUT, "mkernels.o")
    jobs.append(([HIPCC, *HIP_FLAGS, "-x", "hip", "-c", os.path.join(CSRC, "mkernels.hip"), "-o", obj], obj))
    obj = os.path.join(OUT, "mlattice.cpp.o")
    jobs.append(([HIPCC, *HIP_FLAGS, "-x", "hip", "-c", os.path.join(CSRC, "mlattice.cpp"), "-o", obj], obj))
    obj = os.path.join(OUT, "probe.o")
    jobs.append(([HIPCC, *HIP_FLAGS, "-x", "hip", "-c", os.path.join(CSRC, "probe.hip"), "-o", obj], obj))
    obj = os.path.join(OUT, "voxel_gpu.o")
    jobs.append(([HIPCC, *HIP_FLAGS, "-x", "hip", "-c", os.path.join(CSRC, "voxel_gpu.hip"), "-o", obj], obj))
    obj = os.path.join(OUT, "lattice.cpp.o")
    jobs.append(([HIPCC, *HIP_FLAGS, "-x", "hip", "-c", os.path.join(CSRC, "lattice.cpp"), "-o", obj], obj))
    obj = os.path.join(OUT, "voxel.cpp.o")
    jobs.append((["g++", *CXX_FLAGS, "-c", os.path.join(CSRC, "voxel.cpp"), "-o", obj], obj))
    procs = []
    for cmd, obj in jobs:
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)))
    failed = None
    for cmd, pr in procs:
        out, _ = pr.communicate()
        if pr.returncode != 0:
            sys.stderr.write(out)
            failed = failed or cmd
        elif verbose and out.strip():
            # hipcc's host pass warns about the packed-fp32 feature; keep it short
            lines = [l for l in out.splitlines() if "packed-fp32-ops" not in l]
            if lines:
                print("\n".join(lines[-20:]))
    if failed:
        raise RuntimeError(f"build step failed: {' '.join(failed[:3])} ...")
    objs = [obj for _, obj in jobs]
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp, *objs, "-lgomp"], verbose)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True, save_temps="--save-temps" in sys.argv)
    print(LIB)
