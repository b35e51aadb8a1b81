"""Builds libard.so (HIP kernels for gfx950 + the C-ABI) in this directory.

Usage: python build_lib.py   (called by __graft_entry__.build()).
"""
import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libard.so")
OBJ = os.path.join(HERE, "_build")

HIP_SOURCES = ["ard_kernels.hip", "ard_batched_kernels.hip", "ard_batched.hip", "ard_ctx.cpp", "ard_plain.cpp",
               "ard_adjoint_kernels.hip", "ard_adjoint.cpp", "ard_batch.cpp", "ard_output.hip",
               "ard_view.hip", "ard_dct3d.hip", "ard_fft.hip", "ard_check.hip", "ard_fold_direct.hip"]
CXX_SOURCES = ["ard_topology.cpp", "ard_masks.cpp"]
HOST_SOURCES = ["host/ard_host.cpp"]

ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.environ.get("HIPCC", os.path.join(ROCM, "bin", "hipcc"))
ARCH = "gfx950"
# Packed FP32 instructions are disabled for all device code.
COMMON = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-parameter", "-Wno-unused-variable"]
DEVICE = ["--offload-arch=" + ARCH, "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-munsafe-fp-atomics"]


def run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l and l.strip())
    if r.returncode != 0:
        sys.stderr.write(out + "\n")
        raise SystemExit("build failed: " + " ".join(cmd))
    return out


def build(verbose=False, jobs=None):
    os.makedirs(OBJ, exist_ok=True)
    cmds, objs = [], []
    for src in HIP_SOURCES:
        obj = os.path.join(OBJ, src + ".o")
        cmds.append([HIPCC, "-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj] + COMMON + DEVICE)
        objs.append(obj)
    for src in CXX_SOURCES:
        obj = os.path.join(OBJ, src + ".o")
        cmds.append(["g++", "-c", os.path.join(CSRC, src), "-o", obj] + COMMON)
        objs.append(obj)
    for src in HOST_SOURCES:  # the C++ class mirror, over the C-ABI
        obj = os.path.join(OBJ, os.path.basename(src) + ".o")
        cmds.append(["g++", "-c", os.path.join(HERE, src), "-o", obj] + COMMON)
        objs.append(obj)
    # one compiler process per source, at most `jobs` (default: CPUs, up to 8) at a time
    jobs = jobs or min(8, os.cpu_count() or 1)
    with concurrent.futures.ThreadPoolExecutor(max_workers=jobs) as ex:
        for _ in ex.map(run, cmds):
            pass
    tmp = OUT + ".tmp"
    run([HIPCC, "-shared", "-o", tmp] + objs + ["--offload-arch=" + ARCH, "-fPIC"])
    os.replace(tmp, OUT)
    # command-line driver (the reference's main loop without the window)
    run(["g++", "-O2", "-std=c++17", "-D__HIP_PLATFORM_AMD__", "-I" + os.path.join(ROCM, "include"),
         "-o", os.path.join(HERE, "ard_run"), os.path.join(HERE, "host", "ard_run.cpp"),
         "-L" + HERE, "-lard", "-L" + os.path.join(ROCM, "lib"), "-lamdhip64", "-Wl,-rpath,$ORIGIN"])
    # measurement probes behind figures in the design notes (probes/*.hip):
    # matrix-core throughput and the modal pass's streaming floor
    probes = []
    for name in ("mfma_peak", "stream_probe"):
        probes.append([HIPCC, "-x", "hip", os.path.join(HERE, "probes", name + ".hip"), "-o",
                       os.path.join(HERE, "probes", name)] + COMMON + DEVICE)
    with concurrent.futures.ThreadPoolExecutor(max_workers=jobs) as ex:
        for _ in ex.map(run, probes):
            pass
    if verbose:
        print("built", OUT)
    return OUT


if __name__ == "__main__":
    build(verbose=True)
