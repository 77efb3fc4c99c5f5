"""Build recipe for libcsl_hip.so (gfx950) — compiled in-tree with hipcc.

The shared library lands next to this file so it travels with the source tree.
Run ``python -m classical_mcmc_spin_amd.build`` or call :func:`build`.
Objects compile in parallel; the persistent kernel is compiled once per
sites-per-thread value (-DCSL_SPT=n).
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
ROOT = os.path.dirname(HERE)
LIB_NAME = "libcsl_hip.so"
LIB_PATH = os.path.join(HERE, LIB_NAME)

# (source, object suffix, extra flags)
HIP_UNITS = [("csl_kernels.hip", "", []), ("csl_capi.hip", "", []), ("csl_mixed.hip", "", []), ("csl_adjoint.hip", "", []),
             ("csl_dssf.hip", "", []), ("csl_spectrum.hip", "", []), ("csl_quench.hip", "", []),
             ("csl_dopri5.hip", "", []), ("csl_partition.hip", "", []), ("csl_mixed_persist.hip", "", [])] + [
    ("csl_persistent.hip", f"_spt{n}", [f"-DCSL_SPT={n}"]) for n in (1, 2, 3, 4, 6)]
CXX_UNITS = ["gpu_api_shim.cpp", "mixed_gpu_api_shim.cpp"]

HIPCC_FLAGS = [
    "-O3",
    "--offload-arch=gfx950",
    "-std=c++17",
    "-fPIC",
    # packed FP32 ops are kept out of every HIP object
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def hipcc() -> str:
    path = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(path):
        raise RuntimeError("hipcc not found: the HIP toolchain is required to build libcsl_hip.so")
    return path


def build(verbose: bool = False, jobs: int = 0) -> str:
    """Compile every HIP/C++ source into classical_mcmc_spin_amd/libcsl_hip.so."""
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    inc = ["-I", os.path.join(ROOT, "include"), "-I", CSRC]
    cmds, objs = [], []
    for src, suffix, extra in HIP_UNITS:
        obj = os.path.join(objdir, src + suffix + ".o")
        cmds.append([hipcc(), *HIPCC_FLAGS, *extra, *inc, "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    for src in CXX_UNITS:
        obj = os.path.join(objdir, src + ".o")
        cmds.append([hipcc(), "-O2", "-std=c++17", "-fPIC", *inc, "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    jobs = jobs or min(len(cmds), max(1, min(16, os.cpu_count() or 1)))
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(lambda c: _run(c, verbose), cmds))
    tmp = LIB_PATH + ".tmp"
    _run([hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-L/opt/rocm/lib", "-lhipfft", "-Wl,-rpath,/opt/rocm/lib", "-o", tmp], verbose)
    os.replace(tmp, LIB_PATH)
    build_cxx_test(verbose)
    build_h5(verbose)
    return LIB_PATH


# HDF5 C library locations probed for the writer library (prefix, include, lib)
HDF5_PREFIXES = [("/opt/conda", "/opt/conda/include", "/opt/conda/lib"),
                 ("/usr", "/usr/include/hdf5/serial", "/usr/lib/x86_64-linux-gnu/hdf5/serial"),
                 ("/usr", "/usr/include", "/usr/lib/x86_64-linux-gnu")]


def find_hdf5():
    for _, inc, libdir in HDF5_PREFIXES:
        if os.path.exists(os.path.join(inc, "hdf5.h")) and os.path.exists(os.path.join(libdir, "libhdf5.so")):
            return inc, libdir
    return None


def build_h5(verbose: bool = False) -> str:
    """libcsl_h5.so: the HDF5 writers (csrc/csl_h5.cpp) against the HDF5 C
    library; skipped with a message when no HDF5 installation is found."""
    found = find_hdf5()
    if found is None:
        sys.stderr.write("HDF5 C library not found: libcsl_h5.so (HDF5 output) not built\n")
        return ""
    inc, libdir = found
    out = os.path.join(HERE, "libcsl_h5.so")
    tmp = out + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-I", os.path.join(ROOT, "include"), "-I", inc,
          os.path.join(CSRC, "csl_h5.cpp"), "-o", tmp, f"-L{libdir}", "-lhdf5", f"-Wl,-rpath,{libdir}"], verbose)
    os.replace(tmp, out)
    return out


CXX_TESTS = ["test_gpu_api_shim", "test_mixed_gpu_api_shim"]


def build_cxx_test(verbose: bool = False) -> str:
    """C++ programs that drive the gpu:: and mixed_gpu:: drop-in APIs
    (tests/cxx), linked against the in-tree library with an rpath to it.
    The forwarding include directory is on the path, so the reference's
    include names (classical_spin/gpu/...) resolve to the drop-in headers."""
    built = ""
    for name in CXX_TESTS:
        src = os.path.join(ROOT, "tests", "cxx", name + ".cpp")
        exe = os.path.join(ROOT, "tests", "cxx", name)
        if not os.path.exists(src):
            continue
        cmd = ["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "include"),
               "-I", os.path.join(ROOT, "include", "classical_spin_amd_forward"), src, "-o", exe,
               "-L", HERE, "-lcsl_hip", f"-Wl,-rpath,{HERE}", "-Wl,-rpath,$ORIGIN/../../classical_mcmc_spin_amd"]
        _run(cmd, verbose)
        built = exe
    return built


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"build step failed: {' '.join(cmd[:3])} ... ({res.returncode})")
    if verbose:
        msg = "\n".join(l for l in (res.stdout + res.stderr).splitlines() if "packed-fp32-ops" not in l)
        if msg.strip():
            sys.stderr.write(msg + "\n")


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
