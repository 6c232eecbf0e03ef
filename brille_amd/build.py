"""Build the native parts of brille_amd in-tree.

* ``libbrille_amd.so``: the HIP kernels and the C ABI (hipcc, gfx950 only).
* ``_brille*.so``: the pybind11 module with the host-side constructors
  (g++), linked against ``libbrille_amd.so`` with an ``$ORIGIN`` rpath.

Usage: ``python brille_amd/build.py`` (the package refuses to import before
its native library exists, so the script runs by path).
"""
import os
import subprocess
import sys
import sysconfig
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
HOST = os.path.join(CSRC, "host")
OBJ = os.path.join(HERE, "_build")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# gfx950 only; FP64 decisions must not be contracted into FMAs; packed FP32
# ops are disabled for all HIP code.
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(ROOT, "include"),
]
HIP_SOURCES = ["capi.hip", "run.hip", "binned.hip", "batch.hip", "proposal.hip", "noalloc.hip", "sort.hip"]
HOST_SOURCES = ["symmetry.cpp", "lattice.cpp", "polyhedron.cpp", "bz.cpp", "trellis.cpp", "grid.cpp",
                "pybind_module.cpp"]


def _run(cmd):
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in p.stdout.splitlines()
                    if "packed-fp32-ops' is not a recognized feature" not in l)
    if p.returncode != 0:
        raise RuntimeError("command failed: " + " ".join(cmd) + "\n" + out)
    return out


def _ext_suffix():
    return sysconfig.get_config_var("EXT_SUFFIX") or ".so"


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    jobs = []
    hip_objs = []
    for src in HIP_SOURCES:
        path = os.path.join(CSRC, src)
        if not os.path.exists(path):
            continue
        obj = os.path.join(OBJ, src + ".o")
        hip_objs.append(obj)
        jobs.append([HIPCC] + HIP_FLAGS + ["-c", path, "-o", obj])
    import pybind11
    py_inc = sysconfig.get_paths()["include"]
    host_flags = ["-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fvisibility=hidden",
                  "-I", os.path.join(ROOT, "include"), "-I", pybind11.get_include(), "-I", py_inc]
    host_objs = []
    for src in HOST_SOURCES:
        obj = os.path.join(OBJ, src + ".o")
        host_objs.append(obj)
        jobs.append(["g++"] + host_flags + ["-c", os.path.join(HOST, src), "-o", obj])
    workers = min(len(jobs), int(os.environ.get("MAX_JOBS", "8")))
    with ThreadPoolExecutor(max_workers=max(workers, 1)) as ex:
        outs = list(ex.map(_run, jobs))
    if verbose:
        for o in outs:
            if o.strip():
                print(o)
    lib = os.path.join(HERE, "libbrille_amd.so")
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", lib] + hip_objs)
    mod = os.path.join(HERE, "_brille" + _ext_suffix())
    _run(["g++", "-shared", "-fPIC", "-o", mod] + host_objs +
         ["-L", HERE, "-lbrille_amd", "-Wl,-rpath,$ORIGIN"])
    return lib, mod


if __name__ == "__main__":
    for p in build(verbose=True):
        print("built", p)
    sys.exit(0)
