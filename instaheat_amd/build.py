"""Build recipe of the native parts, in tree.

* ``instaheat_amd/libihg.so``: the HIP kernels and the C ABI of
  ``include/ihg.h`` (hipcc, gfx950).
* ``instaheat_amd/libinstaheat.so``: the C99 host (steppers, mk_rhs shim) of
  ``include/instaheat.h``, linked against libihg.
* ``instaheat_amd/libinstaheat_h5.so``: the HDF5 output file of
  ``include/instaheat_h5.h``, built when HDF5's headers and library are
  found (``IH_HDF5_ROOT``, default ``/opt/conda``).
* ``oracle/liboracle.so``: the CPU oracle (test infrastructure).

Run ``python -m instaheat_amd.build`` or ``__graft_entry__.build()``.
"""
import concurrent.futures
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
HOST = os.path.join(HERE, "host")
BUILD = os.path.join(HERE, "_build")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("IHG_OFFLOAD_ARCH", "gfx950")
HIP_FLAGS = [
    "-std=c++17", "-O3", "-fPIC", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result", "-I", os.path.join(ROOT, "include"),
]
CC = os.environ.get("CC", "gcc")
C_FLAGS = ["-std=c99", "-O2", "-fPIC", "-Wall", "-Wextra",
           "-I", os.path.join(ROOT, "include")]

HIP_SOURCES = ["ihg_api.hip", "ihg_fft.hip", "ihg_rhs.hip", "ihg_obs.hip",
               "ihg_step.hip", "ihg_mlp.hip", "ihg_graph.hip", "ihg_ic.hip",
               "ihg_group.hip"]
HOST_SOURCES = ["instaheat.c", "compat.c"]
HDF5_ROOT = os.environ.get("IH_HDF5_ROOT", "/opt/conda")


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                       text=True)
    out = "\n".join(l for l in r.stdout.splitlines()
                    if "packed-fp32-ops" not in l and "is not a recognized feature" not in l)
    if r.returncode != 0:
        raise RuntimeError("command failed:\n" + " ".join(cmd) + "\n" + out)
    return out


def build_hip(jobs=8):
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    cmds = []
    for s in HIP_SOURCES:
        o = os.path.join(BUILD, s.replace(".hip", ".o"))
        objs.append(o)
        cmds.append([HIPCC] + HIP_FLAGS + ["-c", os.path.join(CSRC, s), "-o", o])
    with concurrent.futures.ThreadPoolExecutor(max_workers=jobs) as ex:
        for out in ex.map(_run, cmds):
            if out.strip():
                print(out)
    lib = os.path.join(HERE, "libihg.so")
    tmp = lib + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + objs)
    os.replace(tmp, lib)
    return lib


def build_host():
    srcs = [os.path.join(HOST, s) for s in HOST_SOURCES]
    lib = os.path.join(HERE, "libinstaheat.so")
    tmp = lib + ".tmp"
    _run([CC] + C_FLAGS + ["-shared"] + srcs +
         ["-L", HERE, "-lihg", "-Wl,-rpath,$ORIGIN", "-lm", "-o", tmp])
    os.replace(tmp, lib)
    return lib


def hdf5_found():
    return (os.path.exists(os.path.join(HDF5_ROOT, "include", "hdf5.h")) and
            os.path.exists(os.path.join(HDF5_ROOT, "lib", "libhdf5.so")))


def build_h5():
    """libinstaheat_h5.so, or None when HDF5 is not installed"""
    if not hdf5_found():
        print(f"HDF5 not found under {HDF5_ROOT}: libinstaheat_h5.so skipped")
        return None
    inc = os.path.join(HDF5_ROOT, "include")
    libdir = os.path.join(HDF5_ROOT, "lib")
    lib = os.path.join(HERE, "libinstaheat_h5.so")
    tmp = lib + ".tmp"
    _run([CC] + C_FLAGS + ["-I", inc, "-shared", os.path.join(HOST, "ih_h5.c"),
                           "-L", HERE, "-linstaheat", "-lihg", "-L", libdir,
                           "-lhdf5", "-Wl,-rpath,$ORIGIN",
                           f"-Wl,-rpath,{libdir}", "-o", tmp])
    os.replace(tmp, lib)
    return lib


def build_oracle():
    _run(["make", "-s", "-B", "-C", os.path.join(ROOT, "oracle")])
    return os.path.join(ROOT, "oracle", "liboracle.so")


def build_all(oracle=True):
    libs = [build_hip(), build_host()]
    h5 = build_h5()
    if h5:
        libs.append(h5)
    if oracle:
        libs.append(build_oracle())
    return libs


if __name__ == "__main__":
    for p in build_all(oracle="--no-oracle" not in sys.argv):
        print("built", os.path.relpath(p, ROOT))
