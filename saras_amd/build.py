"""Build of the native parts: the gfx950 HIP library behind include/saras_mg.h and the C++ host
drop-in driver.  Every call recompiles (hipcc cross-compiles without a GPU)."""
import concurrent.futures as cf
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "saras_amd")
CSRC = os.path.join(PKG, "csrc")
HOST = os.path.join(PKG, "host")
OUT = os.path.join(PKG, "_build")
LIB = os.path.join(OUT, "libsaras_mg.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    # packed FP32 instructions stay off in every HIP translation unit
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", "-Wno-unknown-warning-option",
]
HIP_SOURCES = ["mg_kernels.hip", "mg_api.hip", "mg_solve.hip", "ts_kernels.hip", "ts_api.hip", "unet_kernels.hip", "unet_api.hip", "guess_solver.hip", "field_api.hip", "probes.hip"]
CPP_SOURCES = ["grid.cpp"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed: " + " ".join(cmd) + "\n" + r.stdout + r.stderr)
    return r


def _obj(src):
    return os.path.join(OUT, os.path.splitext(os.path.basename(src))[0] + ".o")


def build_lib():
    os.makedirs(OUT, exist_ok=True)
    jobs = []
    for s in HIP_SOURCES:
        jobs.append([HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, s), "-o", _obj(s)])
    for s in CPP_SOURCES:
        jobs.append(["g++", "-O2", "-std=c++17", "-fPIC", "-c", os.path.join(CSRC, s), "-o", _obj(s)])
    with cf.ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        list(ex.map(_run, jobs))
    objs = [_obj(s) for s in HIP_SOURCES + CPP_SOURCES]
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    return LIB


def build_host():
    """C++ drop-in driver over the C ABI (saras_amd/host)."""
    srcs = [os.path.join(HOST, f) for f in sorted(os.listdir(HOST)) if f.endswith(".cc")] if os.path.isdir(HOST) else []
    if not srcs:
        return None
    exe = os.path.join(OUT, "saras_mg_driver")
    _run(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "include"), "-I", HOST, *srcs,
          "-L", OUT, "-lsaras_mg", "-Wl,-rpath,$ORIGIN", "-o", exe])
    return exe


IO_LIB = os.path.join(OUT, "libsaras_io.so")
HDF5_ROOT = os.environ.get("HDF5_ROOT", "/opt/conda")


def build_io():
    """Restart files (include/saras_restart.h): host C++ over the image's serial HDF5."""
    src = os.path.join(PKG, "io", "restart.cc")
    if not os.path.exists(os.path.join(HDF5_ROOT, "include", "hdf5.h")):
        raise RuntimeError(f"hdf5.h not found under {HDF5_ROOT}/include (set HDF5_ROOT)")
    os.makedirs(OUT, exist_ok=True)
    tmp = IO_LIB + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-I", os.path.join(HDF5_ROOT, "include"), src,
          "-L", os.path.join(HDF5_ROOT, "lib"), "-lhdf5", f"-Wl,-rpath,{os.path.join(HDF5_ROOT, 'lib')}", "-o", tmp])
    os.replace(tmp, IO_LIB)
    return IO_LIB


def build_oracle():
    subprocess.run(["make", "-s", "-B", "-C", os.path.join(ROOT, "oracle")], check=True)


def build_all():
    build_lib()
    build_host()
    build_io()
    build_oracle()


if __name__ == "__main__":
    build_all()
    print("built", LIB)
