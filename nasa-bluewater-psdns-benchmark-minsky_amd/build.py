"""Build recipe of the gfx950 library and the Fortran driver.

    python nasa-bluewater-psdns-benchmark-minsky_amd/build.py [-v]

Produces, inside this package directory:
    lib/libpsdns_gpu.so      C ABI (include/psdns_gpu.h), hipcc --offload-arch=gfx950
    lib/psdns_gpu_mod.o/.mod ISO_C_BINDING interface module (amdflang)
    bin/psdns_driver         Fortran RK2 driver linked against libpsdns_gpu.so
Always recompiles (no timestamp skipping), so a call after a source change
is always current.
"""
import concurrent.futures as cf
import os
import pathlib
import shutil
import subprocess
import sys

PKG = pathlib.Path(__file__).resolve().parent
REPO = PKG.parent
CSRC = PKG / "csrc"
FSRC = PKG / "fortran"
LIB = PKG / "lib"
BIN = PKG / "bin"
INCLUDE = REPO / "include"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FC = os.environ.get("FC", "/opt/rocm/bin/amdflang")
ARCH = "gfx950"
HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    # packed FP32 VALU forms are kept out of the generated code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{INCLUDE}",
]
HIP_SOURCES = ["psdns_kernels.hip", "psdns_api.hip", "sgs_kernels.hip", "sgs_api.hip"]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(str(c) for c in cmd), flush=True)
    r = subprocess.run([str(c) for c in cmd], capture_output=True, text=True)
    out = "\n".join(l for l in (r.stdout + r.stderr).splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(map(str, cmd))}\n{out}")
    if verbose and out.strip():
        print(out)


def build_library(verbose=False, defines=(), suffix=""):
    """lib/libpsdns_gpu.so, or with compile-time defines (tuning variants,
    e.g. PSDNS_XMFMA=1) lib/libpsdns_gpu_<suffix>.so for PSDNS_GPU_LIB"""
    obj = PKG / ("_build" + ("_" + suffix if suffix else ""))
    obj.mkdir(exist_ok=True)
    LIB.mkdir(exist_ok=True)
    flags = HIP_FLAGS + [f"-D{d}" for d in defines]
    objs = [obj / (s.rsplit(".", 1)[0] + ".o") for s in HIP_SOURCES]
    with cf.ThreadPoolExecutor(max_workers=len(HIP_SOURCES)) as ex:
        futs = [ex.submit(_run, [HIPCC, *flags, "-c", CSRC / s, "-o", o], verbose) for s, o in zip(HIP_SOURCES, objs)]
        for f in futs:
            f.result()
    name = "libpsdns_gpu" + ("_" + suffix if suffix else "") + ".so"
    tmp = LIB / (name + ".tmp")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp], verbose)
    _check_defined(tmp)
    os.replace(tmp, LIB / name)
    return LIB / name


def _check_defined(so):
    """A kernel whose host-side stub was not emitted leaves an undefined
    psdns:: symbol that only fails when the library is loaded: refuse it."""
    nm = shutil.which("nm") or "/usr/bin/nm"
    r = subprocess.run([nm, "-D", "--undefined-only", str(so)], capture_output=True, text=True)
    bad = [l.split()[-1] for l in r.stdout.splitlines() if "5psdns" in l]
    if bad:
        raise RuntimeError(f"{so}: undefined library symbols (kernel stubs not emitted): {bad[:3]}")


def build_fortran(verbose=False):
    if not pathlib.Path(FC).exists():
        if verbose:
            print(f"{FC} not found: Fortran driver skipped")
        return None
    BIN.mkdir(exist_ok=True)
    mod_src = FSRC / "psdns_gpu_mod.f90"
    drv_src = FSRC / "psdns_driver.f90"
    _run([FC, "-O2", "-c", mod_src, "-J", LIB, "-o", LIB / "psdns_gpu_mod.o"], verbose)
    exe = BIN / "psdns_driver"
    _run([FC, "-O2", "-I", LIB, drv_src, LIB / "psdns_gpu_mod.o", f"-L{LIB}", "-lpsdns_gpu",
          f"-Wl,-rpath,{LIB}", "-Wl,-rpath,$ORIGIN/../lib", "-o", exe], verbose)
    return exe


def build(verbose=False):
    so = build_library(verbose)
    exe = build_fortran(verbose)
    return so, exe


if __name__ == "__main__":
    args = sys.argv[1:]
    if "--define" in args:
        # a tuning variant of the library only: --define NAME=VALUE [--define ...] --suffix S
        defs = [args[i + 1] for i, a in enumerate(args) if a == "--define"]
        suffix = args[args.index("--suffix") + 1]
        print(build_library("-v" in args, defs, suffix))
    else:
        so, exe = build(verbose="-v" in args)
        print(so)
        if exe:
            print(exe)
