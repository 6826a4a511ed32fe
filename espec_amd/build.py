"""Build recipe for libespec_hip.so (gfx950) and the Fortran shim.

``python -m espec_amd.build`` compiles every HIP/C++ source under
espec_amd/csrc into espec_amd/libespec_hip.so with hipcc for gfx950, and the
ISO_C_BINDING shim under espec_amd/fortran into espec_amd/libespec_f.so with
flang.  Builds are in-tree so the shared objects travel with the source tree.
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
FSRC = os.path.join(HERE, "fortran")
LIB = os.path.join(HERE, "libespec_hip.so")
FLIB = os.path.join(HERE, "libespec_f.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLANG = os.environ.get("FLANG", "/opt/rocm/bin/amdflang")
ARCH = "gfx950"

HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "-shared", f"--offload-arch={ARCH}",
    # Packed FP32 VALU forms are kept out of every kernel.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result",
]


def sources():
    out = []
    for f in sorted(os.listdir(CSRC)):
        if f.endswith((".hip", ".cpp")):
            out.append(os.path.join(CSRC, f))
    return out


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {cmd[0]} (exit {r.returncode})")
    return r


def build_hip(verbose=False):
    objdir = os.path.join(ROOT, "_build", "obj")
    os.makedirs(objdir, exist_ok=True)
    objs = []
    for src in sources():
        obj = os.path.join(objdir, os.path.basename(src) + ".o")
        lang = ["-x", "hip"] if src.endswith(".hip") else []
        cflags = [f for f in HIP_FLAGS if f != "-shared"]
        _run([HIPCC] + cflags + ["-c"] + lang + [src, "-o", obj, "-I", os.path.join(ROOT, "include")],
             verbose)
        objs.append(obj)
    tmp = LIB + ".tmp"
    # rocSOLVER (with rocBLAS) serves the one-off dense eigensolve of MTRXDIAG
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + objs
         + ["-L/opt/rocm/lib", "-lrocsolver", "-lrocblas", "-Wl,-rpath,/opt/rocm/lib"], verbose)
    os.replace(tmp, LIB)
    return LIB


def build_fortran(verbose=False):
    srcs = [os.path.join(FSRC, f) for f in sorted(os.listdir(FSRC)) if f.endswith(".f90")] \
        if os.path.isdir(FSRC) else []
    if not srcs:
        return None
    if not os.path.exists(FLANG):
        raise RuntimeError("flang not found; the Fortran shim needs it")
    moddir = os.path.join(ROOT, "_build", "mod")
    os.makedirs(moddir, exist_ok=True)
    tmp = FLIB + ".tmp"
    _run([FLANG, "-O2", "-fPIC", "-shared", "-module-dir", moddir, "-o", tmp] + srcs
         + [f"-L{HERE}", "-lespec_hip", "-Wl,-rpath,$ORIGIN"], verbose)
    os.replace(tmp, FLIB)
    # The test host program that calls the drop-ins as eSPec's main program does.
    drv = os.path.join(ROOT, "tests", "fortran", "dropin_driver.f90")
    if os.path.exists(drv):
        exe = os.path.join(ROOT, "tests", "fortran", "dropin_driver.x")
        _run([FLANG, "-O2", "-module-dir", moddir, "-o", exe + ".tmp", drv, f"-L{HERE}",
              "-lespec_f", "-lespec_hip", f"-Wl,-rpath,$ORIGIN/../../espec_amd"], verbose)
        os.replace(exe + ".tmp", exe)
    return FLIB


def build_oracle(verbose=False):
    """The checkers (test infrastructure): the C oracle always, the reference
    build only where the reference sources are present."""
    make = shutil.which("make") or "make"
    _run([make, "-C", os.path.join(ROOT, "oracle"), "oracle"], verbose)
    if os.path.isdir("/root/reference/src"):
        _run([make, "-C", os.path.join(ROOT, "oracle"), "-j8", "ref"], verbose)


def build(verbose=False):
    build_hip(verbose)
    build_fortran(verbose)
    build_oracle(verbose)


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print("built", LIB)
