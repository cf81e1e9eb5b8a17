"""Build libhft.so (HIP, gfx950) in-tree with hipcc.

Usage: python -m hipft_amd.build   (or __graft_entry__.build())
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "libhft.so")
SOURCES = [os.path.join(HERE, "csrc", "hft_lib.hip")]

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    # keep the reference's rounding: no fused multiply-add contraction
    "-ffp-contract=off",
    # no packed FP32 instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def build_lib(verbose=False):
    tmp = LIB + ".tmp"
    cmd = [hipcc()] + HIPCC_FLAGS + ["-I", os.path.join(ROOT, "include"), "-o", tmp] + SOURCES
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc failed building libhft.so")
    os.replace(tmp, LIB)
    return LIB


FORTRAN_DRIVER = os.path.join(HERE, "fortran", "hft_run")
# the reference's executable over the C ABI: reads hipft.in, writes histories
FORTRAN_HIPFT = os.path.join(HERE, "fortran", "hft_hipft")


def amdflang():
    for cand in (os.environ.get("HFT_FC"), "/opt/rocm/lib/llvm/bin/amdflang",
                 shutil.which("amdflang")):
        if cand and os.path.exists(cand):
            return cand
    return None


def build_fortran(verbose=False):
    """The Fortran drivers (include/hft.f90 + fortran/hft_run.f90, and
    fortran/hft_hipft.f90) linked to libhft.so; skipped when no Fortran
    compiler is installed."""
    fc = amdflang()
    if fc is None:
        if verbose:
            print("no Fortran compiler: Fortran drivers not built", flush=True)
        return None
    import tempfile
    for src, exe in (("hft_run.f90", FORTRAN_DRIVER), ("hft_hipft.f90", FORTRAN_HIPFT)):
        with tempfile.TemporaryDirectory() as mod_dir:
            cmd = [fc, "-O2", "-J", mod_dir, os.path.join(ROOT, "include", "hft.f90"),
                   os.path.join(HERE, "fortran", src), "-L", HERE, "-lhft",
                   "-Wl,-rpath,$ORIGIN/..", "-o", exe]
            if verbose:
                print(" ".join(cmd), flush=True)
            res = subprocess.run(cmd, capture_output=True, text=True, cwd=mod_dir)
        if res.returncode != 0:
            sys.stderr.write(res.stdout + res.stderr)
            raise RuntimeError(f"Fortran driver build failed ({src})")
    return FORTRAN_DRIVER


def build_oracle(verbose=False):
    """Compile the CPU oracle (test infrastructure) with its Makefile."""
    odir = os.path.join(ROOT, "oracle")
    res = subprocess.run(["make", "-C", odir, "-B", "liboracle.so"], capture_output=True, text=True)
    if verbose or res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
    if res.returncode != 0:
        raise RuntimeError("oracle build failed")


if __name__ == "__main__":
    build_lib(verbose=True)
    build_fortran(verbose=True)
    build_oracle(verbose=True)
    print("built", LIB)
