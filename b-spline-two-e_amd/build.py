"""Build recipe for the MI355X R^k builder.

    python b-spline-two-e_amd/build.py          # HIP library + Fortran shim

Produces, in-tree (git-ignored, shipped with the tree to the GPU host):
    b-spline-two-e_amd/lib/libbs2e.so           HIP kernels + C ABI (gfx950)
    b-spline-two-e_amd/lib/libbs2e_fortran.so   Fortran drop-in modules (amdflang)
    b-spline-two-e_amd/lib/bs2e_basis_setup     Fortran driver over the drop-in
    b-spline-two-e_amd/lib/bs2e_rkh_multipoles  Fortran RKH multipole step over the drop-in
"""
from __future__ import annotations

import concurrent.futures as cf
import glob
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
FSRC = os.path.join(HERE, "fortran")
LIB = os.path.join(HERE, "lib")
OBJ = os.path.join(LIB, "obj")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FC = os.environ.get("BS2E_FC", "/opt/rocm/bin/amdflang")
ARCH = "gfx950"
HIP_FLAGS = [
    "-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}",
    # packed FP32 instructions stay disabled for every kernel
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]
# Per-file additions.  The separated-row block fill keeps its FP64 matrix-core
# accumulators in ordinary vector registers (otherwise the compiler moves
# them between accumulation and vector registers around every step).
FILE_FLAGS = {"k_channel.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form"]}


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(cmd)}\n{out}")
    return out


def build_hip(verbose: bool = False) -> str:
    os.makedirs(OBJ, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")))
    objs = [os.path.join(OBJ, os.path.basename(s)[:-4] + ".o") for s in srcs]
    jobs = min(len(srcs), int(os.environ.get("MAX_JOBS", "8")))
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        futs = [ex.submit(_run, [HIPCC, *HIP_FLAGS, *FILE_FLAGS.get(os.path.basename(s), []), "-I", CSRC, "-I",
                                 os.path.join(ROOT, "include"), "-c", s, "-o", o])
                for s, o in zip(srcs, objs)]
        for f in futs:
            out = f.result()
            if verbose and out.strip():
                print(out)
    so = os.path.join(LIB, "libbs2e.so")
    tmp = so + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", *objs, "-o", tmp])
    os.replace(tmp, so)
    return so


def build_fortran(verbose: bool = False) -> str | None:
    """Compile the Fortran drop-in modules against the C ABI (needs amdflang)."""
    srcs = [os.path.join(FSRC, f) for f in ("bs2e_capi.f90", "bs2e_mat_els.f90", "bs2e_basis_setup.f90")]
    if not all(os.path.exists(s) for s in srcs) or shutil.which(FC) is None:
        return None
    mod = os.path.join(OBJ, "fmod")
    os.makedirs(mod, exist_ok=True)
    flags = ["-fdefault-integer-8", "-O2", "-fPIC", "-module-dir", mod, "-I", mod]
    rpath = "-Wl,-rpath,$ORIGIN"          # the .so and the driver sit beside libbs2e.so
    so = os.path.join(LIB, "libbs2e_fortran.so")
    _run([FC, *flags, "-shared", srcs[0], srcs[1], "-o", so, f"-L{LIB}", "-lbs2e", rpath])
    exe = os.path.join(LIB, "bs2e_basis_setup")
    _run([FC, *flags, srcs[2], "-o", exe, f"-L{LIB}", "-lbs2e_fortran", "-lbs2e", rpath])
    rkh = os.path.join(FSRC, "bs2e_rkh_multipoles.f90")
    _run([FC, *flags, rkh, "-o", os.path.join(LIB, "bs2e_rkh_multipoles"), f"-L{LIB}", "-lbs2e_fortran", "-lbs2e",
          rpath])
    return so


def build(verbose: bool = False) -> None:
    build_hip(verbose)
    build_fortran(verbose)


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print("built", os.path.join(LIB, "libbs2e.so"))
