"""Build libmom5_gpu.so (HIP, gfx950) in-tree with hipcc.

    python -m mom5_amd.build            # build (always recompiles)

Flags: -ffp-contract=off keeps every FP64 multiply and add separate, so the
kernels round like the reference's arithmetic; packed FP32 instructions are
disabled (-packed-fp32-ops) as a matter of policy on this target.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libmom5_gpu.so")
SOURCES = ["mom_gpu_api.hip", "k_mdfl.hip", "k_quicker.hip", "k_column.hip", "k_barotropic.hip",
           "k_step.hip", "k_synth.hip", "k_advvel.hip", "k_btvariant.hip",
           "k_quick_step.hip", "k_noalloc.hip"]
ARCH = os.environ.get("MOM5_OFFLOAD_ARCH", "gfx950")


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found (set HIPCC)")


def flags():
    return ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
            "-ffp-contract=off", "-fno-fast-math",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-I", os.path.join(ROOT, "include"), "-I", CSRC,
            "-Wall", "-Wno-unused-function"]


def build(verbose=False, jobs=None):
    cc = hipcc()
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    objs = []
    procs = []
    for src in SOURCES:
        obj = os.path.join(objdir, src.replace(".hip", ".o"))
        cmd = [cc, *flags(), "-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
        objs.append(obj)
    failed = []
    for src, p in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            failed.append((src, out.decode(errors="replace")))
        elif verbose and out:
            print(out.decode(errors="replace"))
    if failed:
        msg = "\n".join(f"--- {s}\n{o}" for s, o in failed)
        raise RuntimeError(f"hipcc failed:\n{msg}")
    tmp = LIB + ".tmp"
    cmd = [cc, "-shared", f"--offload-arch={ARCH}", "-o", tmp, *objs]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


FORTRAN_DIR = os.path.join(HERE, "fortran")
DRIVER = os.path.join(HERE, "ocean_gpu_driver")


def flang():
    for cand in (os.environ.get("FLANG"), "/opt/rocm/bin/amdflang", shutil.which("flang")):
        if cand and os.path.exists(cand):
            return cand
    return None


def build_fortran(verbose=False):
    """Compile the ISO_C_BINDING module and the Fortran host driver with
    flang (-fdefault-real-8 -fdefault-double-8 as MOM is built) and link them
    against libmom5_gpu.so (found next to the executable at run time)."""
    fc = flang()
    if fc is None:
        raise RuntimeError("flang not found (set FLANG)")
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    flags = ["-O2", "-fdefault-real-8", "-fdefault-double-8", "-module-dir", objdir, "-I", objdir]
    mod_o = os.path.join(objdir, "mom_gpu_mod.o")
    drv_o = os.path.join(objdir, "ocean_gpu_driver.o")
    for src, obj in (("mom_gpu_mod.F90", mod_o), ("ocean_gpu_driver.F90", drv_o)):
        cmd = [fc, *flags, "-c", os.path.join(FORTRAN_DIR, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
    tmp = DRIVER + ".tmp"
    cmd = [fc, drv_o, mod_o, "-o", tmp, "-L", HERE, "-lmom5_gpu", "-Wl,-rpath,$ORIGIN"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, DRIVER)
    return DRIVER


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print(LIB)
    print(build_fortran(verbose="-v" in sys.argv))
