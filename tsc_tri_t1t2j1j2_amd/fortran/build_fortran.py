"""Compile the Fortran binding module and its check program with flang
(when a Fortran compiler is present), linking libtsc_hotpath.so."""
import os
import shutil
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)


def build():
    fc = shutil.which("flang") or shutil.which("amdflang") or "/opt/rocm/llvm/bin/flang"
    if not os.path.exists(fc) and not shutil.which(fc):
        print("flang not found: Fortran binding not built")
        return None
    link = ["-L", PKG, "-ltsc_hotpath", "-Wl,-rpath," + PKG]
    out = os.path.join(HERE, "test_hotpath")
    cmd = [fc, "-O2", "-J", HERE, os.path.join(HERE, "tsc_hotpath_mod.f90"), os.path.join(HERE, "test_hotpath.f90"),
           "-o", out] + link
    subprocess.run(cmd, check=True, cwd=HERE)
    # the drop-in subroutines with the reference's signatures, and their host program
    out2 = os.path.join(HERE, "test_dmrg_boundary")
    cmd = [fc, "-O2", "-J", HERE, os.path.join(HERE, "tsc_hotpath_mod.f90"), os.path.join(HERE, "tsc_dmrg_mod.f90"),
           os.path.join(HERE, "test_dmrg_boundary.f90"), "-o", out2] + link
    subprocess.run(cmd, check=True, cwd=HERE)
    return out


if __name__ == "__main__":
    build()
