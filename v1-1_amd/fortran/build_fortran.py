"""Compile the iso_c_binding shim and the Fortran conformance program with
amdflang, linked against libsincdvr.so (rpath relative to the binary)."""
import os
import shutil
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
OUT = os.path.join(REPO, "tests", "fortran", "bin")
FLANG = os.environ.get("FLANG", "/opt/rocm/bin/amdflang")


def build():
    if not shutil.which(FLANG) and not os.path.exists(FLANG):
        return None
    os.makedirs(OUT, exist_ok=True)
    shim_o = os.path.join(OUT, "sincdvr_shim.o")
    subprocess.check_call([FLANG, "-O2", "-fPIC", "-c", os.path.join(HERE, "sincdvr_shim.f90"), "-o", shim_o,
                           "-module-dir", OUT])
    libdir = os.path.join(REPO, "v1-1_amd", "sincdvr")
    rel = os.path.relpath(libdir, OUT)
    exe = os.path.join(OUT, "conformance")
    subprocess.check_call([FLANG, "-O2", os.path.join(REPO, "tests", "fortran", "conformance.f90"), shim_o,
                           "-I", OUT, "-module-dir", OUT, "-L", libdir, "-lsincdvr",
                           "-Wl,-rpath,$ORIGIN/" + rel, "-o", exe])
    return exe


if __name__ == "__main__":
    print(build())
