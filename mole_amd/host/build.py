"""Builds the C++ host layer's test program (mole_amd/host/_build/host_tests)
against libmole_amd.so with g++; the headers under host/include are
header-only."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
OUT = os.path.join(HERE, "_build", "host_tests")


def build(verbose=False):
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    libdir = os.path.join(ROOT, "mole_amd")
    cmd = ["g++", "-O2", "-std=c++17", "-Wall", "-Wextra",
           "-I", os.path.join(HERE, "include"), "-I", os.path.join(ROOT, "include"),
           os.path.join(HERE, "tests", "host_tests.cpp"), "-o", OUT,
           "-L", libdir, "-lmole_amd", f"-Wl,-rpath,{libdir}", "-Wl,-rpath,$ORIGIN/../..",
           "-L/opt/rocm/lib", "-Wl,-rpath,/opt/rocm/lib", "-lamdhip64"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"g++ failed:\n{r.stderr[-6000:]}")
    if verbose:
        print(f"built {OUT}")
    return OUT


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
