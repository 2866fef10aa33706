"""Builds tests/cpp/dropin_check.cpp against include/tdag.hpp and libtdag.so (g++, C++20)."""
import os
import subprocess

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBDIR = os.path.join(REPO, "i3d2025-voxel-transform-dag_amd", "tdag_amd")


def build(out_dir: str) -> str:
    exe = os.path.join(out_dir, "dropin_check")
    cmd = ["g++", "-std=c++20", "-O1", "-Wall", "-Werror", "-I", os.path.join(REPO, "include"), os.path.join(REPO, "tests", "cpp", "dropin_check.cpp"),
           "-o", exe, "-L", LIBDIR, "-ltdag", f"-Wl,-rpath,{LIBDIR}"]
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError(proc.stderr)
    return exe
