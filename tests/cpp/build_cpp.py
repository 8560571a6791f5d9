"""Builds the C++ facade programs of tests/cpp with g++ against the in-tree
libuntangle_hip.so (run path relative to the binary, so the build travels with
the tree).  No GPU is needed to build."""
from __future__ import annotations

import pathlib
import subprocess

HERE = pathlib.Path(__file__).resolve().parent
ROOT = HERE.parents[1]
PKG = ROOT / "untangling26_amd"
PROGRAMS = ["facade_two_sheets"]


def build(force: bool = False) -> dict:
    out = {}
    lib = PKG / "libuntangle_hip.so"
    for name in PROGRAMS:
        src, exe = HERE / f"{name}.cpp", HERE / name
        deps = [src, ROOT / "include" / "untangle.hpp", ROOT / "include" / "untangle.h"]
        if force or not exe.exists() or exe.stat().st_mtime < max(p.stat().st_mtime for p in deps):
            cmd = ["g++", "-std=c++17", "-O2", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
                   f"-I{ROOT / 'include'}", str(src), "-o", str(exe), f"-L{PKG}", "-luntangle_hip",
                   "-L/opt/rocm/lib", "-lamdhip64", "-Wl,-rpath,$ORIGIN/../../untangling26_amd",
                   "-Wl,-rpath,/opt/rocm/lib"]
            r = subprocess.run(cmd, capture_output=True, text=True)
            if r.returncode != 0:
                raise RuntimeError(f"g++ failed for {name}:\n{r.stderr}")
        out[name] = exe
    return out


if __name__ == "__main__":
    for k, v in build(force=True).items():
        print(k, v)
