"""Builds the host checks of tests/host with g++ (no GPU, no hipcc): the
plain library librh.so and librh_asan.so with AddressSanitizer and
UndefinedBehaviorSanitizer.  Both compile the package's own host/device
headers (csrc/region_view.hpp), so they run the kernels' per-record code."""
from __future__ import annotations

import pathlib
import subprocess

HERE = pathlib.Path(__file__).resolve().parent
CSRC = HERE.parents[1] / "untangling26_amd" / "csrc"
SRC = HERE / "region_host.cpp"
BASE = ["g++", "-std=c++17", "-g", "-fPIC", "-shared", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
        f"-I{CSRC}", "-ffp-contract=off"]
VARIANTS = {
    "librh.so": ["-O2"],
    "librh_asan.so": ["-O1", "-fno-omit-frame-pointer", "-fsanitize=address,undefined"],
}


def _deps():
    return [SRC, CSRC / "region_view.hpp", CSRC / "contact_rec.hpp", CSRC / "ut_math.hpp"]


def build(force: bool = False) -> dict:
    out = {}
    newest = max(p.stat().st_mtime for p in _deps())
    for name, flags in VARIANTS.items():
        lib = HERE / name
        if force or not lib.exists() or lib.stat().st_mtime < newest:
            cmd = BASE + flags + [str(SRC), "-o", str(lib)]
            r = subprocess.run(cmd, capture_output=True, text=True)
            if r.returncode != 0:
                raise RuntimeError(f"g++ failed for {name}:\n{r.stderr}")
        out[name] = lib
    return out


def asan_runtime() -> str:
    return subprocess.run(["g++", "-print-file-name=libasan.so"], capture_output=True, text=True,
                          check=True).stdout.strip()


if __name__ == "__main__":
    for k, v in build(force=True).items():
        print(k, v)
