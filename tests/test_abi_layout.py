"""The ctypes mirrors of the C-ABI structs (untangling26_amd/_lib.py) must have
the exact layout of include/untangle.h: a shifted field hands the kernels a
wrong pointer.  A small C program compiled against the header prints sizeof
and offsetof of every field; the ctypes offsets must match."""
import ctypes as C
import pathlib
import subprocess

import pytest

from untangling26_amd import _lib

ROOT = pathlib.Path(__file__).resolve().parents[1]

PAIRS = {
    "ut_pcg_params": _lib.PcgParams,
    "ut_pcg_report": _lib.PcgReport,
    "ut_contact_args": _lib.ContactArgs,
    "ut_assemble_args": _lib.AssembleArgs,
    "ut_prp_params": _lib.PrpParams,
    "ut_solver_desc": _lib.SolverDesc,
    "ut_solver_params": _lib.SolverParams,
    "ut_step_report": _lib.StepReport,
    "ut_prp_choice": _lib.PrpChoice,
    "ut_prp_stats": _lib.PrpStats,
}


@pytest.fixture(scope="module")
def c_layout(tmp_path_factory):
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "untangle.h"', "int main(void) {"]
    for cname, cls in PAIRS.items():
        lines.append(f'printf("{cname} sizeof %zu\\n", sizeof({cname}));')
        for f in cls._fields_:
            lines.append(f'printf("{cname} {f[0]} %zu\\n", offsetof({cname}, {f[0]}));')
    lines += ["return 0;", "}"]
    d = tmp_path_factory.mktemp("abi")
    src, exe = d / "layout.c", d / "layout"
    src.write_text("\n".join(lines))
    subprocess.run(["gcc", "-std=c11", "-I", str(ROOT / "include"), str(src), "-o", str(exe)], check=True,
                   capture_output=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    table = {}
    for ln in out.splitlines():
        s, f, v = ln.split()
        table[(s, f)] = int(v)
    return table


@pytest.mark.parametrize("cname", sorted(PAIRS))
def test_struct_layout_matches_header(c_layout, cname):
    cls = PAIRS[cname]
    assert C.sizeof(cls) == c_layout[(cname, "sizeof")], cname
    for f in cls._fields_:
        assert getattr(cls, f[0]).offset == c_layout[(cname, f[0])], (cname, f[0])
    # every header field is mirrored: the mirror's last field ends where the C struct's padding starts
    last = cls._fields_[-1][0]
    assert getattr(cls, last).offset + getattr(cls, last).size <= C.sizeof(cls)
