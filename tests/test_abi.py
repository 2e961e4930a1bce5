"""The device library's C ABI: it loads (no GPU needed for that) and exports
exactly the entry points include/dfn.h declares; the Python mirror binds them all."""
import ctypes as C
import os
import re

import pytest

from demystification_dfn_amd import _abi
from demystification_dfn_amd.device import LIB_PATH

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    src = open(os.path.join(ROOT, "include", "dfn.h")).read()
    return sorted(set(re.findall(r"\bint\s+(dfn_\w+)\s*\(", src)))


def test_header_matches_python_mirror():
    names = [n[len("dfn_"):] for n in declared()]
    assert sorted(names) == sorted(_abi.SIGNATURES)


@pytest.mark.skipif(not os.path.exists(LIB_PATH), reason="device library not built")
def test_library_loads_and_exports_every_declared_symbol():
    lib = C.CDLL(LIB_PATH)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_problem_struct_layout_matches_header():
    # the ctypes mirror must have one field per struct member of dfn_problem
    src = open(os.path.join(ROOT, "include", "dfn.h")).read()
    body = src[src.index("typedef struct dfn_problem {"):src.index("} dfn_problem;")]
    count = 0
    for line in body.splitlines()[1:]:
        line = line.split("/*")[0].strip()
        if not line or not line.endswith(";"):
            continue
        decl = line.rstrip(";")
        decl = re.sub(r"^(const\s+)?(int32_t|int64_t|double)\s*", "", decl)
        count += len([d for d in decl.split(",") if d.strip()])
    assert count == len(_abi.DfnProblem._fields_)
