"""The C-ABI library loads and exports every entry point include/*.h
declares (no GPU calls), and the ctypes layouts match the header."""
import ctypes as C
import os
import re
import subprocess

import pytest

from luxion_amd import _ffi as F

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if not fn.endswith(".h"):
            continue
        txt = open(os.path.join(ROOT, "include", fn)).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        for m in re.finditer(r"^[a-z_][\w\s\*]*?\b(lx_\w+)\s*\(", txt, flags=re.M):
            names.add(m.group(1))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(F.LIB_PATH)
    names = declared_functions()
    assert len(names) >= 40
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_ffi_signature_table_covers_header():
    names = set(declared_functions())
    assert names == set(F.SIGNATURES), sorted(names ^ set(F.SIGNATURES))


def test_abi_version_and_clean_error_state():
    lib = F.load()
    assert lib.lx_abi_version() == 4
    assert lib.lx_last_error() == b""


def _header_structs():
    txt = open(os.path.join(ROOT, "include", "luxion_mi355x.h")).read()
    return sorted(set(re.findall(r"typedef struct (lx_\w+) \{", txt)))


def test_struct_layouts_match_the_compiled_header(tmp_path):
    """A C++ program compiled against include/luxion_mi355x.h prints sizeof
    and offsetof of every field of every struct the header defines; the
    ctypes mirror in luxion_amd/_ffi.py must agree field for field."""
    structs = _header_structs()
    assert len(structs) >= 15
    missing = [s for s in structs if not hasattr(F, s)]
    assert not missing, missing
    lines = ['#include "luxion_mi355x.h"', "#include <cstddef>", "#include <cstdio>", "int main() {"]
    for s in structs:
        lines.append(f'    std::printf("{s} size %zu\\n", sizeof({s}));')
        for fname, _ in getattr(F, s)._fields_:
            lines.append(f'    std::printf("{s} {fname} %zu\\n", offsetof({s}, {fname}));')
    lines += ["    return 0;", "}"]
    src = tmp_path / "layout.cpp"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run(["g++", "-std=c++17", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    checked = 0
    for line in filter(None, out):
        s, key, val = line.split()
        cls = getattr(F, s)
        want = C.sizeof(cls) if key == "size" else getattr(cls, key).offset
        assert int(val) == want, (s, key, int(val), want)
        checked += 1
    assert checked == sum(1 + len(getattr(F, s)._fields_) for s in structs)


def test_integration_example_is_the_compiled_consumer():
    """The C++ example in INTEGRATION.md is, line for line, the block that
    tests/capi_consumer.cpp compiles (and the GPU suite runs)."""
    doc = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    block = doc.split("```cpp", 1)[1].split("```", 1)[0].strip().split("\n")
    assert block[0] == '#include "luxion_mi355x.h"'
    snippet = "\n".join(block[1:]).strip()
    src = open(os.path.join(ROOT, "tests", "capi_consumer.cpp")).read()
    body = src.split("// INTEGRATION-BEGIN", 1)[1].split("// INTEGRATION-END", 1)[0].strip()
    assert snippet == body


def test_null_arguments_are_rejected_without_gpu():
    lib = F.load()
    assert lib.lx_scene_create(None, None, None) == 1
    assert b"null" in lib.lx_last_error()
    assert lib.lx_render_iterate(None, 1, 1) == 1


def test_tuning_json_errors_without_gpu(tmp_path):
    """lx_tuning_load validates before touching a context: a null context is
    rejected, as is a missing file (no device needed for either)."""
    from luxion_amd import _ffi as F
    lib = F.load()
    assert lib.lx_tuning_load(None, str(tmp_path / "missing.json").encode()) == 1  # LX_ERR_INVALID
    assert b"null" in lib.lx_last_error()
