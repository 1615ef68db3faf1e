"""The C-ABI library loads and exports every entry point include/*.h
declares (no GPU calls), and the ctypes layouts match the header."""
import ctypes as C
import os
import re

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


def test_struct_sizes_match_c_layout():
    # sizes from the header's field lists (x86-64 SysV alignment)
    assert C.sizeof(F.lx_tex_transform) == 20
    assert C.sizeof(F.lx_material_desc) == 4 * (3 + 1 + 5 + 2 + 1 + 5 + 1 + 5 + 3 + 1 + 1 + 5 + 5 + 1)
    assert C.sizeof(F.lx_camera_desc) == 4 * (2 + 12 + 2)
    assert C.sizeof(F.lx_ray) == 32
    assert C.sizeof(F.lx_hit) == 20
    assert C.sizeof(F.lx_instance_desc) == 56


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
