"""The C-ABI library loads without a GPU and exports every entry point that
include/fpsi.h declares (no compute calls)."""
import os
import re

from flowpsi_amd import capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    txt = open(os.path.join(ROOT, "include", "fpsi.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(fpsi_[a-z_]+)\s*\(", txt)))


def test_header_declares_entry_points():
    names = declared()
    assert "fpsi_create" in names and "fpsi_newton_step" in names


def test_library_exports_every_declared_symbol():
    L = capi.load()
    missing = [n for n in declared() if not hasattr(L, n)]
    assert not missing, missing


def test_python_binding_lists_every_symbol():
    assert sorted(capi.EXPORTS) == declared()


def test_struct_layouts_match_header():
    """ctypes mirrors of the ABI structs carry the header's fields in order."""
    txt = open(os.path.join(ROOT, "include", "fpsi.h")).read()
    body = re.search(r"typedef struct \{([^{}]*)\} fpsi_params;", txt).group(1)
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    fields = []
    for decl in body.split(";"):
        decl = decl.strip()
        if not decl:
            continue
        names = decl.split(None, 1)[1]
        for nm in names.split(","):
            fields.append(re.sub(r"\[.*\]", "", nm).strip())
    assert fields == [f[0] for f in capi.Params._fields_]
