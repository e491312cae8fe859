"""The C-ABI library loads without a GPU and exports every function that include/*.h declares
(no compute calls)."""
import ctypes as C
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DECL = re.compile(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_]*\s*\*?\s*(lkd?_[A-Za-z0-9_]+)\s*\(", re.M)


def declared():
    names = set()
    inc = os.path.join(ROOT, "include")
    for f in sorted(os.listdir(inc)):
        if f.endswith(".h"):
            text = open(os.path.join(inc, f)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            names.update(DECL.findall(text))
    return names


def test_headers_declare_the_entry_points():
    names = declared()
    for must in ("lk_kkt_create", "lk_kkt_factor", "lk_kkt_solve", "lk_ipm_solve", "lk_arena_create",
                 "lk_verify_optimal", "lkd_chol_step", "lkd_trsv_level", "lkd_assemble_pairs"):
        assert must in names
    assert len(names) > 60


def test_library_exports_every_declared_function(native):
    L = native.lib()
    missing = [n for n in sorted(declared()) if not hasattr(L, n)]
    assert not missing, f"declared but not exported: {missing}"


def test_version_and_error_string(native):
    L = native.lib()
    L.lk_version.restype = C.c_int
    assert L.lk_version() > 0
    assert isinstance(L.lk_last_error(), bytes)
