"""The C ABI library loads and exports every symbol include/*.h declares
(no compute calls: this runs without a GPU)."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for h in ("th2d.h", "th2d_bag.h", "thirring_compat.h"):
        path = os.path.join(ROOT, "include", h)
        if not os.path.exists(path):
            continue
        txt = open(path).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        for m in re.finditer(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\(", txt):
            name = m.group(1)
            if name in ("if", "sizeof", "return", "defined"):
                continue
            # declarations: a return type precedes the name on the line
            line = txt[txt.rfind("\n", 0, m.start()) + 1:m.start()]
            if re.search(r"\b(int|void|double|float|char|const char|long)\b[\s*]*$", line.strip() + " ") or line.strip().endswith("*"):
                names.add(name)
    return sorted(names)


def test_library_exports_all_declared_symbols():
    from thirring2d_amd import _native
    lib = _native.lib()
    missing = [n for n in declared_symbols() if not hasattr(lib, n)]
    assert not missing, missing
    assert len(declared_symbols()) > 40


def test_python_binding_covers_header():
    from thirring2d_amd import _native
    hdr = set(n for n in declared_symbols() if n.startswith("th2d_"))
    assert hdr == set(_native.SIGNATURES), sorted(hdr ^ set(_native.SIGNATURES))
