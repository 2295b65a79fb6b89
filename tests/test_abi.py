"""The C ABI libraries load and export every symbol include/*.h declares
(no compute calls: this runs without a GPU).

include/th2d.h and include/th2d_bag.h are libth2d.so; include/thirring_hmc.h
and include/thirring_bag.h (the reference's own signatures) are
libthirring_hmc.so and libthirring_bag.so, which also carry the
reference's globals."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "thirring2d_amd")

LIBS = {
    "th2d.h": "libth2d.so",
    "th2d_bag.h": "libth2d.so",
    "thirring_hmc.h": "libthirring_hmc.so",
    "thirring_bag.h": "libthirring_bag.so",
}

TYPES = r"\b(int|void|double|float|char|long|complex)\b"


def _strip(txt):
    return re.sub(r"/\*.*?\*/", "", txt, flags=re.S)


def declared_functions(header):
    txt = _strip(open(os.path.join(ROOT, "include", header)).read())
    names = set()
    for m in re.finditer(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\(", txt):
        name = m.group(1)
        if name in ("if", "sizeof", "return", "defined"):
            continue
        line = txt[txt.rfind("\n", 0, m.start()) + 1:m.start()].strip()
        if re.search(TYPES + r"[\s*]*$", line + " ") or line.endswith("*"):
            names.add(name)
    return names


def declared_variables(header):
    txt = _strip(open(os.path.join(ROOT, "include", header)).read())
    names = set()
    for m in re.finditer(r"^extern\s+[A-Za-z_ ]+?\s+([^;(]+);", txt, flags=re.M):
        for part in m.group(1).split(","):
            names.add(part.strip().lstrip("*").split("[")[0].strip())
    return names


def declared_symbols():
    return sorted(set().union(*(declared_functions(h) for h in ("th2d.h", "th2d_bag.h"))))


@pytest.mark.parametrize("header", sorted(LIBS))
def test_library_exports_all_declared_symbols(header):
    path = os.path.join(PKG, LIBS[header])
    if LIBS[header] == "libth2d.so":
        from thirring2d_amd import _native
        lib = _native.lib()
    else:
        from thirring2d_amd import _native
        _native.lib()  # the dependency, loaded from the same directory
        lib = ctypes.CDLL(path)
    funcs = declared_functions(header)
    missing = [n for n in sorted(funcs) if not hasattr(lib, n)]
    assert not missing, missing
    for v in declared_variables(header):
        ctypes.c_int.in_dll(lib, v)  # raises if the symbol is absent
    assert len(funcs) > 10


def test_python_binding_covers_header():
    from thirring2d_amd import _native
    hdr = set(n for n in declared_symbols() if n.startswith("th2d_"))
    assert hdr == set(_native.SIGNATURES), sorted(hdr ^ set(_native.SIGNATURES))


def test_reference_signatures_present():
    """The compatibility headers declare the entry points the reference's
    callers use (hmc.c:73-746; Thirring.h:84-107, fermionbag.c:49-62)."""
    hmc = declared_functions("thirring_hmc.h")
    for n in ("fm_mul", "fm_conjugate_mul", "fermion_matrix", "determinant", "fmdm_invert_cg",
              "random_pseudofermion", "stochastic_vector", "pseudofermion_action",
              "stochastic_md_action", "random_momentum", "momentum_step", "gauge_step",
              "calc_gauge_action", "update_gauge"):
        assert n in hmc, n
    bag = declared_functions("thirring_bag.h")
    for n in ("fM_index", "fM", "fM_transpose", "cg_propagator", "calc_Dinv",
              "update_background", "determinant", "determinant_mu", "det_added_link",
              "det_removed_link", "det_moved_monomer", "new_link", "new_monomer", "removed_link",
              "removed_monomer", "moved_source_monomer", "not_moved_source_monomer"):
        assert n in bag, n
