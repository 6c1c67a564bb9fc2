"""The C ABI: libgearhip.so loads without a GPU and exports every function
include/*.h declares, and the ctypes table binds each of them."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DECL = re.compile(r"^\s*(?:[A-Za-z_][\w ]*?[\s*]+)(gh_\w+)\s*\(", re.M)


def _declared():
    names = set()
    inc = os.path.join(ROOT, "include")
    for f in sorted(os.listdir(inc)):
        if not f.endswith(".h"):
            continue
        text = open(os.path.join(inc, f)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        if "GH_MATH_FN" in text:      # header-only inline functions, not exports
            continue
        names.update(m.group(1) for m in DECL.finditer(text) if "typedef" not in m.group(0))
    return names


def _lib():
    path = os.path.join(ROOT, "gear_amd", "libgearhip.so")
    if not os.path.exists(path):
        pytest.skip("library not built")
    return ctypes.CDLL(path)


def test_every_declared_symbol_is_exported():
    names = _declared()
    assert len(names) >= 15, names
    h = _lib()
    missing = [n for n in sorted(names) if not hasattr(h, n)]
    assert not missing, missing


def test_binding_table_covers_the_header():
    from gear_amd import _lib as L
    assert _declared() <= set(L.SIGNATURES), sorted(_declared() - set(L.SIGNATURES))


def test_host_only_calls():
    h = _lib()
    assert h.gh_version() >= 100
    h.gh_qcqp_pack_bytes.restype = ctypes.c_size_t
    assert h.gh_qcqp_pack_bytes(50, 21) > 0
    assert h.gh_qcqp_pack_bytes(65, 21) == 0        # unsupported size
    h.gh_dag_workspace_bytes.restype = ctypes.c_size_t
    assert h.gh_dag_workspace_bytes(None, 4) == 0
    assert h.gh_dag_plan_destroy(None) == 0
