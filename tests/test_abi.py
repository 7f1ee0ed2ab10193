"""The C-ABI library loads and exports every entry point that include/cpfem.h declares
(no compute call; no GPU needed)."""
import os
import re

import pytest

from conftest import ROOT


def _declared():
    src = open(os.path.join(ROOT, "include", "cpfem.h")).read()
    return sorted(set(re.findall(r"\b(cpfem_[a-z0-9_]+)\s*\(", src)))


def test_header_and_binding_agree():
    from cpfem_amd import _native
    assert sorted(_native.EXPORTED) == _declared()


def test_library_exports_every_symbol():
    from cpfem_amd import _native
    if not os.path.exists(_native.LIB_PATH):
        pytest.skip("library not built (run __graft_entry__.build())")
    L = _native.load_library()
    for name in _declared():
        assert hasattr(L, name), name
    assert L.cpfem_version() == 1
