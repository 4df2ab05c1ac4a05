"""CPU tier: the C-ABI library builds, loads and exports every entry point
that include/itg.h declares (no compute calls: there is no GPU here)."""
import os
import re

import pytest

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "itg.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b(itg_[a-z0-9_]+)\s*\(", text)
    return sorted(set(names))


def test_header_parses_and_lists_core_entry_points():
    names = declared_functions()
    for n in ["itg_init", "itg_contract", "itg_permute", "itg_truncate", "itg_eigh_trunc",
              "itg_svd_trunc", "itg_qr", "itg_heff2_apply", "itg_mlp_forward"]:
        assert n in names


def test_library_exports_every_declared_symbol():
    from itg import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libitg.so not built (run __graft_entry__.build())")
    L, missing = _lib.load_only()
    assert not missing
    for n in declared_functions():
        assert hasattr(L, n), n


def test_binding_covers_header():
    from itg import _lib
    assert set(declared_functions()) == set(_lib.EXPORTED)
