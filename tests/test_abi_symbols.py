"""The C-ABI library loads without a GPU and exports every function include/exgame.h
declares (no compute call is made here)."""
import ctypes
import os
import re

from exgame_amd import _lib
from exgame_amd import build as libbuild

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "exgame.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(exg_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_boundary():
    names = declared_functions()
    for must in ("exg_world_create", "exg_world_step", "exg_store_create", "exg_slot_insert", "exg_dispatch",
                 "exg_gravity_accumulate", "exg_last_error"):
        assert must in names


def test_library_exports_every_declared_symbol():
    if not os.path.exists(_lib.LIB_PATH):
        libbuild.build()
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_oracle_is_not_linked_into_the_product():
    # the product library must not carry the oracle's symbols
    if not os.path.exists(_lib.LIB_PATH):
        libbuild.build()
    lib = ctypes.CDLL(_lib.LIB_PATH)
    assert not hasattr(lib, "orc_world_new")
