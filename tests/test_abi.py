"""The C-ABI library loads and exports every entry point include/diffmom.h declares
(no compute calls: this runs without a GPU)."""
import ctypes
import os
import re

from conftest import ROOT


def declared():
    src = open(os.path.join(ROOT, "include", "diffmom.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(dm_[a-z0-9_]+)\s*\(", src)))


def test_header_declares_entry_points():
    names = declared()
    for must in ("dm_create", "dm_build_rwg", "dm_rwg_create", "dm_assemble_z", "dm_apply", "dm_mf_apply",
                 "dm_excitation_planewave", "dm_solve", "dm_radiation_vectors"):
        assert must in names


def test_library_exports_every_declared_symbol(dm):
    lib = ctypes.CDLL(dm.LIB_PATH)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_error_string_without_device(dm):
    lib = dm.lib()
    assert lib.dm_version() >= 100
    assert isinstance(lib.dm_last_error(), bytes)
