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


def test_every_declared_entry_point_has_a_ctypes_signature(dm):
    """The package binds every header function with explicit argtypes (an
    unbound one would take Python floats / arrays without conversion)."""
    dm.lib()
    missing = [n for n in declared() if n not in dm.SIGNATURES]
    assert not missing, missing


def test_error_string_without_device(dm):
    lib = dm.lib()
    assert lib.dm_version() >= 100
    assert isinstance(lib.dm_last_error(), bytes)


def c_layout(tmp_path, struct, fields):
    """(sizeof, [offsetof ...]) of a header struct, from a program built with gcc."""
    src = tmp_path / "layout.c"
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "diffmom.h"', 'int main(void) {',
             f'  printf("%zu\\n", sizeof({struct}));']
    lines += [f'  printf("%zu\\n", offsetof({struct}, {f}));' for f in fields]
    lines += ['  return 0;', '}']
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    import subprocess
    subprocess.run(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = [int(v) for v in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()]
    return out[0], out[1:]


JL_TYPES = {"Cint": ctypes.c_int, "Cdouble": ctypes.c_double, "Ptr{Cvoid}": ctypes.c_void_p}


def julia_fields(struct):
    """(name, ctypes type) of a struct in the Julia shim, in declaration order."""
    src = open(os.path.join(ROOT, "diffmom.jl_amd", "julia", "DiffMoMHIP.jl")).read()
    body = re.search(r"^(?:mutable )?struct " + struct + r"\n(.*?)^end", src, flags=re.S | re.M).group(1)
    return [(n, JL_TYPES[t]) for n, t in re.findall(r"(\w+)::([\w{}]+)", body)]


def test_option_structs_match_the_header(dm, tmp_path):
    """The ctypes and Julia mirrors of dm_solve_opts / dm_solve_stats have the
    header's fields, in its order, at its offsets (a shorter mirror makes the
    library read past the caller's struct)."""
    for cls in (dm.dm_solve_opts, dm.dm_solve_stats):
        names = [f for f, _ in cls._fields_]
        size, offs = c_layout(tmp_path, cls.__name__, names)
        assert ctypes.sizeof(cls) == size
        assert [getattr(cls, f).offset for f in names] == offs
        jl = julia_fields(cls.__name__)
        assert [n for n, _ in jl] == names
        assert [t for _, t in jl] == [t for _, t in cls._fields_]


def _jl_signatures():
    import json
    return json.load(open(os.path.join(ROOT, "tests", "golden", "jl_signatures.json")))


def test_julia_shim_takes_the_reference_keywords():
    """Every keyword of the reference functions the shim replaces
    (tests/golden/jl_signatures.json, read from the reference's source by
    tests/golden/make_jl_signatures.py) is a keyword of the shim's method."""
    import importlib.util
    spec = importlib.util.spec_from_file_location("mk", os.path.join(ROOT, "tests", "golden", "make_jl_signatures.py"))
    mk = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mk)
    shim = open(os.path.join(ROOT, "diffmom.jl_amd", "julia", "DiffMoMHIP.jl")).read()
    for name, ref_kw in _jl_signatures().items():
        mine = mk.keywords(shim, name)
        missing = [k for k in ref_kw if k not in mine]
        assert not missing, (name, missing)


def test_python_mirror_takes_the_reference_keywords(dm):
    import inspect
    for name, ref_kw in _jl_signatures().items():
        params = inspect.signature(getattr(dm, name)).parameters
        missing = [k for k in ref_kw if k not in params]
        assert not missing, (name, missing)
