"""The C-ABI library loads on a host without a GPU and exports every entry
point that include/music_hip.h declares (no compute call is made)."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "music-code_amd", "lib", "libmusic_hip.so")
HDR = os.path.join(ROOT, "include", "music_hip.h")


def declared():
    txt = open(HDR).read()
    txt = re.sub(r"/\*.*?\*/", " ", txt, flags=re.S)
    names = re.findall(r"\b(mh_[a-z0-9_]+)\s*\(", txt)
    return sorted(set(names))


def test_library_exports_every_declared_entry_point():
    lib = ctypes.CDLL(LIB)
    names = declared()
    assert len(names) > 40
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_fortran_module_binds_exported_names():
    """Every bind(C) name of the Fortran interface module exists in the library."""
    lib = ctypes.CDLL(LIB)
    mod = open(os.path.join(ROOT, "music-code_amd", "fortran", "music_hip_mod.f90")).read()
    names = sorted(set(re.findall(r'bind\(C,\s*name="(mh_[a-z0-9_]+)"\)', mod)))
    assert names
    assert [n for n in names if not hasattr(lib, n)] == []
