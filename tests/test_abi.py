"""The C ABI library loads and exports every symbol include/brille_amd.h declares (CPU only:
no compute calls)."""
import ctypes
import os
import re

from brille_amd import capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(ROOT, "include", "brille_amd.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(brille_[a-z0-9_]+)\s*\(", src)))


def test_header_and_binding_agree():
    assert declared_functions() == sorted(capi.SIGNATURES)


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(capi.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing


def test_abi_version_without_a_device():
    assert capi.lib().brille_abi_version() == 1


def test_error_text_is_reported():
    buf = ctypes.create_string_buffer(64)
    rc = capi.lib().brille_grid_create(None, None)
    assert rc == 1
    capi.lib().brille_last_error(buf, 64)
    assert b"null" in buf.value


def test_oracle_is_not_linked_into_the_product():
    """The product library never references the oracle's symbols."""
    data = open(capi.LIB_PATH, "rb").read()
    assert b"brille_oracle" not in data
