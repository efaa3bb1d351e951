"""The C ABI library loads and exports every entry point include/eimp.h declares;
argument errors come back as status codes (no GPU needed)."""
import ctypes as C
import os
import re

import pytest

from eagleimp_amd import engine

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    src = open(os.path.join(ROOT, "include", "eimp.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(eimp_[a-z_0-9]+)\s*\(", src)))


def test_header_matches_binding_list():
    assert declared_symbols() == sorted(engine.EXPORTS)


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(engine.LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_error_paths_without_device():
    L = engine.lib()
    assert L.eimp_engine_initialize(None, 2, 8, 10, 10, 10, 64, None) == -1
    assert b"null" in L.eimp_last_error()
    assert L.eimp_engine_block_size(None, None) == -1
    assert L.eimp_engine_run_batch(None, 0, None, C.c_float(0.125), 0, None) == -4
    assert L.eimp_engine_destroy(None) == 0
    assert L.eimp_pbwt_variant_count() > 0
    nt, cpw, mk = C.c_int(), C.c_int(), C.c_int()
    assert L.eimp_pbwt_variant_info(0, C.byref(nt), C.byref(cpw), C.byref(mk)) == 0
    assert mk.value == nt.value // 64 * cpw.value * 64
    assert L.eimp_pbwt_variant_info(10_000, None, None, None) == -1


def test_variant_table_covers_config_k():
    vs = engine.pbwt_variants()
    assert max(v["max_k"] for v in vs) >= 16384
    assert any(v["max_k"] >= 10048 and v["threads"] == 512 for v in vs)


@pytest.mark.skipif(engine.device_count() > 0, reason="a device is visible")
def test_create_fails_cleanly_without_device():
    h = C.c_void_p()
    rc = engine.lib().eimp_engine_create(0, C.byref(h))
    assert rc != 0 and not h.value
    with pytest.raises(engine.EimpError):
        engine.Engine(0)
