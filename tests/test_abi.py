"""CPU: the C ABI library loads, exports every entry point include/gdb.h
declares, and reports errors without a GPU. No kernel is launched here."""
import ctypes as C
import os
import re

import pytest

from gpudbmetalbenchmark_amd import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    txt = open(os.path.join(ROOT, "include", "gdb.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(gdb_[a-z0-9_]+)\s*\(", txt)))


def test_header_declares_the_pipelines():
    names = declared_functions()
    for n in ("gdb_open", "gdb_close", "gdb_table_create", "gdb_q6", "gdb_q1", "gdb_q3",
              "gdb_q9", "gdb_q13", "gdb_select_lt", "gdb_sum", "gdb_join_count"):
        assert n in names


def test_library_exports_every_declared_symbol():
    lib = abi.load()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_prototypes_cover_header():
    assert set(declared_functions()) <= set(abi.PROTOTYPES), \
        sorted(set(declared_functions()) - set(abi.PROTOTYPES))


def test_strerror():
    lib = abi.load()
    assert lib.gdb_strerror(0) == b"ok"
    assert lib.gdb_strerror(abi.GDB_E_ADMISSION).startswith(b"request exceeds")
    assert lib.gdb_strerror(-999) == b"unknown error"


def test_open_without_gpu_fails_cleanly():
    lib = abi.load()
    ctx = C.c_void_p()
    rc = lib.gdb_open(0, 0, C.byref(ctx))
    if rc == abi.GDB_OK:  # a GPU is visible: this test is about the CPU-only case
        lib.gdb_close(ctx)
        pytest.skip("GPU present")
    assert rc == abi.GDB_E_NODEVICE
    assert b"no HIP device" in lib.gdb_last_error()
    assert lib.gdb_open(0, 0, None) == abi.GDB_E_INVALID


def test_null_arguments_rejected():
    lib = abi.load()
    assert lib.gdb_q6(None, None, None, None, None) == abi.GDB_E_INVALID
    assert lib.gdb_q1(None, None, None, None, None) == abi.GDB_E_INVALID
    lib.gdb_close(None)
    lib.gdb_table_free(None)


def test_calendar_and_names():
    import datetime
    lib = abi.load()
    d0 = datetime.date(1992, 1, 1)
    for day in range(0, 3000, 11):
        assert lib.gdb_year_of_day(day) == (d0 + datetime.timedelta(days=day)).year
    assert lib.gdb_nation_name(18) == b"CHINA"
    assert lib.gdb_segment_name(1) == b"BUILDING"
