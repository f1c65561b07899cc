"""CPU: the C ABI library loads, exports every entry point include/gdb.h
declares, and reports errors without a GPU. No kernel is launched here."""
import ctypes as C
import os
import re

import pytest

from gpudbmetalbenchmark_amd import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


HEADERS = sorted(f for f in os.listdir(os.path.join(ROOT, "include")) if f.endswith(".h"))


def declared_functions():
    names = set()
    for h in HEADERS:
        txt = open(os.path.join(ROOT, "include", h)).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        names |= set(re.findall(r"\b(gdb_[a-z0-9_]+)\s*\(", txt))
    return sorted(names)


def test_header_declares_the_pipelines():
    names = declared_functions()
    for n in ("gdb_open", "gdb_close", "gdb_table_create", "gdb_q6", "gdb_q1", "gdb_q3",
              "gdb_q9", "gdb_q13", "gdb_select_lt", "gdb_sum", "gdb_join_count", "gdb_q2",
              "gdb_q4", "gdb_q5", "gdb_q7", "gdb_q8", "gdb_q10", "gdb_q11", "gdb_q12", "gdb_q14",
              "gdb_q15", "gdb_q16", "gdb_q17", "gdb_q18", "gdb_q19", "gdb_q20", "gdb_q21",
              "gdb_q22"):
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


# C struct name -> ctypes class, for the layout check below
LAYOUTS = {
    "gdb_tpch_db": "TpchDb", "gdb_q2_params": "Q2Params", "gdb_q2_row": "Q2Row",
    "gdb_q2_result": "Q2Result", "gdb_q4_params": "Q4Params", "gdb_q4_result": "Q4Result",
    "gdb_q5_params": "Q5Params", "gdb_q5_result": "Q5Result", "gdb_q7_params": "Q7Params",
    "gdb_q7_result": "Q7Result", "gdb_q8_params": "Q8Params", "gdb_q8_result": "Q8Result",
    "gdb_q10_params": "Q10Params", "gdb_q10_row": "Q10Row", "gdb_q10_result": "Q10Result",
    "gdb_q11_params": "Q11Params", "gdb_q11_row": "Q11Row", "gdb_q12_params": "Q12Params",
    "gdb_q12_result": "Q12Result", "gdb_q14_params": "Q14Params", "gdb_q14_result": "Q14Result",
    "gdb_q15_params": "Q15Params", "gdb_q15_result": "Q15Result", "gdb_q16_params": "Q16Params",
    "gdb_q16_row": "Q16Row", "gdb_q17_params": "Q17Params", "gdb_q17_result": "Q17Result",
    "gdb_q18_params": "Q18Params", "gdb_q18_row": "Q18Row", "gdb_q18_result": "Q18Result",
    "gdb_q19_block": "Q19Block", "gdb_q19_params": "Q19Params", "gdb_q19_result": "Q19Result",
    "gdb_q20_params": "Q20Params", "gdb_q21_params": "Q21Params", "gdb_q21_row": "Q21Row",
    "gdb_q21_result": "Q21Result", "gdb_q22_params": "Q22Params", "gdb_q22_row": "Q22Row",
    "gdb_q22_result": "Q22Result", "gdb_timing": "Timing", "gdb_q3_result": "Q3Result",
    "gdb_q1_result": "Q1Result", "gdb_q9_result": "Q9Result", "gdb_q13_result": "Q13Result",
}


def test_struct_layouts_match_ctypes(tmp_path):
    """sizeof and every field offset of the header structs, from the C compiler,
    equal the ctypes declarations the Python host side passes across the ABI."""
    import shutil
    import subprocess
    if not shutil.which("gcc"):
        pytest.skip("no gcc")
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "gdb_tpch.h"',
             "int main(void) {"]
    for cname, py in LAYOUTS.items():
        cls = getattr(abi, py)
        lines.append(f'printf("{cname} __sizeof__ %zu\\n", sizeof({cname}));')
        for f, _ in cls._fields_:
            lines.append(f'printf("{cname} {f} %zu\\n", offsetof({cname}, {f}));')
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)],
                   check=True)
    got = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    for line in filter(None, got):
        cname, field, val = line.split()
        cls = getattr(abi, LAYOUTS[cname])
        exp = C.sizeof(cls) if field == "__sizeof__" else getattr(cls, field).offset
        assert int(val) == exp, (cname, field, int(val), exp)


def test_set_kernel_timing_null_context():
    lib = abi.load()
    assert lib.gdb_set_kernel_timing(None, 0) == abi.GDB_E_INVALID
