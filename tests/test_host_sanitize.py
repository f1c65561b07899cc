"""CPU: the library's host-side sources under AddressSanitizer and
UndefinedBehaviorSanitizer.

The seeded generators (gen.cpp) and the .tbl ingest (tbl.cpp, the replacement
for the reference's MappedFile and parse*ColumnChunk loaders,
src/infra.h:59-171) run on the CPU only, so they are compiled here with g++
and both sanitizers (no GPU code on that line) together with a driver that
calls every generator at edge sizes and parses valid and malformed .tbl text.
Any out-of-bounds access, leak, signed overflow or other undefined behaviour
aborts the driver with a report.
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "gpudbmetalbenchmark_amd", "csrc")


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not installed")
def test_generators_and_tbl_ingest_are_sanitizer_clean(tmp_path):
    exe = str(tmp_path / "sanitize_host")
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined",
           "-fno-sanitize-recover=all", "-fno-omit-frame-pointer", "-D__HIP_PLATFORM_AMD__",
           "-I/opt/rocm/include", "-I" + os.path.join(ROOT, "include"), "-I" + CSRC,
           os.path.join(CSRC, "tbl.cpp"), os.path.join(CSRC, "gen.cpp"),
           os.path.join(ROOT, "tests", "host_sanitize", "sanitize_host.cpp"),
           "-lpthread", "-o", exe]
    b = subprocess.run(cmd, capture_output=True, text=True)
    assert b.returncode == 0, b.stderr[-4000:]
    work = tmp_path / "files"
    work.mkdir()
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = subprocess.run([exe, str(work)], capture_output=True, text=True, env=env, timeout=300)
    report = r.stdout + r.stderr
    assert r.returncode == 0, report[-4000:]
    assert "all checks passed" in r.stdout
    assert "runtime error" not in report and "AddressSanitizer" not in report, report[-4000:]
