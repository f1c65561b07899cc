"""CPU: the library's host-side sources under the sanitizers.

- The seeded generators (gen.cpp) and the .tbl ingest (tbl.cpp, the
  replacement for the reference's MappedFile and parse*ColumnChunk loaders,
  src/infra.h:59-171) run on the CPU only, so they are compiled here with g++
  and AddressSanitizer + UndefinedBehaviorSanitizer together with a driver that
  calls every generator at edge sizes and parses valid and malformed .tbl text.
- The serving scheduler (serve.cpp) is compiled unchanged against a fake
  executor (simulated streams and events, batch executors whose results encode
  their request) and driven by concurrent client threads: under ASan + UBSan
  with g++, and under ThreadSanitizer with the ROCm LLVM clang++ (whose runtime
  understands the condition-variable waits libstdc++ makes; g++ 11's does not
  and reports false double locks).

Any out-of-bounds access, leak, signed overflow, data race or other undefined
behaviour aborts a driver with a report. No GPU code is compiled here.
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "gpudbmetalbenchmark_amd", "csrc")
DRIVERS = os.path.join(ROOT, "tests", "host_sanitize")
ROCM_CLANG = "/opt/rocm/llvm/bin/clang++"
COMMON = ["-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer", "-D__HIP_PLATFORM_AMD__",
          "-I/opt/rocm/include", "-I" + os.path.join(ROOT, "include"), "-I" + CSRC]
ASAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all"]
ENV = dict(ASAN_OPTIONS="detect_leaks=1", UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1",
           TSAN_OPTIONS="halt_on_error=1:second_deadlock_stack=1")


def build_and_run(tmp_path, compiler, flags, sources, args=()):
    exe = str(tmp_path / "driver")
    b = subprocess.run([compiler] + COMMON + flags + sources + ["-lpthread", "-o", exe],
                       capture_output=True, text=True)
    assert b.returncode == 0, b.stderr[-4000:]
    r = subprocess.run([exe] + list(args), capture_output=True, text=True,
                       env=dict(os.environ, **ENV), timeout=300)
    report = r.stdout + r.stderr
    assert r.returncode == 0, report[-6000:]
    assert "all checks passed" in r.stdout
    for marker in ("runtime error", "AddressSanitizer", "LeakSanitizer", "ThreadSanitizer"):
        assert marker not in report, report[-6000:]
    return r.stdout


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not installed")
def test_generators_and_tbl_ingest_asan_ubsan(tmp_path):
    work = tmp_path / "files"
    work.mkdir()
    build_and_run(tmp_path, "g++", ASAN,
                  [os.path.join(CSRC, "tbl.cpp"), os.path.join(CSRC, "gen.cpp"),
                   os.path.join(DRIVERS, "sanitize_host.cpp")], [str(work)])


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not installed")
def test_serving_scheduler_fake_executor_asan_ubsan(tmp_path):
    out = build_and_run(tmp_path, "g++", ASAN,
                        [os.path.join(CSRC, "serve.cpp"), os.path.join(DRIVERS, "sanitize_serve.cpp")])
    print(out)


@pytest.mark.skipif(not os.path.exists(ROCM_CLANG), reason="ROCm clang++ not installed")
def test_serving_scheduler_fake_executor_tsan(tmp_path):
    out = build_and_run(tmp_path, ROCM_CLANG, ["-fsanitize=thread"],
                        [os.path.join(CSRC, "serve.cpp"), os.path.join(DRIVERS, "sanitize_serve.cpp")])
    print(out)
