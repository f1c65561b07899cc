"""Build recipe for libgdb.so (the engine) and, for tests, the CPU oracle.

    python -m gpudbmetalbenchmark_amd.build          # engine + oracle + driver
    python -m gpudbmetalbenchmark_amd.build --engine # engine only
    ... --incremental                                # only changed sources

The engine is every csrc/*.hip and csrc/*.cpp compiled by hipcc for gfx950 and
linked into gpudbmetalbenchmark_amd/libgdb.so next to this file, so it travels
with the source tree. Object files go to _build/ at the repo root.
"""
import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
INCLUDE = os.path.join(ROOT, "include")
BUILD = os.path.join(ROOT, "_build")
LIB = os.path.join(PKG, "libgdb.so")
DRIVER = os.path.join(PKG, "gdb_bench")
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_LIB = os.path.join(ORACLE_DIR, "liboracle.so")

ARCH = os.environ.get("GDB_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")
# Packed-FP32 instructions are disabled for every HIP translation unit.
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-Wall",
             "-Wno-unused-function", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             f"-I{INCLUDE}", f"-I{CSRC}"]
LINK_FLAGS = ["-shared", "-fPIC", f"--offload-arch={ARCH}", "-lpthread"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("build step failed:\n" + " ".join(cmd) + "\n" + r.stdout + r.stderr)
    return r


def _sources(exts):
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC)
                  if os.path.splitext(f)[1] in exts and not f.startswith("driver"))


def _headers_mtime():
    hs = [os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".h")]
    hs += [os.path.join(INCLUDE, f) for f in os.listdir(INCLUDE) if f.endswith(".h")]
    return max(os.path.getmtime(h) for h in hs)


def build_engine(jobs=8, incremental=False):
    """Compiles every source (incremental=True: only sources newer than their
    object or than any header) and links libgdb.so."""
    os.makedirs(BUILD, exist_ok=True)
    srcs = _sources((".hip", ".cpp"))
    hdr = _headers_mtime() if incremental else None
    objs = []
    cmds = []
    for s in srcs:
        o = os.path.join(BUILD, os.path.basename(s) + ".o")
        objs.append(o)
        if incremental and os.path.exists(o) and \
                os.path.getmtime(o) > max(os.path.getmtime(s), hdr):
            continue
        lang = ["-x", "hip"] if s.endswith(".hip") else []
        cmds.append([HIPCC] + HIP_FLAGS + lang + ["-c", s, "-o", o])
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(_run, cmds))
    tmp = LIB + ".tmp"
    _run([HIPCC] + LINK_FLAGS + objs + ["-o", tmp])
    os.replace(tmp, LIB)
    return LIB


def build_driver():
    src = os.path.join(CSRC, "driver.cpp")
    if not os.path.exists(src):
        return None
    srcs = sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC)
                  if f.startswith("driver") and f.endswith(".cpp"))
    _run(["g++", "-O2", "-std=c++17", f"-I{INCLUDE}"] + srcs + ["-o", DRIVER, f"-L{PKG}", "-lgdb",
          "-Wl,-rpath,$ORIGIN", "-lpthread"])
    return DRIVER


def build_oracle():
    _run(["make", "-s", "-C", ORACLE_DIR, "CC=gcc"])
    return ORACLE_LIB


def build_all(jobs=8, incremental=False):
    out = [build_engine(jobs, incremental), build_oracle()]
    d = build_driver()
    if d:
        out.append(d)
    return out


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--engine", action="store_true")
    ap.add_argument("-j", type=int, default=8)
    ap.add_argument("--incremental", action="store_true",
                    help="recompile only sources changed since their object (or any header)")
    a = ap.parse_args()
    if a.engine:
        print(build_engine(a.j, a.incremental))
    else:
        for p in build_all(a.j, a.incremental):
            print(p)
    sys.exit(0)
