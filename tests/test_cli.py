"""CPU: the gdb_bench CLI keeps the reference's argument grammar and exit
codes (src/main.cpp:48-72, :190-203): `help` exits 0; an unknown query, or a
scale word it does not know (taken as a query name, as the reference does),
exits 1 with "Unknown query" before any device is opened; the SF100 registry
has no primitives; a TPC-H query without a TPC-H database is refused."""
import os
import subprocess

import pytest

from gpudbmetalbenchmark_amd import build

EXE = build.DRIVER


def run(*args, cwd=None):
    if not os.path.exists(EXE):
        pytest.skip("gdb_bench not built")
    return subprocess.run([EXE, *args], capture_output=True, text=True, timeout=60, cwd=cwd)


def test_help_exits_zero():
    r = run("help")
    assert r.returncode == 0 and "sf100" in r.stdout


@pytest.mark.parametrize("args", [("q99",), ("sf1", "nosuch"), ("sf5",), ("q1", "sf5")])
def test_unknown_query_or_scale_exits_one(args):
    r = run(*args)
    assert r.returncode == 1
    assert "Unknown query" in r.stderr
    assert "Device:" not in r.stdout  # refused before the GPU is touched


def test_sf100_has_no_primitives():
    r = run("sf100", "selection")
    assert r.returncode == 1 and "Unknown query for sf100" in r.stderr


def test_tpch_query_needs_a_database(tmp_path):
    r = run("q14", cwd=str(tmp_path))  # no data/SF-1 here and no --data
    assert r.returncode == 1 and "needs a TPC-H database" in r.stderr
