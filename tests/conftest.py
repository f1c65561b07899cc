"""Shared fixtures. GPU tests are marked `gpu` and run only on a machine with
an MI355X; everything else runs on the CPU."""
import json
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests", "golden")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (run with -m gpu)")


@pytest.fixture(scope="session")
def golden():
    with open(os.path.join(ROOT, "tests", "golden", "golden.json")) as f:
        return json.load(f)


@pytest.fixture(scope="session")
def flat(golden):
    from gpudbmetalbenchmark_amd import generate
    return generate.lineitem_flat(golden["flat"]["seed"], golden["flat"]["n"], receipt=True)


@pytest.fixture(scope="session")
def joinset(golden):
    from gpudbmetalbenchmark_amd import generate
    j = golden["join"]
    return generate.join_set(j["seed"], j["customers"], j["orders"], j["parts"], j["suppliers"])


@pytest.fixture(scope="session")
def engine():
    from gpudbmetalbenchmark_amd import Engine
    e = Engine(0)
    yield e
    e.close()
