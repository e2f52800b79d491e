"""Shared fixtures.  Tests marked `gpu` need a HIP device (MI355X) and call
the product through its C ABI (libsc23ad.so); the others run on the CPU:
the oracle against the golden fixtures, the host-side setup, the generators
and the ABI surface."""
from __future__ import annotations

import json
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "sc23-ad_amd"), os.path.join(ROOT, "oracle"),
          os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")


@pytest.fixture(scope="session")
def golden():
    with open(os.path.join(ROOT, "tests", "golden", "reference.json")) as f:
        return json.load(f)
