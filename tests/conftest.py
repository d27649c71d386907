"""Shared test setup: import paths, markers, and a build when nothing is built."""
import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(REPO, "gpu-canonical-labeling_amd")
for p in (REPO, PKG, os.path.join(REPO, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(REPO, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")
    need = [os.path.join(PKG, "libgcl.so"), os.path.join(PKG, "gpu_canon_label"),
            os.path.join(REPO, "oracle", "liboracle.so")]
    if not all(os.path.exists(p) for p in need):
        import __graft_entry__
        __graft_entry__.build()


def gpu_available():
    import gcl
    return gcl.device_count() > 0


@pytest.fixture(scope="session")
def golden():
    import json
    with open(os.path.join(GOLDEN, "golden.json")) as f:
        return json.load(f)


def golden_graph(name):
    import gcl
    return gcl.read_dimacs(os.path.join(GOLDEN, "graphs", name + ".dimacs"))


def golden_output(entry):
    with open(os.path.join(GOLDEN, "expected", entry["output"]), "rb") as f:
        return f.read()
