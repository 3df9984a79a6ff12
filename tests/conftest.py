import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "music-code_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLD = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


@pytest.fixture(scope="session")
def golden():
    import json
    with open(os.path.join(GOLD, "ref_vectors.json")) as f:
        return json.load(f)


@pytest.fixture(scope="session")
def hip():
    """The HIP library on cuda:0; fails loudly when it is missing."""
    from musichip import hip as H
    return H.Context(0)
