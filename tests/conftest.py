import json
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "gromacs-fmm_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def golden():
    with open(os.path.join(ROOT, "tests", "golden", "spc27_refdata.json")) as fh:
        return json.load(fh)


def announce(msg):
    """Names the case before a GPU step, flushed, so a failure can be attributed."""
    print(f"[case] {msg}", flush=True)
