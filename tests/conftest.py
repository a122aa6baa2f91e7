import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")
    config.addinivalue_line("markers", "ref: needs the reference build in oracle/_ref")


def pytest_collection_modifyitems(config, items):
    from oracle import oracle as ora
    if not ora.ref_available():
        skip = pytest.mark.skip(reason="oracle/_ref not built (make -C oracle ref)")
        for it in items:
            if "ref" in it.keywords:
                it.add_marker(skip)
