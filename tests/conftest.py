import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line(
        'markers', 'gpu: needs a real MI355X (select with -m gpu)')


def pytest_collection_modifyitems(config, items):
    # GPU tests need the built libraries; CPU-only tests never touch them.
    pass
