import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, 'tests', 'golden')
LIH_ROOT = os.path.join(GOLDEN, 'lih_bulk08', 'lih_bulk08')


def pytest_configure(config):
    config.addinivalue_line('markers', 'gpu: needs an MI355X (runs through libmgw.so on cuda:0)')
    config.addinivalue_line('markers', 'slow: long-running (full benchmark sizes)')


@pytest.fixture(scope='session')
def engine():
    from molgw_amd.gw import Engine
    e = Engine(0)
    yield e
    e.close()
