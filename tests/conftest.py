import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, 't-nalg_amd')
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line('markers', 'gpu: needs an MI355X (ROCm) GPU and the built HIP library')
    config.addinivalue_line('markers', 'slow: longer CPU oracle runs')


def pytest_collection_modifyitems(config, items):
    # GPU tests print the case they run (a fault would otherwise leave no trace)
    pass


@pytest.fixture(scope='session')
def gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip('no GPU')
    from tnalg_amd import _lib
    _lib.load()
    return torch.device('cuda:0')
