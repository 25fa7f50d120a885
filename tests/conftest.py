import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, 'tests', 'golden')


def pytest_configure(config):
    config.addinivalue_line(
        'markers', 'gpu: needs an AMD Instinct GPU and the built libpercolate_hip.so')


@pytest.fixture(scope='session')
def golden():
    """Loader of the committed fixtures under tests/golden/."""
    def load(name):
        path = os.path.join(GOLDEN, name)
        if name.endswith('.json'):
            with open(path) as f:
                return json.load(f)
        return dict(np.load(path, allow_pickle=False))
    return load


@pytest.fixture(scope='session')
def gpu():
    """torch with a visible ROCm GPU, or skip."""
    import torch
    if not torch.cuda.is_available():
        pytest.skip('no ROCm GPU visible')
    return torch
