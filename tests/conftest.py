'''Shared test setup: the repository root and this folder on the path, the `gpu` marker, and a fixed
torch seed per test (the reference's `tests/conftest.py:7-10`).'''
import os
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, HERE)


def pytest_configure(config):
    config.addinivalue_line('markers', 'gpu: runs HIP kernels on an MI355X')
    config.addinivalue_line('markers', 'slow: long CPU runs')


@pytest.fixture(autouse=True)
def seed_generator():
    import torch
    torch.manual_seed(0)
