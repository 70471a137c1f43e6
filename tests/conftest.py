"""pytest configuration: the `gpu` marker (tests that need an MI355X and the built library), and the
shared fixtures: the CPU oracle (oracle/) and the device library (sincfeyn_amd.sfx)."""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, 'oracle'))


def pytest_configure(config):
    config.addinivalue_line('markers', 'gpu: needs an AMD GPU (MI355X) and sincfeyn_amd/_lib/libsfx.so')


def _oracle(nc):
    import oracle as O
    if not O.available():
        pytest.fail('the oracle is not built: run make -C oracle (or __graft_entry__.build())')
    return O.Oracle(nc=nc)


@pytest.fixture(scope='session')
def oracle():
    """the oracle with the reference kernels built without FP contraction: the floating-point semantics
    of the reference's GPU path, which the device reproduces point by point"""
    return _oracle(True)


@pytest.fixture(scope='session')
def oracle_ref():
    """the oracle with the reference kernels as the reference's default CPU build compiles them"""
    return _oracle(False)


@pytest.fixture(scope='session')
def sfx():
    from sincfeyn_amd import sfx as S
    S.init(0)
    return S
