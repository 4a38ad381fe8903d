"""Shared fixtures.  GPU tests are marked ``gpu``; they call the library
through its C ABI (ctypes) and compare with the oracle on the CPU."""
import importlib
import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if REPO not in sys.path:
    sys.path.insert(0, REPO)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the built libedpp.so")


def edpp_module():
    return importlib.import_module("efficient-dpp_amd")


@pytest.fixture(scope="session")
def edpp():
    return edpp_module()


@pytest.fixture(scope="session")
def orc():
    from oracle import pyoracle
    return pyoracle


@pytest.fixture(scope="session")
def torch_gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU visible")
    return torch


@pytest.fixture(scope="session")
def ctx(edpp, torch_gpu):
    """One library context on torch's current stream of cuda:0."""
    torch = torch_gpu
    torch.cuda.set_device(0)
    c = edpp.Context(0, stream=torch.cuda.current_stream().cuda_stream)
    yield c
    c.close()
