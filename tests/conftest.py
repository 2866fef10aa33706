"""Shared pytest setup: import paths and the `gpu` marker."""
import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG_DIR = os.path.join(REPO, "i3d2025-voxel-transform-dag_amd")
ORACLE_DIR = os.path.join(REPO, "oracle")
for p in (PKG_DIR, ORACLE_DIR, os.path.join(REPO, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU; runs the HIP path through the C-ABI")


@pytest.fixture(scope="session")
def oracle():
    import oracle_ctypes

    oracle_ctypes.build()
    return oracle_ctypes


@pytest.fixture(scope="session")
def gpu_ctx():
    import torch

    import tdag_amd

    # torch's bundled HIP runtime must initialise the device before libtdag's ROCm runtime does;
    # tests hand torch device buffers to the library.
    torch.cuda.init()
    ctx = tdag_amd.Context(0)
    yield ctx
    ctx.close()
