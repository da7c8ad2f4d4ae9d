"""Shared pytest setup.

Markers: ``gpu`` -- needs a visible MI355X (HIP device); everything else runs
on the CPU.  The oracle (oracle/) is imported only from tests as the checker.
"""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG_DIR = os.path.join(ROOT, "fft3d-fpga_amd")
for p in (ROOT, PKG_DIR):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP GPU (MI355X)")


@pytest.fixture(scope="session")
def orc():
    from oracle import oracle
    oracle.lib()
    return oracle


@pytest.fixture(scope="session")
def fl():
    import fftfpga_amd
    fftfpga_amd.load()
    return fftfpga_amd


@pytest.fixture(scope="session")
def gpu_lib(fl):
    """Library initialised on the GPU through the drop-in entry point."""
    rc = fl.fpga_initialize("AMD HIP", fl.LIB_PATH, False)
    if rc != 0:
        pytest.fail(f"fpga_initialize returned {rc} on a GPU run: {fl.last_error()}")
    yield fl
    fl.fpga_final()
