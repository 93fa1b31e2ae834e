"""Test configuration: paths, the `gpu` marker and helpers mirroring the reference's test/conftest.py."""

import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "bempp-cl_amd"), ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU and the built libbempp_hip.so")


class Helpers(object):
    """Helpers of the reference test suite (test/conftest.py:147-174)."""

    @staticmethod
    def load_npy_data(name):
        return np.load(os.path.join(GOLDEN, name + ".npy"))

    @staticmethod
    def load_npz_data(name):
        return np.load(os.path.join(GOLDEN, name + ".npz"))

    @staticmethod
    def load_grid(name):
        from bempp_amd.api.grid import import_grid

        return import_grid(os.path.join(GOLDEN, name + ".msh"))

    @staticmethod
    def default_tolerance(precision):
        return 6e-4 if precision == "single" else 1e-10


@pytest.fixture
def helpers():
    return Helpers


@pytest.fixture
def default_parameters():
    from bempp_amd.api.utils.parameters import DefaultParameters

    p = DefaultParameters()
    p.quadrature.regular = 4
    p.quadrature.singular = 4
    return p


@pytest.fixture(scope="session")
def sphere_grid():
    return Helpers.load_grid("sphere")


@pytest.fixture(scope="session")
def hip_available():
    import torch
    from bempp_amd.core import hip_backend

    if not torch.cuda.is_available():
        pytest.fail("GPU test run without a visible GPU")
    hip_backend.lib()
    return True
