import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
TESTS = os.path.dirname(os.path.abspath(__file__))
if TESTS not in sys.path:
    sys.path.insert(0, TESTS)

GOLD = os.path.join(TESTS, "golden")
REF = os.path.join(GOLD, "reference")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")
    config.addinivalue_line("markers", "slow: longer CPU test")


def ref_mol(name, basis, nelec):
    from librint_amd import mol
    return mol.read_reference_txt(os.path.join(REF, "mol_" + name, basis + ".txt"), nelec, name + "/" + basis)


def ref_truth(name, what):
    import numpy as np
    return np.loadtxt(os.path.join(REF, "truth", name, what + ".txt"))


@pytest.fixture(scope="session", autouse=True)
def _torch_device_first():
    """torch carries its own HIP runtime, which must bring the device up
    before the library's runtime does; any test may import the library, so
    on a GPU machine torch's device is initialised before the first test."""
    import torch
    if torch.cuda.is_available():
        torch.zeros(1, device="cuda:0")


@pytest.fixture(scope="session")
def gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch.device("cuda:0")
