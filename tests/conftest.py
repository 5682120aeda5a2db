import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")
    config.addinivalue_line("markers", "slow: minutes of CPU oracle work")


def gpu_available() -> bool:
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def water_svp():
    from dftax_amd.system import WATER, Molecule
    return Molecule.from_xyz(WATER, "def2-svp", spherical=True)


@pytest.fixture(scope="session")
def water_dimer_tzvp():
    from dftax_amd.system import Molecule
    geom = ("O 0.000 0.000 0.000; H 0.757 0.587 0.000; H -0.757 0.587 0.000;"
            "O 0.300 -0.400 2.900; H 0.900 -0.100 3.600; H -0.500 0.100 3.100")
    return Molecule.from_xyz(geom, "def2-tzvp", spherical=True)
