import ctypes
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

REPO = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(REPO))

GOLDEN = REPO / "tests" / "golden"
ORACLE_LIB = REPO / "oracle" / "liboracle.so"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def _ensure_oracle():
    if not ORACLE_LIB.exists():
        subprocess.run(["make", "-C", str(REPO / "oracle"), str(ORACLE_LIB)], check=True,
                       stdout=subprocess.DEVNULL)
    return ORACLE_LIB


@pytest.fixture(scope="session")
def oracle():
    from oracle import tde_oracle
    _ensure_oracle()
    return tde_oracle.Oracle(ORACLE_LIB)


@pytest.fixture(scope="session")
def golden():
    return {k: np.load(GOLDEN / f"{k}.npz") for k in ("g1_points", "g2_sources", "g3_blocks",
                                                      "g4_halfspace")}


@pytest.fixture(scope="session")
def pq3d_lib():
    from pyquake3d_amd import _lib
    if not _lib.LIB_PATH.exists():
        from pyquake3d_amd.build import build_library
        build_library()
    return _lib.load()


@pytest.fixture(scope="session")
def gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU test selected but no GPU is visible")
    return torch.device("cuda:0")
