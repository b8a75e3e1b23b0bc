import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built libabcmb_hip.so")


@pytest.fixture(scope="session")
def fiducial():
    # config 1: flat LCDM, massless neutrinos (N_eff = 3.044)
    return {"h": 0.6736, "omega_b": 0.02237, "omega_cdm": 0.1200, "A_s": 2.1e-9, "n_s": 0.9649,
            "tau_reion": 0.0544, "Neff": 3.044}
