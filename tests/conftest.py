import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import __graft_entry__  # noqa: E402

__graft_entry__.load()

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


def have_gpu():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def le_case():
    from fesomc_amd import meshio
    d = os.path.join(GOLDEN, "le_medium")
    # The reference's energy series was produced by a build where Mt=int(dt/dt_2D)
    # with dt=0.5, dt_2D=0.1 came out as 4 (IEEE double gives exactly 5).
    txt = open(os.path.join(d, "fv_run.dat")).read() + "\n Mt = 4\n"
    return txt, meshio.read_mesh(d)
