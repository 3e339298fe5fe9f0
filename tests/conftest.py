import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
GOLDEN = ROOT / "tests" / "golden"
MODELS = GOLDEN / "models"
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


@pytest.fixture(scope="session")
def synthetic18(tmp_path_factory):
    """The seeded 18-orbital triangular model (3 atoms x 6 orbitals)."""
    from xatu_amd import synthetic
    p = tmp_path_factory.mktemp("models") / "random18.model"
    return str(synthetic.write_model(p, natoms=3, orb_per_atom=6))


@pytest.fixture(scope="session")
def synthetic36(tmp_path_factory):
    """The seeded 36-orbital model (4 atoms x 9 orbitals)."""
    from xatu_amd import synthetic
    p = tmp_path_factory.mktemp("models") / "random36.model"
    return str(synthetic.write_model(p, natoms=4, orb_per_atom=9))
