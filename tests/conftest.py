import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")
    # scalars must be taken from 0-d/one-element records explicitly (a future NumPy error)
    config.addinivalue_line("filterwarnings", "error:Conversion of an array with ndim > 0 to a scalar:DeprecationWarning")
