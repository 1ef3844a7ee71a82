import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line(
        "markers", "gpu: needs an AMD GPU (MI355X) and the built libnemo_hip.so")
    config.addinivalue_line(
        "markers", "slow: takes more than a few seconds")
