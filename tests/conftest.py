import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP engine through the C ABI)")
    config.addinivalue_line("markers", "slow: long-running CPU oracle case")


def _ensure_oracle():
    from oracle import oracle as orc

    if not os.path.exists(orc.library_path()):
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle")], check=True, capture_output=True)


_ensure_oracle()
