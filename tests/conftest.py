import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libhexa on the device)")


def gpu_available() -> bool:
    try:
        from v119_2025_12_19_hexatop_aws_amd import hexa
        hexa.device_info()
        return True
    except Exception:
        return False
