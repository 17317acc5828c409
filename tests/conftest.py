import os
import sys

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))

import oracle_path  # noqa: E402,F401


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) to run")
