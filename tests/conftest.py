import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a ROCm GPU and the built libjqmc_hip.so")
    config.addinivalue_line("markers", "slow: long CPU oracle runs")


def _print_case(item):
    # name the case before any GPU work starts, so a crashed run still identifies it
    sys.stdout.write(f"\n[case] {item.nodeid}\n")
    sys.stdout.flush()


def pytest_runtest_setup(item):
    if item.get_closest_marker("gpu") is not None:
        _print_case(item)
        import torch

        if not torch.cuda.is_available():
            pytest.skip("no GPU visible")
