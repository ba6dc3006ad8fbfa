import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP) device")
    config.addinivalue_line("markers", "slow: takes more than a few seconds on the CPU")


def gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def golden_dir():
    return os.path.join(ROOT, "tests", "golden")


# Parity figures: the GPU tests print the errors they measure (lines with "rel-L2", "rel " or "of the peak").  Those
# lines of passing tests are repeated in the terminal summary, so a quiet run (-q) still records the measured values.
_PARITY_LINES = []
_PARITY_KEYS = ("rel-L2", " rel ", "of the peak", "rel-")


def pytest_runtest_logreport(report):
    if report.when == "call" and report.passed and report.capstdout:
        for line in report.capstdout.splitlines():
            if any(k in line for k in _PARITY_KEYS):
                _PARITY_LINES.append(f"{report.nodeid.split('::')[-1]}: {line.strip()}")


def pytest_terminal_summary(terminalreporter):
    if _PARITY_LINES:
        terminalreporter.section("measured parity figures")
        for line in _PARITY_LINES:
            terminalreporter.write_line(line)
