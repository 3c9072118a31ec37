import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))


def _ensure_built():
    for path, lib in (("dopamine_amd/csrc", "dopamine_amd/libdopamine_hip.so"), ("oracle", "oracle/liboracle.so"),
                      ("tests/hostcheck", "tests/hostcheck/libhostcheck.so")):
        if not os.path.exists(os.path.join(ROOT, lib)):
            subprocess.run(["make", "-C", os.path.join(ROOT, path), "-j8"], check=True)


_ensure_built()


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")
