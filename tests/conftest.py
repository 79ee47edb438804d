import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "dgfem-acoustic_amd")
for p in (PKG, os.path.dirname(os.path.abspath(__file__))):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")
    # build what is missing (never rebuilds what exists)
    if not os.path.exists(os.path.join(ROOT, "oracle", "liboracle.so")):
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle")], check=True)
    if not os.path.exists(os.path.join(PKG, "libdgkhost.so")):
        subprocess.run([sys.executable, os.path.join(PKG, "build.py"), "host"], check=True)


REFERENCE = "/root/reference"


@pytest.fixture
def reference_dir():
    if not os.path.isdir(REFERENCE):
        pytest.skip("reference tree not mounted")
    return REFERENCE
