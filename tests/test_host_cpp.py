"""The C++ host layer (mole_amd/host/include/mole: the reference's traits,
Runner and VmcRunner over the C ABI): its test program's host-only checks,
which need no GPU."""
import os
import subprocess

import pytest

HOST = os.path.join(os.path.dirname(__file__), "..", "mole_amd", "host")
BIN = os.path.join(HOST, "_build", "host_tests")


@pytest.fixture(scope="module")
def host_bin():
    if not os.path.exists(BIN):
        from mole_amd.host import build
        build.build()
    return BIN


def test_host_layer_checks(host_bin):
    r = subprocess.run([host_bin], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "host tests passed" in r.stdout
