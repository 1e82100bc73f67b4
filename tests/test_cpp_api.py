"""The C++ API (include/gpufhe.hpp) through tests/cpp/test_api.cpp, a program
that calls it the way the reference's C++ tests call the reference's API and
checks every result against the oracle."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
BIN = os.path.join(HERE, "cpp", "test_api")


def build_test_api():
    """Compiles tests/cpp/test_api against the built library and the oracle."""
    cmd = ["hipcc", "-O2", "-std=c++17", os.path.join(HERE, "cpp", "test_api.cpp"), "-o", BIN,
           "-L" + os.path.join(ROOT, "gpu-resident-library_amd"), "-lgpufhe_amd",
           "-L" + os.path.join(ROOT, "oracle"), "-loracle",
           "-Wl,-rpath,$ORIGIN/../../gpu-resident-library_amd", "-Wl,-rpath,$ORIGIN/../../oracle"]
    subprocess.check_call(cmd)


def _run(group, timeout):
    if not os.path.exists(BIN):
        pytest.fail("tests/cpp/test_api is not built (run __graft_entry__.build())")
    r = subprocess.run([BIN, group], capture_output=True, text=True, timeout=timeout)
    print(r.stdout[-4000:])
    fails = [l for l in r.stdout.splitlines() if l.startswith("FAIL")]
    assert r.returncode == 0 and not fails, "\n".join(fails) or r.stderr[-2000:]
    assert "PASS" in r.stdout


def test_cpp_api_host():
    if not os.path.exists(BIN):
        build_test_api()
    _run("cpu", 300)


@pytest.mark.gpu
def test_cpp_api_device():
    _run("gpu", 240)
