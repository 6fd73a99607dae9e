"""The C++ host classes of include/neon_hip.hpp (tests/cpp/host_api_test.cpp, built by
neon_amd/build.py with g++ against libneon_hip.so).

Without a GPU the program checks the factory's errors and that creating a context fails
loudly with NO_DEVICE (exit 77); on the GPU it also runs the reference's tridiagonal golden
system through both solvers, uniaxial neo-Hooke patch tests against their closed forms (with the
step hook) and the natural frequencies of a bar against the 1-D consistent-mass chain."""
import pathlib
import subprocess

import pytest

BIN = pathlib.Path(__file__).resolve().parent / "cpp" / "host_api_test"


def _run():
    if not BIN.exists():
        pytest.fail(f"{BIN} is not built (neon_amd/build.py build_host_test)")
    return subprocess.run([str(BIN)], capture_output=True, text=True, timeout=300)


def test_host_classes_without_device():
    from neon_amd import _capi

    if _capi.device_count() > 0:
        pytest.skip("a GPU is present: covered by the gpu test")
    r = _run()
    assert r.returncode == 77, r.stdout + r.stderr
    assert "FAIL" not in r.stdout


@pytest.mark.gpu
def test_host_classes_on_device():
    r = _run()
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "FAIL" not in r.stdout
