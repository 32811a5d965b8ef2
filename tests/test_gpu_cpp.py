"""Runs the C++ tests of the drop-in backend (tests/cpp/) on the GPU.  They
instantiate parafaft::MI355Backend<T> and the one-rank ParaFaFT_C2C /
ParaFaFT_R2C drivers exactly as the reference's own test executables do and
return their failure count."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
BUILD = os.path.join(HERE, "cpp", "_build")


EXES = ["test_mi355_backend", "test_mi355_nd", "test_pfxvector", "test_mi355_mpi"]


@pytest.mark.parametrize("exe", EXES + [f"asan/{e}" for e in EXES])
def test_cpp_executable(exe):
    """asan/*: the same tests with AddressSanitizer and UBSan on the host
    code (leak detection off: the HIP runtime keeps allocations to exit)."""
    path = os.path.join(BUILD, exe)
    if not os.path.exists(path):
        pytest.fail(f"{path} not built (run __graft_entry__.build())")
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0")
    r = subprocess.run([path], capture_output=True, text=True, timeout=110, env=env)
    print(r.stdout[-4000:])
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-2000:]
    assert "0 failure(s)" in r.stdout
