"""Host-side code under AddressSanitizer and UndefinedBehaviorSanitizer: the CPU oracle (the
step in both formulations, neighbour lists against brute force, non-finite and out-of-range
coordinates, the FP64 evaluation, the surrogate's forward/gradient/Adam), the host kernel
classes and the host particle container (removal semantics against the reference's).  Built with gcc/g++ from tests/native/Makefile; no GPU involved."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
NATIVE = os.path.join(HERE, "native")


@pytest.fixture(scope="module")
def binaries():
    if shutil.which("gcc") is None or shutil.which("g++") is None:
        pytest.skip("host compilers not available")
    r = subprocess.run(["make", "-C", NATIVE], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout + r.stderr
    return os.path.join(NATIVE, "_build")


@pytest.mark.parametrize("name", ["oracle_sanitize", "kernels_sanitize", "particles_check"])
def test_sanitized_run(binaries, name):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1", OMP_NUM_THREADS="4")
    r = subprocess.run([os.path.join(binaries, name)], capture_output=True, text=True,
                       timeout=300, env=env)
    assert r.returncode == 0 and "sanitize ok" in r.stdout, r.stdout + r.stderr
