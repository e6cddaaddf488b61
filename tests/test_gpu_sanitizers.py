"""The C ABI's host side under AddressSanitizer and UndefinedBehaviorSanitizer on the GPU: the plan
builder (classes, families, flow lists, wave segments, set CSR, work items), plans in parts, the
per-trade PFE rows, every valuation variant and the autotuner, a batch, the adjoint's linear form, the
XVA reducer on a resident cube and a G2++ plan, run by tests/native/api_selftest.cpp with libeqhip's
host code instrumented (the device code is not). The binary is built beforehand on the CPU with
`make -C tests/native api_selftest`."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BINARY = os.path.join(ROOT, "tests", "native", "_build", "api_selftest")


def test_c_abi_host_side_under_sanitizers():
    if not os.path.exists(BINARY):
        pytest.skip("not built: make -C tests/native api_selftest")
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:halt_on_error=1",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    print("running the sanitized C ABI self-test", flush=True)
    r = subprocess.run([BINARY], env=env, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-4000:]
    assert "C ABI self-test passed" in r.stdout
