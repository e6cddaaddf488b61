"""The oracle and the C++ host under AddressSanitizer and UndefinedBehaviorSanitizer (CPU): the oracle's
self-test (oracle/selftest.c, built by `make -C oracle sanitize`) runs every restated entry point on small
mixed inputs; the host's (tests/native/host_selftest.cpp, the host sources compiled with the sanitizers and
linked to the uninstrumented libeqhip.so) runs the registry, contexts, market, Greek grammar and, without a
device, the no-device failure of every pricing route. Both abort on the first finding."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN_ENV = dict(os.environ, ASAN_OPTIONS="halt_on_error=1:abort_on_error=0:detect_leaks=1",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")


def host_selftest_binary():
    """tests/native/_build/host_selftest (g++ with -fsanitize=address,undefined, tests/native/Makefile)."""
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "tests", "native"), "host_selftest"], check=True,
                   capture_output=True)
    return os.path.join(ROOT, "tests", "native", "_build", "host_selftest")


def test_oracle_under_sanitizers():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "sanitize"], check=True, capture_output=True)
    r = subprocess.run([os.path.join(ROOT, "oracle", "_san", "eqr_selftest")], env=SAN_ENV, capture_output=True,
                       text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-4000:]
    assert "oracle self-test passed" in r.stdout


def test_host_under_sanitizers():
    if not os.path.exists(os.path.join(ROOT, "engine_quant_amd", "libeqhip.so")):
        pytest.fail("libeqhip.so is not built (__graft_entry__.build())")
    binary = host_selftest_binary()
    from engine_quant_amd import eqhip
    gpu = eqhip.is_gpu_backend_available()
    env = dict(SAN_ENV)
    if gpu:  # the HIP runtime keeps allocations until exit, which the leak check would report
        env["ASAN_OPTIONS"] = "halt_on_error=1:detect_leaks=0"
    r = subprocess.run([binary] + (["--gpu"] if gpu else []), env=env, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-4000:]
    assert "host self-test passed" in r.stdout
