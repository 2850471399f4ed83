"""Sanitizer runs (AddressSanitizer + UndefinedBehaviorSanitizer, any finding
aborts): the CPU oracle driven over every function (oracle/sanitize_main.cpp,
built by `make -C oracle sanitize`), and the library's host code (argument
checks, planning, workspace carving, executors; device code unchanged) driven
through the C ABI (tests/csrc/host_sanitize_main.cpp against the
host-sanitized library, both built by __graft_entry__.build()): validation
paths on any machine, every entry point family on a GPU."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle")
ORACLE_BIN = os.path.join(ORACLE, "_build", "oracle_sanitize")
DRIVER = os.path.join(ROOT, "tests", "_build", "host_sanitize")


def _run(path, leaks):
    env = dict(os.environ, ASAN_OPTIONS=f"detect_leaks={1 if leaks else 0}", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([path], capture_output=True, text=True, env=env, timeout=600)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-4000:]
    assert "runtime error" not in out and "AddressSanitizer" not in out, out[-4000:]
    return r.stdout


def test_oracle_under_sanitizers():
    subprocess.check_call(["make", "-s", "-C", ORACLE, "sanitize"])
    assert "oracle sanitizer run: ok" in _run(ORACLE_BIN, leaks=True)


def _driver():
    if not os.path.exists(DRIVER):
        pytest.fail(f"{DRIVER} missing: run __graft_entry__.build()")
    return DRIVER


def test_library_host_code_validation_under_sanitizers():
    # (The HIP runtime keeps allocations until exit: leak reports are off.)
    assert "library host code under sanitizers: ok" in _run(_driver(), leaks=False)


@pytest.mark.gpu
def test_library_host_code_on_device_under_sanitizers(gpu):
    assert "(validation and device calls)" in _run(_driver(), leaks=False)
