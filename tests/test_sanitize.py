"""Host sanitizer builds (AddressSanitizer + UndefinedBehaviorSanitizer),
run in the CPU test set: the library's device-free logic
(old-gpuintegration_amd/csrc/host_logic.cpp with tests/sanitize_host.cpp)
and the whole CPU oracle (oracle/cuhre_oracle.c with tests/sanitize_oracle.c).
A sanitizer finding aborts the program, so both must exit 0 and say ok."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.parametrize("where,binary,ok", [
    ("old-gpuintegration_amd", "_asan/sanitize_host", "host checks ok"),
    ("oracle", "_asan/sanitize_oracle", "oracle checks ok"),
])
def test_sanitized_build_runs_clean(where, binary, ok):
    d = os.path.join(ROOT, where)
    subprocess.run(["make", "-s", "-C", d, "sanitize"], check=True, capture_output=True, timeout=300)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([os.path.join(d, binary)], capture_output=True, text=True, timeout=300, env=env)
    assert r.returncode == 0, r.stdout + r.stderr
    assert ok in r.stdout
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr
