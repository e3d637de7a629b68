"""The library's host code under AddressSanitizer, UBSan and LeakSanitizer:
tests/host_asan/vx_host_check (built on the CPU by tests/host_asan/build.sh,
sanitizers in host code only) drives contexts, chunks, meshes, batches in
both forms with arena growth, the server under a budget with work in flight
at destruction, the world tick loop, frustum masks and draw lists through the
C ABI, and must end with no sanitizer report.  Allocations the HIP runtime
keeps for the process are suppressed (tests/host_asan/lsan.supp)."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "host_asan")
EXE = os.path.join(HERE, "vx_host_check")


def test_host_code_under_sanitizers():
    if not os.path.exists(EXE):
        pytest.skip("tests/host_asan/build.sh has not been run in this tree")
    print("[case] host check under ASan/UBSan/LSan", flush=True)
    env = dict(os.environ,
               ASAN_OPTIONS="detect_leaks=1:halt_on_error=1",
               LSAN_OPTIONS=f"suppressions={os.path.join(HERE, 'lsan.supp')}:print_suppressions=0",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    p = subprocess.run([EXE], capture_output=True, text=True, timeout=240, env=env)
    report = (p.stdout + p.stderr)[-6000:]
    assert p.returncode == 0, report
    assert "host check ok" in p.stdout, report
    for marker in ("ERROR: AddressSanitizer", "ERROR: LeakSanitizer", "runtime error:"):
        assert marker not in p.stderr, report
