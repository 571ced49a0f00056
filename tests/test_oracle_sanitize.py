"""The CPU oracle under AddressSanitizer and UndefinedBehaviorSanitizer: every entry point
driven on a small system (oracle/sanity_main.c, built by `make -C oracle sanitize` with
gcc, host code only).  Any out-of-bounds access, use after free, leak or undefined
behaviour aborts the driver."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_oracle_clean_under_sanitizers():
    b = subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "-s", "sanitize"], capture_output=True, text=True)
    assert b.returncode == 0, b.stderr
    env = dict(os.environ, OMP_NUM_THREADS="2", ASAN_OPTIONS="detect_leaks=1:abort_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = subprocess.run([os.path.join(ROOT, "oracle", "_san", "cro_sanity")], capture_output=True, text=True, env=env,
                       timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "sanity: ok" in r.stdout
