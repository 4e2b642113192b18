"""The host-side path-tracing code under AddressSanitizer and UndefinedBehaviorSanitizer
(the reference ran compute-sanitizer memcheck over its renderer, Tools/verify_bvh.sh:9-10):
tests/sanitize/host_driver.cpp runs rt_core.h (sampler, intersection, slab tests,
integrator) and traverse.h (2/4/8-wide and quantized traversal) with the oracle on edge-case
scenes, checks hits against brute force and its image against the oracle's, and any
sanitizer report fails the run.  CPU only."""
import subprocess
from pathlib import Path

HERE = Path(__file__).resolve().parent / "sanitize"


def test_host_code_is_clean_under_asan_and_ubsan():
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True, timeout=600)
    r = subprocess.run([str(HERE / "build" / "host_driver")], capture_output=True, text=True, timeout=600,
                       env={"ASAN_OPTIONS": "detect_leaks=1:abort_on_error=1", "UBSAN_OPTIONS": "print_stacktrace=1",
                            "PATH": "/usr/bin:/bin"})
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-4000:]
    assert "ERROR: AddressSanitizer" not in out and "runtime error:" not in out, out[-4000:]
    assert "0 failures" in r.stdout
