"""The library's host code under AddressSanitizer and UndefinedBehaviorSanitizer.

tests/fake_device/sanitize_main.cpp drives the scene helpers, the parameter
editor (every key, spaceball events), the file writers (including a path
that cannot be opened and names cut at a small buffer), the serving
scheduler over the host-memory device with injected launch and device
failures, and a small render and volume of the CPU oracle. The build is g++
with -fsanitize=address,undefined (no device code); any report fails.
"""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
FAKE_DIR = os.path.join(HERE, "fake_device")


def test_host_code_is_clean_under_sanitizers(tmp_path):
    exe = os.path.join(FAKE_DIR, "host_sanitize")
    subprocess.run(["make", "-s", "-C", FAKE_DIR, exe], check=True)
    env = dict(os.environ)
    env["ASAN_OPTIONS"] = "detect_leaks=1:abort_on_error=0:exitcode=23"
    env["UBSAN_OPTIONS"] = "print_stacktrace=1:halt_on_error=1"
    env["OMP_NUM_THREADS"] = "2"
    p = subprocess.run([exe, str(tmp_path)], capture_output=True, text=True, env=env, timeout=600)
    report = p.stdout + p.stderr
    assert p.returncode == 0, report[-4000:]
    assert "runtime error" not in report and "AddressSanitizer" not in report, report[-4000:]
    assert "host code clean under the sanitizers" in p.stdout
    # the writers ran: their files exist with the reference's sizes
    assert os.path.getsize(tmp_path / "exps.raw") == 5 * 4 * 3 * 4
    assert os.path.getsize(tmp_path / "p.raw") == 7 * 5 * 36
    assert (tmp_path / "r.ppm").read_text().startswith("P3\n7 5\n255\n")
