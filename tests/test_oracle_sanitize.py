"""The CPU oracle under AddressSanitizer, LeakSanitizer and
UndefinedBehaviorSanitizer: oracle/sanitize_driver.c runs every gradient
form, limiter and flow model on a small sheared box (iterations, forces,
node values) in a sanitized build; any out-of-bounds access, leak or
undefined operation fails the run."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc is needed to build the sanitized oracle")
def test_oracle_is_clean_under_address_and_undefined_behaviour_sanitizers(tmp_path):
    exe = tmp_path / "orc_san"
    build = subprocess.run(
        ["gcc", "-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
         "-fno-sanitize-recover=undefined", "-ffp-contract=off", "-Wno-alloc-size-larger-than",
         "-o", str(exe), os.path.join(ROOT, "oracle", "phl_oracle.c"), os.path.join(ROOT, "oracle", "sanitize_driver.c"),
         "-lm"], capture_output=True, text=True)
    assert build.returncode == 0, build.stderr[-2000:]
    run = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300,
                         env=dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
                                  UBSAN_OPTIONS="print_stacktrace=1"))
    assert run.returncode == 0, run.stderr[-4000:]
    assert "sanitized oracle runs: ok" in run.stdout
    assert "runtime error" not in run.stderr
