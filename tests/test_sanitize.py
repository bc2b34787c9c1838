"""AddressSanitizer + UndefinedBehaviorSanitizer builds of the host-side code
(the OBJ and RGBE readers, the BVH builder and the generators of
libmambart, compiled with g++ from their sources) and of the oracle (gcc),
each driven by a harness in tests/sanitize/ over well-formed and malformed
inputs.  Any invalid access, overflow or undefined operation aborts the
harness with a sanitizer report; the test requires a clean exit."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "mambatracer_amd", "csrc")
ORACLE = os.path.join(ROOT, "oracle")
SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer", "-g", "-O1"]


def _build_and_run(tmp_path, cmd, exe):
    res = subprocess.run(cmd + ["-o", str(exe)], capture_output=True, text=True)
    assert res.returncode == 0, res.stderr[-4000:]
    env = dict(os.environ, ASAN_OPTIONS="abort_on_error=0:detect_leaks=1", UBSAN_OPTIONS="print_stacktrace=1")
    run = subprocess.run([str(exe)], capture_output=True, text=True, env=env, timeout=600)
    print(run.stdout[-2000:])
    assert run.returncode == 0, (run.stdout[-2000:], run.stderr[-6000:])
    assert "runtime error" not in run.stderr and "ERROR: AddressSanitizer" not in run.stderr, run.stderr[-6000:]


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_host_code_under_sanitizers(tmp_path):
    srcs = [os.path.join(ROOT, "tests", "sanitize", "host_check.cpp")] + \
        [os.path.join(CSRC, f) for f in ("mt_io.cpp", "mt_host.cpp", "mt_error.cpp")]
    _build_and_run(tmp_path, ["g++", "-std=c++17", *SAN, *srcs, "-lpthread"], tmp_path / "host_check")


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
def test_oracle_under_sanitizers(tmp_path):
    srcs = [os.path.join(ROOT, "tests", "sanitize", "oracle_check.c")] + \
        [os.path.join(ORACLE, f) for f in sorted(os.listdir(ORACLE)) if f.endswith(".c")]
    _build_and_run(tmp_path, ["gcc", "-std=gnu11", "-ffp-contract=off", "-I" + CSRC, *SAN, *srcs, "-lm",
                              "-lpthread"], tmp_path / "oracle_check")
