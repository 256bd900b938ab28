"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer (CPU only):
the product's plan builders (cpp_projects_amd/csrc/plan.cpp, lowered for every
init form at the workload shapes) and the oracle's CPU routines, compiled
with g++ -fsanitize=address,undefined -fno-sanitize-recover=all and run by
tests/cpp/sanitize_driver.cpp; any report fails the test."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("g++") is None, reason="no host compiler")
def test_plan_and_oracle_under_asan_ubsan(tmp_path):
    exe = tmp_path / "sanitize_driver"
    flags = ["-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
             "-ffp-contract=off"]
    objs = []
    for src, lang in ((os.path.join(ROOT, "oracle", "sps_oracle.c"), ["-x", "c", "-std=c11"]),
                      (os.path.join(ROOT, "cpp_projects_amd", "csrc", "plan.cpp"), ["-std=c++17"]),
                      (os.path.join(ROOT, "tests", "cpp", "sanitize_driver.cpp"), ["-std=c++17"])):
        o = tmp_path / (os.path.basename(src) + ".o")
        r = subprocess.run(["g++", *flags, *lang, "-I", os.path.join(ROOT, "include"), "-c", src, "-o", str(o)],
                           capture_output=True, text=True)
        assert r.returncode == 0, r.stderr[-3000:]
        objs.append(str(o))
    r = subprocess.run(["g++", *flags, *objs, "-lm", "-o", str(exe)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([str(exe)], capture_output=True, text=True, env=env, timeout=600)
    assert r.returncode == 0 and "sanitize driver ok" in r.stdout, (r.stdout + r.stderr)[-4000:]
