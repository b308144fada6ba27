"""CPU test of the MATLAB gateway's argument conversion and context cache.

The gateway (marie-tools_amd/csrc/mex/vie_hip_mex.cpp) can only be compiled
where MATLAB is installed; everything it does with its arguments lives in
plain C++ (mex/vie_mex_args.cpp) that this test drives through a g++ test
program built with AddressSanitizer and UndefinedBehaviorSanitizer: idxS of
every numeric class MARIE uses (int32 in the shipped body models, double
elsewhere; geo_body_domain.m:13), unsorted idxS, the content fingerprint of
a T-struct, and the context cache (content keys, LRU eviction, exact material
comparison).
"""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "marie-tools_amd", "csrc")
EXE = os.path.join(CSRC, "_build", "test_vie_mex_args")


def test_mex_argument_conversion_and_cache():
    # g++ only (no GPU code); builds in about a second
    subprocess.run(["make", "-s", "-C", CSRC, "mexargs-test"], check=True, timeout=300)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    r = subprocess.run([EXE], capture_output=True, text=True, timeout=120, env=env)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.strip().splitlines()[-1] == "ALL OK", r.stdout
    assert "runtime error" not in r.stderr, r.stderr
