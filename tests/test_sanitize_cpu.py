"""Host AddressSanitizer + UndefinedBehaviorSanitizer run of the C ABI.

The library built with `python 3d_ravesim_amd/build.py --sanitize`
(libravesim_hip_san.so: host code instrumented, device code as usual) is
driven by tests/sanitize/abi_check.cpp, an executable carrying the sanitizer
runtimes.  Here, with no GPU, the checker runs the argument checks, error
strings, run controls and cache queries; tests/test_gpu_sanitize.py runs its
device half.  Any heap misuse or undefined behaviour in the host layer aborts
the checker with a report and a non-zero status.
"""

import importlib.util
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _build_module():
    spec = importlib.util.spec_from_file_location("_rsim_build", os.path.join(ROOT, "3d_ravesim_amd", "build.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


SAN_ENV = {
    # the HIP runtime's own allocations live until exit: leaks are not checked
    "ASAN_OPTIONS": "detect_leaks=0:halt_on_error=1:abort_on_error=0",
    "UBSAN_OPTIONS": "halt_on_error=1:print_stacktrace=1",
}


def run_checker(mode, timeout):
    B = _build_module()
    if not os.path.exists(B.LIB_SAN):
        pytest.skip("no host-sanitizer build (python 3d_ravesim_amd/build.py --sanitize)")
    # build.py --sanitize links the checker with the library; link it here only
    # when it is missing (a host-only clang++ step of a few seconds)
    exe = B.CHECK_EXE if os.path.exists(B.CHECK_EXE) else B.build_checker()
    env = dict(os.environ, **SAN_ENV)
    p = subprocess.run([exe, mode], capture_output=True, text=True, timeout=timeout, env=env)
    print(p.stderr[-4000:])
    assert p.returncode == 0, p.stderr[-4000:]
    assert f"abi_check {mode}: OK" in p.stdout
    assert "ERROR: AddressSanitizer" not in p.stderr
    assert "runtime error:" not in p.stderr
    return p


def test_c_abi_host_layer_is_clean_under_asan_and_ubsan():
    run_checker("cpu", timeout=120)
