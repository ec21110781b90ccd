"""Device half of the host-sanitizer check (tests/sanitize/abi_check.cpp gpu):
volumes, fused and plain slice loops, cone-beam tables built by host threads,
chains, batches, captured graphs, per-run cancel and progress, both detector
forms, the adjoint's checkpoint workspace and cache release, with the
library's host code under AddressSanitizer and UndefinedBehaviorSanitizer."""

import pytest

from test_sanitize_cpu import run_checker

pytestmark = pytest.mark.gpu


def test_c_abi_device_runs_are_clean_under_host_asan_and_ubsan():
    print("abi_check gpu", flush=True)
    run_checker("gpu", timeout=240)
