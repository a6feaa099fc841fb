"""The library's host code under AddressSanitizer: tests/asan/sar_asan_driver
(built by tests/asan/build.sh, which __graft_entry__.build() runs) drives the
C ABI on small problems -- plans from host, device and grid positions in four
interpolation / precision modes, fused and rocFFT spectral stages, the
reference-record entry points, resampling, signal decoding and bad
arguments -- with ASan on the host side (device code as in the product).
Any heap overflow, use after free or double free in the plan set-up or the
C ABI ends the driver with an ASan report and a non-zero status."""
import os
import subprocess
from pathlib import Path

import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

DRIVER = Path(__file__).resolve().parent / "asan" / "sar_asan_driver"


def test_host_code_under_address_sanitizer():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    if not DRIVER.exists():
        pytest.skip("ASan driver not built (tests/asan/build.sh)")
    env = dict(os.environ)
    env["ASAN_OPTIONS"] = ("detect_leaks=0:halt_on_error=1:verify_asan_link_order=0:"
                           "protect_shadow_gap=0")
    print("case: ASan host driver", flush=True)
    r = subprocess.run([str(DRIVER)], capture_output=True, text=True, timeout=240, env=env)
    print(r.stdout[-4000:])
    print(r.stderr[-4000:])
    assert r.returncode == 0
    assert "all checks passed" in r.stdout
    assert "AddressSanitizer" not in r.stderr
