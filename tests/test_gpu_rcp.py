"""The reciprocal of the intersection test (hpt::rcp_rn: v_rcp_f32 and one
Newton step in fused multiply-adds, the division for out-of-range inputs)
equals the correctly rounded division 1.f / d for all 2^32 float bit
patterns on this GPU: tests/cpp/rcp_exhaustive, built by
__graft_entry__.build() with the library's compiler flags."""
from __future__ import annotations

import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PROGRAM = os.path.join(ROOT, "tests", "cpp", "rcp_exhaustive")


def test_rcp_rn_equals_division_for_every_float():
    if not os.path.exists(PROGRAM):
        pytest.fail("tests/cpp/rcp_exhaustive is not built: run __graft_entry__.build()")
    print("\n[rcp] checking all 2^32 floats", flush=True)
    r = subprocess.run([PROGRAM], capture_output=True, text=True, timeout=120)
    print(r.stdout.strip(), flush=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "checked 4294967296 mismatches 0" in r.stdout
