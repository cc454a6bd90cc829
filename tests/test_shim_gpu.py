"""The C++ drop-in header (include/cbx_combblas.hpp) called from genuine
CombBLAS SpDCCols objects, compared with the reference's own CPU kernels.
The check program is built from the reference tree by oracle/Makefile
(target `ref`) into oracle/_ref/shim_check and travels with the checkout."""
from __future__ import annotations

import subprocess
from pathlib import Path

import pytest

pytestmark = pytest.mark.gpu
PROG = Path(__file__).resolve().parents[1] / "oracle" / "_ref" / "shim_check"


@pytest.mark.skipif(not PROG.exists(), reason="oracle/_ref/shim_check not built (needs the reference tree)")
def test_dropin_header_matches_reference_kernels():
    r = subprocess.run([str(PROG)], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "shim ok" in r.stdout
