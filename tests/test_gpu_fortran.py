"""The Fortran host binding (iso_c_binding module) driving the HIP library."""
import os
import subprocess

import pytest

from tsc_tri_t1t2j1j2_amd import capi

pytestmark = pytest.mark.gpu
EXE = os.path.join(os.path.dirname(capi.LIB_PATH), "fortran", "test_hotpath")


@pytest.mark.skipif(not os.path.exists(EXE), reason="Fortran check program not built (no flang at build time)")
def test_fortran_host_program():
    out = subprocess.run([EXE], capture_output=True, text=True, timeout=120)
    print(out.stdout, out.stderr)
    assert out.returncode == 0
    lines = [l for l in out.stdout.splitlines() if l.startswith(("PASS", "FAIL"))]
    assert len(lines) == 4 and all(l.startswith("PASS") for l in lines), out.stdout
