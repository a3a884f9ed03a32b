"""CPU checks of the C-ABI library: it loads, exports every declared
symbol, the struct layouts match, and host-only entry points agree with
the oracle. No GPU compute is called here."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

from oracle import wigner
from tsc_tri_t1t2j1j2_amd import capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "tsc_hotpath.h")


def _declared():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(tsc_[a-z0-9_]+)\s*\(", src)))


def test_header_matches_binding_list():
    assert _declared() == sorted(capi.EXPORTS)


def test_library_exports_every_declared_symbol():
    assert os.path.exists(capi.LIB_PATH), "build() first"
    out = subprocess.run(["nm", "-D", "--defined-only", capi.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = set(line.split()[-1] for line in out.splitlines() if line.strip())
    missing = [s for s in _declared() if s not in exported]
    assert not missing, missing
    L = capi.lib()
    for s in _declared():
        assert getattr(L, s) is not None


class _Contrib(ctypes.Structure):
    _fields_ = [("a_off", ctypes.c_int64), ("b_off", ctypes.c_int64), ("lda", ctypes.c_int32),
                ("ldb", ctypes.c_int32), ("k", ctypes.c_int32), ("a_buf", ctypes.c_int8), ("b_buf", ctypes.c_int8),
                ("a_op", ctypes.c_int8), ("b_op", ctypes.c_int8), ("coef_re", ctypes.c_double),
                ("coef_im", ctypes.c_double)]


class _Out(ctypes.Structure):
    _fields_ = [("c_off", ctypes.c_int64), ("ldc", ctypes.c_int32), ("m", ctypes.c_int32), ("n", ctypes.c_int32),
                ("c_buf", ctypes.c_int32), ("cbeg", ctypes.c_int32), ("cend", ctypes.c_int32),
                ("mode", ctypes.c_int32)]


def test_struct_layouts_match_c():
    for S, dt in ((_Contrib, capi.CONTRIB_DTYPE), (_Out, capi.OUTBLOCK_DTYPE)):
        assert ctypes.sizeof(S) == dt.itemsize
        for name, _ in S._fields_:
            assert getattr(S, name).offset == dt.fields[name][1], name


def test_w6j_library_matches_exact_oracle():
    rng = np.random.default_rng(7)
    checked = 0
    for _ in range(30000):
        a = [int(x) for x in rng.integers(0, 14, 6)]
        ref = wigner.w6j(*a)
        got = capi.w6j(*a)
        assert abs(got - ref) <= 1e-13 * max(1.0, abs(ref)), (a, got, ref)
        checked += ref != 0.0
    assert checked > 100


def test_fortran_module_binds_every_entry_point():
    """The Fortran binding (the reference's host language) declares an
    interface for every function of include/tsc_hotpath.h."""
    import os
    import re
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    hdr = open(os.path.join(root, "include", "tsc_hotpath.h")).read()
    declared = set(re.findall(r"^(?:int|double|void \*|const char \*)\s*(tsc_[a-z0-9_]+)\(", hdr, re.M))
    mod = open(os.path.join(root, "tsc_tri_t1t2j1j2_amd", "fortran", "tsc_hotpath_mod.f90")).read()
    bound = set(re.findall(r'bind\(c, name="(tsc_[a-z0-9_]+)"\)', mod))
    assert declared and declared <= bound, sorted(declared - bound)
