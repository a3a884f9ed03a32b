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


def _random_spectrum(rng, nsec, dmax):
    dims = rng.integers(1, dmax + 1, size=nsec)
    twos = rng.integers(0, 5, size=nsec)
    eig = []
    for d in dims:
        w = np.sort(rng.exponential(size=d) ** 6)
        if rng.random() < 0.3:
            w[: d // 2] = 0.0   # rank-deficient sector
        eig.append(w)
    return dims, twos, eig


@pytest.mark.parametrize("seed", range(6))
def test_select_states_matches_the_kept_rule(seed):
    """tsc_select_states (host code of the library) against the oracle's
    restatement of den_mat.f90:255-364 on random spectra with ties and
    zeros, across the kept_min / kept_max / trun_err regimes."""
    rng = np.random.default_rng(seed)
    dims, twos, eig = _random_spectrum(rng, 12, 30)
    n = int(dims.sum())
    for kmin, kmax, terr in [(0, 5, 1e-3), (3, 40, 1e-6), (10, 12, 0.0), (n, n + 3, 1e-12), (n + 5, n + 9, 1e-9),
                             (1, n + 4, 0.0)]:
        kept, ns, err, ent, ent1 = truncation_select(eig, twos, kmin, kmax, terr)
        got = capi.select_states(dims, twos, np.concatenate(eig), kmin, kmax, terr)
        assert list(got["kept"]) == list(kept) and got["num_state"] == ns
        assert abs(got["trun_err"] - err) < 1e-14 and abs(got["entropy"] - ent) < 1e-12
        assert abs(got["entropy1"] - ent1) < 1e-12
        assert got["num_state"] <= n and int(np.sum(got["kept"])) == got["num_state"]
        assert all(0 <= k <= d for k, d in zip(got["kept"], dims))


def test_select_states_stops_at_the_end_of_the_spectrum():
    """trun_err = 0 and kept_max past the spectrum: the count never passes
    n, whatever the rounding remainder of the discarded weight."""
    rng = np.random.default_rng(11)
    for _ in range(200):
        dims, twos, eig = _random_spectrum(rng, int(rng.integers(1, 6)), 9)
        n = int(dims.sum())
        got = capi.select_states(dims, twos, np.concatenate(eig), 0, n + 10, 0.0)
        assert got["num_state"] <= n and int(np.sum(got["kept"])) == got["num_state"]


def test_select_states_rejects_negative_parameters():
    for kmin, kmax, terr in [(-1, 5, 1e-6), (1, -5, 1e-6), (1, 5, -1e-6)]:
        with pytest.raises(capi.TscError, match=">= 0"):
            capi.select_states([2], [0], [0.3, 0.7], kmin, kmax, terr)


def truncation_select(eig, twos, kmin, kmax, terr):
    from oracle import truncation
    return truncation.select(eig, twos, kmin, kmax, terr)


def test_fortran_descriptor_types_match_the_c_structs(tmp_path):
    """The bind(c) mirrors in tsc_hotpath_mod.f90 have the C structs' sizes
    (the planner and block-model descriptors a Fortran host packs)."""
    import shutil
    from tsc_tri_t1t2j1j2_amd import native
    fc = shutil.which("flang") or "/opt/rocm/llvm/bin/flang"
    if not os.path.exists(fc) and not shutil.which(fc):
        pytest.skip("no Fortran compiler")
    mod = os.path.join(ROOT, "tsc_tri_t1t2j1j2_amd", "fortran", "tsc_hotpath_mod.f90")
    prog = tmp_path / "sizes.f90"
    names = ["tsc_contrib", "tsc_outblock", "tsc_sector", "tsc_op_block", "tsc_op_desc", "tsc_pair_desc",
             "tsc_term_desc", "tsc_sb_desc", "tsc_site_ops", "tsc_model_desc", "tsc_bond"]
    body = "\n".join(f"  type({n}) :: v{i}" for i, n in enumerate(names))
    prints = "\n".join(f"  print '(i0)', c_sizeof(v{i})" for i in range(len(names)))
    prog.write_text(f"program sizes\n  use iso_c_binding\n  use tsc_hotpath\n  implicit none\n{body}\n{prints}\nend program\n")
    exe = tmp_path / "sizes"
    subprocess.run([fc, "-J", str(tmp_path), "-c", mod, "-o", str(tmp_path / "m.o")], check=True, cwd=tmp_path)
    pkg = os.path.dirname(capi.LIB_PATH)
    subprocess.run([fc, "-J", str(tmp_path), str(prog), str(tmp_path / "m.o"), "-o", str(exe), "-L", pkg,
                    "-ltsc_hotpath", "-Wl,-rpath," + pkg], check=True, cwd=tmp_path)
    got = [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    want = [capi.CONTRIB_DTYPE.itemsize, capi.OUTBLOCK_DTYPE.itemsize] + [
        ctypes.sizeof(t) for t in (native.Sector, native.OpBlock, native.OpDesc, native.PairDesc, native.TermDesc,
                                   native.SbDesc, native.SiteOps, native.ModelDesc, native.Bond)]
    assert got == want, (got, want)


def test_work_units_cover_every_tile_once_and_spread_over_the_xcds(tmp_path):
    """The library's work-unit builder, compiled into a host program: every
    tile of every output block gets units whose contribution ranges tile the
    block's range exactly once, and the per-XCD work stays balanced, also
    for one block much larger than an XCD's share (tests/native)."""
    import shutil
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc):
        pytest.skip("no hipcc")
    src = os.path.join(ROOT, "tests", "native", "work_units_check.cpp")
    pkg = os.path.dirname(capi.LIB_PATH)
    exe = tmp_path / "work_units_check"
    subprocess.run([hipcc, "-O1", "-std=c++17", "-I", os.path.join(ROOT, "include"), "-I",
                    os.path.join(pkg, "csrc"), src, "-o", str(exe), "-L", pkg, "-ltsc_hotpath",
                    "-Wl,-rpath," + pkg], check=True, capture_output=True, timeout=300)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0 and out.stdout.startswith("ok"), out.stdout + out.stderr


class _RecordingLib:
    """Stand-in for the library that records create/destroy calls (no GPU)."""

    def __init__(self):
        self.calls = []
        self.next = 1

    def _new(self, kind, out):
        out._obj.value = self.next
        self.calls.append(("create", kind, self.next))
        self.next += 1
        return 0

    def tsc_ctx_create(self, device, out):
        return self._new("ctx", out)

    def tsc_vec_create(self, ctx, n, out):
        return self._new("vec", out)

    def tsc_plan_create(self, ctx, *args):
        return self._new("plan", args[-1])

    def tsc_ctx_destroy(self, h):
        self.calls.append(("destroy", "ctx", h.value if hasattr(h, "value") else h))
        return 0

    def tsc_vec_destroy(self, h):
        self.calls.append(("destroy", "vec", h.value if hasattr(h, "value") else h))
        return 0

    def tsc_plan_destroy(self, h):
        self.calls.append(("destroy", "plan", h.value if hasattr(h, "value") else h))
        return 0


@pytest.mark.parametrize("order", ["ctx_first", "objects_first", "gc"])
def test_teardown_never_frees_after_the_context(monkeypatch, order):
    """Vectors and plans released after their context (the order Python may
    finalise objects in at interpreter exit, or a garbage-collected cycle)
    issue no device free: the library never sees a vector or plan destroy
    after the destroy of its context."""
    import gc
    rec = _RecordingLib()
    monkeypatch.setattr(capi, "_lib", rec)
    ctx = capi.Context(0)
    v = capi.Vec(ctx, 4)
    z = np.zeros(0, capi.CONTRIB_DTYPE)
    zo = np.zeros(0, capi.OUTBLOCK_DTYPE)
    p = capi.Plan(ctx, np.zeros(1, np.complex128), 1, 0, z, zo, z, zo, 1, 1)
    if order == "ctx_first":
        ctx.close()
        v.close()
        p.close()
    elif order == "objects_first":
        p.close()
        v.close()
        ctx.close()
    else:
        ring = [ctx, v, p]
        ring.append(ring)   # a cycle: finalisers run in the collector's order
        del ctx, v, p, ring
        gc.collect()
    kinds = [(a, k) for (a, k, _) in rec.calls]
    assert ("destroy", "ctx") in kinds
    after = kinds[kinds.index(("destroy", "ctx")) + 1:]
    assert not [c for c in after if c[0] == "destroy"], rec.calls
