"""Per-length files in the reference's layouts and checkpoint/restart
(block_to_disk, model_to_disk, truns_to_disk, wave_to_disk, basis_to_disk,
DMRG_code/block.f90:1547-2076; restart.dat, main.f90:36-39, 138-175)."""
import os
import struct

import numpy as np
import pytest

import fake_device
from tsc_tri_t1t2j1j2_amd import measure as meas
from tsc_tri_t1t2j1j2_amd import persist
from tsc_tri_t1t2j1j2_amd.blocks import SITE_SECTORS, SuperBasis, TotalBasis, TotalBlock, site_operators
from tsc_tri_t1t2j1j2_amd.dmrg import DMRG


def _rec(body):
    n = struct.pack("<i", len(body))
    return n + body + n


def _i(*v):
    return struct.pack("<" + "i" * len(v), *v)


def test_site_operator_record_is_the_references_block(tmp_path):
    """c+ of one site written by block_to_disk: len 1, num 1, dim 1, up_dif
    1, down_dif 1, then the sub-block acting on the empty site (N 0, 2S 0),
    down_dif +1, 1 x 1, -sqrt2 (st_elec_up, dmrg_sub.f90:742-764), each a
    Fortran record framed by its 4-byte length."""
    p = str(tmp_path / "op")
    w = persist.RecordWriter(p)
    persist.write_total_block(w, site_operators()["c"], list(SITE_SECTORS), 1)
    w.close()
    expect = b"".join([_rec(_i(1)), _rec(_i(1)), _rec(_i(1)), _rec(_i(1)), _rec(_i(1)), _rec(b""),
                       _rec(_i(0, 0, 1)), _rec(_i(1, 1)), _rec(struct.pack("<dd", -np.sqrt(2.0), 0.0))])
    assert open(p, "rb").read() == expect
    length, tb = persist.read_total_block(persist.RecordReader(p), list(SITE_SECTORS))
    assert length == 1 and (tb.up_dif, tb.k2) == (1, 1) and list(tb.blocks) == [(0, 1)]


def test_round_trips(tmp_path):
    rng = np.random.default_rng(3)
    secs = [(2, 2, 3), (2, 0, 2), (1, 1, 4), (0, 0, 1)]
    tb = TotalBlock(1, 1)
    tb.blocks[(0, 2)] = rng.normal(size=(3, 4)) + 1j * rng.normal(size=(3, 4))
    tb.blocks[(2, 3)] = rng.normal(size=(4, 1)) + 1j * rng.normal(size=(4, 1))
    p = str(tmp_path / "blk")
    w = persist.RecordWriter(p)
    persist.write_total_block(w, tb, secs, 5)
    w.close()
    length, back = persist.read_total_block(persist.RecordReader(p), secs)
    # every (bra, ket) pair the rank and particle change allow is written,
    # in ket order and then bra 2S ascending, a zero block where the
    # operator has none
    # (update_trun_ndia's sub-block list)
    assert length == 5 and list(back.blocks) == [(1, 2), (0, 2), (2, 3)]
    assert not back.blocks[(1, 2)].any()
    for k in tb.blocks:
        assert np.array_equal(back.blocks[k], tb.blocks[k])
    # basis text: the formats of basis_to_disk, read back to the same structure
    bs = TotalBasis(secs)
    persist.basis_to_disk(str(tmp_path / "bs"), bs, 6)
    lines = open(tmp_path / "bs").read().split("\n")
    assert lines[:3] == ["len=     6", f"num={bs.num:6d}", f"dim={bs.dim:6d}"]
    length, sectors = persist.basis_from_disk(str(tmp_path / "bs"))
    assert length == 6 and [(N, J, d) for (N, J, d, _) in sectors] == bs.plain_sectors()
    for (N, J, d, subs), (_, _, _, mine) in zip(sectors, bs.sectors):
        assert [(p, sd) for (p, sd, _, _) in subs] == [(p, sd) for (_, _, p, sd) in mine]
        assert [bl for (_, _, bl, _) in subs] == [bs.block_sectors[a][:2] for (a, _, _, _) in mine]
    # truncation: env matrices are stored conjugated and read back as they were
    from tsc_tri_t1t2j1j2_amd.wavetrans import Trunc
    U = {x: rng.normal(size=(d, min(d, 2))) + 1j * rng.normal(size=(d, min(d, 2)))
         for x, (_, _, d, _) in enumerate(bs.sectors) if x % 2 == 0}
    order = sorted(U)
    tr = Trunc(bs, U, order)
    for env in (False, True):
        persist.truns_to_disk(str(tmp_path / "tr"), tr, 6, env=env)
        r = persist.RecordReader(str(tmp_path / "tr"))
        assert [r.int1() for _ in range(5)] == [6, len(order), sum(U[x].shape[1] for x in order), 0, 0]
        r.empty()
        r.ints(3)
        rows, cols = r.ints(2)
        stored = r.matrix(rows, cols)
        assert np.array_equal(stored, np.conj(U[order[0]]) if env else U[order[0]])
        _, back = persist.truns_from_disk(str(tmp_path / "tr"), bs, env=env)
        assert back.order == order and all(np.array_equal(back.U[x], U[x]) for x in order)
    # wave function
    sb = SuperBasis(bs, TotalBasis(secs), 4, 0)
    psi = rng.normal(size=sb.dim) + 1j * rng.normal(size=sb.dim)
    persist.wave_to_disk(str(tmp_path / "eigvec1"), sb, bs, TotalBasis(secs), psi, 4, 5)
    sl, el, subs = persist.wave_from_disk(str(tmp_path / "eigvec1"))
    assert (sl, el, len(subs)) == (4, 5, sb.num)
    assert np.array_equal(persist.psi_in_super_basis(subs, sb, bs, TotalBasis(secs)), psi)
    persist.write_restart(str(tmp_path / "restart.dat"), 1, 7, 0)
    assert persist.read_restart(str(tmp_path / "restart.dat")) == (1, 7, 0)


def _dmrg(ctx):
    return DMRG(ctx, 3, 3, Ne=8, kept_min=4, kept_max=12, sweeps=1, lanzero=1e-10, trun_err=1e-14,
                native_planner=False)


def test_restart_after_an_interrupted_run(monkeypatch, tmp_path):
    """A run stopped after its first left-to-right half-sweep and restarted
    with restart1 = 2 from its files reaches the energy of an uninterrupted
    run (the restart uses the reference's tighter restart tolerance)."""
    ctx = fake_device.install(monkeypatch)
    full = _dmrg(ctx)
    e_full = full.run(checkpoint=str(tmp_path / "a"))
    stop = _dmrg(ctx)

    def crash(*a, **k):
        raise RuntimeError("stopped")

    monkeypatch.setattr(stop, "right_to_left", crash)
    with pytest.raises(RuntimeError):
        stop.run(checkpoint=str(tmp_path / "b"))
    n_done = len(stop.history)
    again = _dmrg(ctx)
    loaded = persist.model_from_disk(str(tmp_path / "b" / persist.block_name("sys", 5)), "sys", list(range(1, 6)),
                                     stop.slot_sites("sys", 5))
    assert loaded.sites == stop.sys_blocks[5].sites
    e = again.run(checkpoint=str(tmp_path / "b"), restart=(2, 0, 0))
    assert again.history[:n_done] == [tuple(h) for h in stop.history]
    assert len(again.history) == len(full.history)
    assert abs(e - e_full) < 1e-6, (e, e_full)
    with pytest.raises(ValueError):
        _dmrg(ctx).run(restart=(1, 0, 0))


def test_measurement_replay_from_files(monkeypatch, tmp_path):
    """restart1 = 4: the measurement of the stored final state and stored
    truncations equals the measurement made inside the run."""
    ctx = fake_device.install(monkeypatch)
    d = DMRG(ctx, 4, 2, Ne=6, kept_min=4, kept_max=10, sweeps=1, lanzero=1e-10, trun_err=1e-14, native_planner=False)
    d.run(checkpoint=str(tmp_path))
    out = d.measure()
    rep = meas.measure_from_disk(ctx, str(tmp_path), 4, 2)
    assert rep["l_end"] == out["l_end"] and rep["js"] == out["js"]
    for kind in ("SS", "nn", "cc"):
        assert rep[kind] == out[kind]
    assert rep["density"] == out["density"]
    assert rep["pairs"] == out["pairs"]
    assert os.path.exists(tmp_path / "sysbs3") and os.path.exists(tmp_path / "envtruns3")


def test_step_logs_in_reference_formats(monkeypatch, tmp_path):
    """Energy_All.dat (eigen.f90:192 format), Kept_State.dat and entropy1.dat
    (den_mat.f90:342-352 formats): one line per solve / per truncation."""
    import re
    ctx = fake_device.install(monkeypatch)
    d = _dmrg(ctx)
    d.run(log_dir=str(tmp_path))
    en = open(tmp_path / "Energy_All.dat").read().splitlines()
    assert len(en) == len(d.history)
    m = re.fullmatch(r"Sys_len=(.{3})  E0=  (.{18})  Iter=(.{4})  time=  (.{12})", en[-1])
    assert m and abs(float(m.group(2)) - d.history[-1][2]) < 1e-11
    ks = open(tmp_path / "Kept_State.dat").read().splitlines()
    m = re.fullmatch(r"Sys_len=(.{3})  Num_State=(.{5})  Err=(.{18})", ks[-1])
    assert m and int(m.group(2)) == d.history[-1][4]
    assert abs(float(m.group(3)) - d.history[-1][3]) <= 1e-8 * abs(d.history[-1][3]) + 1e-300
    ent = open(tmp_path / "entropy1.dat").read().splitlines()
    assert len(ent) == len(ks) and all(len(line) == 8 + 20 + 27 for line in ent)
