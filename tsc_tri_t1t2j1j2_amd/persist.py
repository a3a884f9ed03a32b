"""Per-length DMRG files in the reference's on-disk layout, and
checkpoint/restart of the driver.

The reference keeps every block, truncation, basis and wave function on
disk, one file per block length (hyx15903/TSC_Tri_t1t2J1J2
DMRG_code/block.f90:1547-2076: wave_to_disk, block_to_disk, model_to_disk,
truns_to_disk, basis_to_disk; names from Get_model_name / Get_wave_name,
DMRG_code/dmrg_sub.f90:657-709), and restarts from them as restart.dat
says (main.f90:36-39, 138-175). The files here have the same names and
record layouts:

* unformatted sequential records (a 4-byte little-endian length before and
  after each record, the layout of gfortran and ifort), default 4-byte
  integers, complex128 matrices in column-major order;
* Total_Block: len | num | dim | up_dif | down_dif, then per sub-block an
  empty record, (ket N, ket 2S, 2S_bra - 2S_ket), (rows, cols), the matrix;
* eigvec<k>: sys_len | env_len | num | dim, then per sector an empty
  record, (sys N, 2S), (env N, 2S), (sys dim, env dim), the psi block;
* systruns<L> / envtruns<L>: one Total_Block, a sub-block per kept enlarged
  sector (its N, 2S; rows = enlarged dim, cols = kept);
* sysbs<L> / envbs<L>: the formatted Total_Basis text.

Numbers follow this build's conventions: env operators act as psi . O^T, so
env operators and env truncation matrices are the complex conjugates of the
reference's; the writers store env truncations conjugated (the reference's
convention) and the readers undo it. Block operator files (sysblocks<L>)
hold this build's operators (c+, S, n with the reduced elements of
blocks.site_operators) in Total_Block records, one triple per site the
block still carries, listed in a sidecar JSON file; they restart this
driver, they are not interchangeable with the reference's model files.
"""
import json
import struct

import numpy as np

from .blocks import SuperBasis, TotalBasis, TotalBlock, TotalModel
from .resident import model_to_host
from .wavetrans import Trunc

NAMES = ("c", "S", "n")


# ---------------------------------------------------------------------
# unformatted sequential records

class RecordWriter:
    def __init__(self, path):
        self.f = open(path, "wb")

    def record(self, *parts):
        """One record: integers (4-byte) and numpy arrays, concatenated."""
        chunks = []
        for p in parts:
            if isinstance(p, (int, np.integer)):
                chunks.append(struct.pack("<i", int(p)))
            else:
                a = np.asarray(p)
                if a.dtype == np.complex128:
                    chunks.append(np.asfortranarray(a).tobytes(order="F"))
                else:
                    chunks.append(np.asarray(a, np.int32).tobytes())
        body = b"".join(chunks)
        n = struct.pack("<i", len(body))
        self.f.write(n + body + n)

    def close(self):
        self.f.close()


class RecordReader:
    def __init__(self, path):
        with open(path, "rb") as fh:
            self.data = fh.read()
        self.pos = 0

    def record(self):
        (n,) = struct.unpack_from("<i", self.data, self.pos)
        body = self.data[self.pos + 4:self.pos + 4 + n]
        (n2,) = struct.unpack_from("<i", self.data, self.pos + 4 + n)
        if n2 != n:
            raise ValueError(f"corrupt record at byte {self.pos}: markers {n} and {n2}")
        self.pos += 8 + n
        return body

    def ints(self, count=None):
        b = self.record()
        v = [int(x) for x in np.frombuffer(b, "<i4")]
        if count is not None and len(v) != count:
            raise ValueError(f"expected {count} integers, record holds {len(v)}")
        return v

    def int1(self):
        return self.ints(1)[0]

    def matrix(self, rows, cols):
        b = self.record()
        if len(b) != 16 * rows * cols:
            raise ValueError(f"matrix record of {len(b)} bytes for {rows} x {cols}")
        return np.frombuffer(b, np.complex128).reshape((rows, cols), order="F").copy()

    def empty(self):
        if self.record():
            raise ValueError("expected an empty record")


# ---------------------------------------------------------------------
# Total_Block

def _dense(m):
    return np.asarray(m, np.complex128)


def write_total_block(w, tb, sectors, length):
    """block_to_disk (block.f90:1715-1737) of an operator on a block with
    sectors [(N, 2S, dim)]; sub-blocks in (ket, bra) order."""
    keys = sorted(tb.blocks, key=lambda k: (k[1], k[0]))
    w.record(int(length))
    w.record(len(keys))
    w.record(int(sum(sectors[k][2] for (_, k) in keys)))
    w.record(int(tb.up_dif))
    w.record(int(tb.k2))
    for (b, k) in keys:
        m = _dense(tb.blocks[(b, k)])
        w.record()
        w.record(int(sectors[k][0]), int(sectors[k][1]), int(sectors[b][1] - sectors[k][1]))
        w.record(int(m.shape[0]), int(m.shape[1]))
        w.record(m)


def read_total_block(r, sectors):
    """block_from_disk for a block with sectors [(N, 2S, dim)]; returns
    (length, TotalBlock)."""
    index = {(s[0], s[1]): i for i, s in enumerate(sectors)}
    length, num, _dim, up_dif, k2 = (r.int1() for _ in range(5))
    tb = TotalBlock(up_dif, k2)
    for _ in range(num):
        r.empty()
        n_up, n_down, ddif = r.ints(3)
        rows, cols = r.ints(2)
        k = index[(n_up, n_down)]
        b = index[(n_up + up_dif, n_down + ddif)]
        tb.blocks[(b, k)] = r.matrix(rows, cols)
    return length, tb


# ---------------------------------------------------------------------
# wave function, truncations, bases

def wave_to_disk(path, super_basis, sys_bs, env_bs, psi, sys_len, env_len):
    """wave_to_disk (block.f90:1547-1576): psi (host vector in the super
    basis layout) of a superblock whose sys part has sys_len sites (block +
    site) and env part env_len."""
    w = RecordWriter(path)
    try:
        w.record(int(sys_len))
        w.record(int(env_len))
        w.record(super_basis.num)
        w.record(int(super_basis.dim))
        psi = np.asarray(psi, np.complex128)
        for i, (x, y, dx, dy) in enumerate(super_basis.sub):
            w.record()
            w.record(sys_bs.sectors[x][0], sys_bs.sectors[x][1])
            w.record(env_bs.sectors[y][0], env_bs.sectors[y][1])
            w.record(dx, dy)
            o = super_basis.offset[i]
            w.record(psi[o:o + dx * dy].reshape((dx, dy), order="F"))
    finally:
        w.close()


def wave_from_disk(path):
    """(sys_len, env_len, [(sys QN, env QN, block)]) of an eigvec file."""
    r = RecordReader(path)
    sys_len, env_len, num, _dim = (r.int1() for _ in range(4))
    subs = []
    for _ in range(num):
        r.empty()
        sq = tuple(r.ints(2))
        eq = tuple(r.ints(2))
        dx, dy = r.ints(2)
        subs.append((sq, eq, r.matrix(dx, dy)))
    return sys_len, env_len, subs


def psi_in_super_basis(subs, super_basis, sys_bs, env_bs):
    """The host psi vector of super_basis from wave_from_disk's sectors."""
    psi = np.zeros(super_basis.dim, np.complex128)
    got = {(sq, eq): m for (sq, eq, m) in subs}
    if len(got) != super_basis.num:
        raise ValueError("stored wave function has another super basis")
    for i, (x, y, dx, dy) in enumerate(super_basis.sub):
        m = got[((sys_bs.sectors[x][0], sys_bs.sectors[x][1]), (env_bs.sectors[y][0], env_bs.sectors[y][1]))]
        if m.shape != (dx, dy):
            raise ValueError("stored wave function sector of another shape")
        o = super_basis.offset[i]
        psi[o:o + dx * dy] = m.ravel(order="F")
    return psi


def truns_to_disk(path, trun, length, env=False):
    """truns_to_disk (block.f90:1935-1951) of the truncation that built a
    block of `length` sites; env truncations are stored conjugated."""
    w = RecordWriter(path)
    try:
        bs = trun.basis
        w.record(int(length))
        w.record(len(trun.order))
        w.record(int(sum(trun.U[x].shape[1] for x in trun.order)))
        w.record(0)
        w.record(0)
        for x in trun.order:
            u = np.asarray(trun.U[x], np.complex128)
            w.record()
            w.record(bs.sectors[x][0], bs.sectors[x][1], 0)
            w.record(u.shape[0], u.shape[1])
            w.record(np.conj(u) if env else u)
    finally:
        w.close()


def truns_from_disk(path, basis, env=False):
    """A Trunc over `basis` (the enlarged basis the truncation acts on)."""
    r = RecordReader(path)
    length, num, _dim, _ud, _dd = (r.int1() for _ in range(5))
    U, order = {}, []
    for _ in range(num):
        r.empty()
        n_up, n_down, _ = r.ints(3)
        rows, cols = r.ints(2)
        x = basis.index[(n_up, n_down)]
        if rows != basis.sectors[x][2]:
            raise ValueError(f"truncation of sector {(n_up, n_down)} has {rows} rows, the basis {basis.sectors[x][2]}")
        u = r.matrix(rows, cols)
        U[x] = np.conj(u) if env else u
        order.append(x)
    return length, Trunc(basis, U, order)


def basis_to_disk(path, basis, length):
    """basis_to_disk (block.f90:1979-2022), formats (A4,I6), (A11,I4),
    (A13,I4), (A10,I4,2X,A12,I4)."""
    lines = [f"len={length:6d}", f"num={basis.num:6d}", f"dim={basis.dim:6d}"]
    for (N, J, d, subs) in basis.sectors:
        lines += ["", f"idx={len(subs):6d}", f"num={len(subs):6d}", f"dim={d:6d}", f"new_num_up={N:4d}",
                  f"new_num_down={J:4d}"]
        for (a, s, p, sd) in subs:
            bn, bj, _ = basis.block_sectors[a]
            sn, sj, _ = basis.site_sectors[s]
            lines += [f"pos={p:6d}", f"dim={sd:6d}", f"bl_num_up={bn:4d}  bl_num_down={bj:4d}",
                      f"st_num_up={sn:4d}  st_num_down={sj:4d}"]
    with open(path, "w") as fh:
        fh.write("\n".join(lines) + "\n")


def basis_from_disk(path):
    """(length, [(N, 2S, dim, [(spos, sdim, block QN, site QN)])])."""
    with open(path) as fh:
        lines = fh.read().split("\n")
    pos = 0

    def take(width):
        nonlocal pos
        line = lines[pos]
        pos += 1
        return int(line[width:])

    length, num, _dim = take(4), take(4), take(4)
    out = []
    for _ in range(num):
        pos += 1       # blank record
        _idx, nsub, d = take(4), take(4), take(4)
        N, J = take(11), take(13)
        subs = []
        for _ in range(nsub):
            p, sd = take(4), take(4)
            bl, st = lines[pos], lines[pos + 1]
            pos += 2
            subs.append((p, sd, (int(bl[10:14]), int(bl[28:32])), (int(st[10:14]), int(st[28:32]))))
        out.append((N, J, d, subs))
    return length, out


# ---------------------------------------------------------------------
# blocks of this driver

def model_to_disk(path, model):
    """The block Hamiltonian and the (c+, S, n) operators of every site the
    block carries, as Total_Block records (model_to_disk's layout,
    block.f90:1802-1844); the sites and sectors in path + '.json'."""
    m = model_to_host(model)
    w = RecordWriter(path)
    try:
        w.record(len(m.sites))
        write_total_block(w, m.ham, m.sectors, len(m.sites))
        op_sites = sorted(m.ops)
        for i in op_sites:
            for name in NAMES:
                write_total_block(w, m.ops[i][name], m.sectors, len(m.sites))
    finally:
        w.close()
    with open(path + ".json", "w") as fh:
        json.dump({"sites": m.sites, "sectors": [list(s) for s in m.sectors], "op_sites": op_sites}, fh)


def model_from_disk(path):
    with open(path + ".json") as fh:
        meta = json.load(fh)
    sectors = [tuple(s) for s in meta["sectors"]]
    r = RecordReader(path)
    length = r.int1()
    if length != len(meta["sites"]):
        raise ValueError("block file and its site list disagree")
    _, ham = read_total_block(r, sectors)
    ops = {}
    for i in meta["op_sites"]:
        ops[i] = {}
        for name in NAMES:
            _, ops[i][name] = read_total_block(r, sectors)
    return TotalModel(meta["sites"], sectors, ham, ops)


# ---------------------------------------------------------------------
# file names (Get_model_name, Get_wave_name)

def block_name(side, length):
    return f"{side}blocks{length}"


def trun_name(side, length):
    return f"{side}truns{length}"


def basis_name(side, length):
    return f"{side}bs{length}"


def wave_name(k=1):
    return f"eigvec{k}"


def read_restart(path):
    """restart.dat: 'restart1 sys_len_start' then 'restart2' (main.f90:36-38)."""
    with open(path) as fh:
        a = fh.readline().split()
        b = fh.readline().split()
    return int(a[0]), int(a[1]) if len(a) > 1 else 0, int(b[0]) if b else 0


def write_restart(path, restart1, sys_len_start, restart2=0):
    with open(path, "w") as fh:
        fh.write(f"{restart1} {sys_len_start}\n{restart2}\n")


__all__ = ["RecordWriter", "RecordReader", "write_total_block", "read_total_block", "wave_to_disk", "wave_from_disk",
           "psi_in_super_basis", "truns_to_disk", "truns_from_disk", "basis_to_disk", "basis_from_disk",
           "model_to_disk", "model_from_disk", "read_restart", "write_restart", "SuperBasis", "TotalBasis"]
