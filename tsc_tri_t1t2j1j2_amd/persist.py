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
* sysbs<L> / envbs<L>: the formatted Total_Basis text;
* sysblocks<L> / envblocks<L> (model_to_disk): len, the block Hamiltonian,
  then per operator slot (INB1 / nleg11 of dmrg_setup0, lattice.op_slots)
  elec_up, elec_down, spin_sd, num_sn.

Contents are the reference's numbers, converted from this build's
conventions where they differ: env operators here act as psi . O^T, so env
operators and env truncation matrices are the complex conjugates of the
reference's, and an enlarged env basis here orders its fermion modes block
then site where the reference puts the site first, which flips the sign of
the product states with an odd block and an odd site (env_signs): env
truncation rows and the env columns of a wave function carry that sign.
Writers apply both conversions and readers undo them. The reference's
model files also hold elec_down, which this build derives from c+
(elec_down), and slots for sites whose last neighbour is the block's last
site; a checkpointed run carries those sites (DMRG.slot_sites).
"""
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
    sectors [(N, 2S, dim)]. Sub-blocks in the order update_trun_dia /
    update_trun_ndia leave them (block.f90:3465-3620): for each ket sector
    in block order, the bra sectors of 2S from 2S_ket - 2k to 2S_ket + 2k
    that exist, a zero matrix where the operator has no such block."""
    index = {(s[0], s[1]): i for i, s in enumerate(sectors)}
    keys = []
    for k, (n, j2, _d) in enumerate(sectors):
        for jb in range(j2 - tb.k2, j2 + tb.k2 + 1, 2):
            b = index.get((n + tb.up_dif, jb))
            if b is not None:
                keys.append((b, k))
    extra = set(tb.blocks) - set(keys)
    if extra:
        raise ValueError(f"operator blocks {sorted(extra)} do not match its sectors")
    w.record(int(length))
    w.record(len(keys))
    w.record(int(sum(sectors[k][2] for (_, k) in keys)))
    w.record(int(tb.up_dif))
    w.record(int(tb.k2))
    for (b, k) in keys:
        m = tb.blocks.get((b, k))
        m = np.zeros((sectors[b][2], sectors[k][2]), np.complex128) if m is None else _dense(m)
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

def env_signs(basis):
    """Per enlarged-env sector x, the +-1 vector over its rows that takes
    this build's env basis to the reference's: (-1)^(N_block N_site).

    This build orders the fermion modes of an enlarged env as block then
    site, like the sys side; the reference puts the site first (its
    update_block_ndia_env gives block operators the sign (-1)^N_site and
    update_site_ndia_env gives site operators none, block.f90:2828-3459,
    against (-1)^N_block on site operators in update_site_ndia_sys). The
    two product states differ by exactly that sign; everything else (the
    6j embedding, the coupling order block (x) site) is the same, so
    renormalised env operators agree and only the rows of an env
    truncation and the env columns of a wave function change."""
    out = []
    for (_N, _J, d, subs) in basis.sectors:
        v = np.ones(d)
        for (a, st, p, sd) in subs:
            if basis.block_sectors[a][0] % 2 and basis.site_sectors[st][0] % 2:
                v[p:p + sd] = -1.0
        out.append(v)
    return out


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
        sg = env_signs(env_bs)
        for i, (x, y, dx, dy) in enumerate(super_basis.sub):
            w.record()
            w.record(sys_bs.sectors[x][0], sys_bs.sectors[x][1])
            w.record(env_bs.sectors[y][0], env_bs.sectors[y][1])
            w.record(dx, dy)
            o = super_basis.offset[i]
            w.record(psi[o:o + dx * dy].reshape((dx, dy), order="F") * sg[y][None, :])
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
    sg = env_signs(env_bs)
    got = {(sq, eq): m for (sq, eq, m) in subs}
    if len(got) != super_basis.num:
        raise ValueError("stored wave function has another super basis")
    for i, (x, y, dx, dy) in enumerate(super_basis.sub):
        m = got[((sys_bs.sectors[x][0], sys_bs.sectors[x][1]), (env_bs.sectors[y][0], env_bs.sectors[y][1]))]
        if m.shape != (dx, dy):
            raise ValueError("stored wave function sector of another shape")
        o = super_basis.offset[i]
        psi[o:o + dx * dy] = (m * sg[y][None, :]).ravel(order="F")
    return psi


def truns_to_disk(path, trun, length, env=False):
    """truns_to_disk (block.f90:1935-1951) of the truncation that built a
    block of `length` sites; env truncations are stored conjugated and with
    the rows in the reference's fermion order (env_signs)."""
    w = RecordWriter(path)
    try:
        bs = trun.basis
        sg = env_signs(bs) if env else None
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
            w.record(sg[x][:, None] * np.conj(u) if env else u)
    finally:
        w.close()


def truns_from_disk(path, basis, env=False):
    """A Trunc over `basis` (the enlarged basis the truncation acts on)."""
    r = RecordReader(path)
    length, num, _dim, _ud, _dd = (r.int1() for _ in range(5))
    sg = env_signs(basis) if env else None
    U, order = {}, []
    for _ in range(num):
        r.empty()
        n_up, n_down, _ = r.ints(3)
        rows, cols = r.ints(2)
        x = basis.index[(n_up, n_down)]
        if rows != basis.sectors[x][2]:
            raise ValueError(f"truncation of sector {(n_up, n_down)} has {rows} rows, the basis {basis.sectors[x][2]}")
        u = r.matrix(rows, cols)
        U[x] = sg[x][:, None] * np.conj(u) if env else u
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
# block (model) files

def _par(x2):
    """(-1)^(x2/2) for an even doubled exponent."""
    return -1.0 if (x2 // 2) % 2 else 1.0


def elec_down(c, sectors):
    """The reference's elec_down (the rank-1/2 annihilation tensor c~,
    with c~_{+1/2} = -c_down and c~_{-1/2} = c_up, st_elec_down of
    dmrg_sub.f90:779-791) from its elec_up c+ on a block with sectors
    [(N, 2S, dim)]. With the Wigner-Eckart convention <J'M'|T^k_q|JM> =
    (-1)^(J'-M') (J' k J; -M' q M) <J'||T||J> and c~_q = (-1)^(k+q)
    (c+_{-q})^dagger, a symmetry of the 3j symbol gives
    <J'||c~||J> = (-1)^(3/2 + J - J') conj(<J||c+||J'>)."""
    out = TotalBlock(-c.up_dif, c.k2)
    for (b, k), m in c.blocks.items():
        # c+ block <b||c+||k>: c~ block with bra k, ket b
        out.blocks[(k, b)] = _par(sectors[b][1] - sectors[k][1] + 3) * np.conj(_dense(m)).T
    return out


def _conj_block(tb):
    return TotalBlock(tb.up_dif, tb.k2, {k: np.conj(_dense(v)) for k, v in tb.blocks.items()})


def model_to_disk(path, model, side, slots):
    """model_to_disk (block.f90:1802-1844) in the reference's layout and
    conventions: len; the block Hamiltonian; then for each operator slot
    (slots: chain sites in slot order, lattice.op_slots) the slot site's
    elec_up (c+), elec_down (c~), spin_sd (S) and num_sn (n), the four
    Total_Blocks model_to_disk writes for the t-J model (tjmodel = 1,
    v123 = 1, no ring terms). Env blocks are written conjugated (this
    build's env operators act as psi . O^T). Sites the block carries
    beyond the slots (partners of measured pairs) follow as a record
    holding their count and chain numbers and their (c+, S, n) blocks; the
    reference's reader stops after the slots and never sees them."""
    m = model_to_host(model)
    env = side == "env"
    missing = [i for i in slots if i not in m.ops]
    if missing:
        raise ValueError(f"{side} block of {m.len} sites carries no operators of slot sites {missing}")
    conv = _conj_block if env else (lambda tb: tb)
    L = len(m.sites)
    w = RecordWriter(path)
    try:
        w.record(L)
        write_total_block(w, conv(m.ham), m.sectors, L)
        for i in slots:
            up = conv(m.ops[i]["c"])
            for tb in (up, elec_down(up, m.sectors), conv(m.ops[i]["S"]), conv(m.ops[i]["n"])):
                write_total_block(w, tb, m.sectors, L)
        extra = sorted(set(m.ops) - set(slots))
        if extra:
            w.record(len(extra), *extra)
            for i in extra:
                for name in NAMES:
                    write_total_block(w, conv(m.ops[i][name]), m.sectors, L)
    finally:
        w.close()


def model_from_disk(path, side, sites, slots):
    """model_from_disk (block.f90:1848-1898) of a file in the reference's
    layout, as a TotalModel: sectors from the Hamiltonian's sub-blocks,
    the slot sites' c+, S, n (elec_down is implied by c+), and the extra
    sites model_to_disk appends. sites: the block's chain sites in this
    build's fermion order (sys 1..L, env N, N-1, ..); slots: its slot sites
    (lattice.op_slots)."""
    env = side == "env"
    conv = _conj_block if env else (lambda tb: tb)
    r = RecordReader(path)
    L = r.int1()
    if L != len(sites):
        raise ValueError(f"block file of {L} sites, expected {len(sites)}")
    # the Hamiltonian's header and sub-blocks give the sectors
    _len, num, _dim, up_dif, k2 = (r.int1() for _ in range(5))
    if (up_dif, k2) != (0, 0):
        raise ValueError("the first block of a model file is not a scalar")
    sectors, hblocks = [], {}
    for x in range(num):
        r.empty()
        n, j2, dd = r.ints(3)
        rows, cols = r.ints(2)
        if dd != 0 or rows != cols:
            raise ValueError("Hamiltonian sub-block off the diagonal")
        sectors.append((n, j2, rows))
        hblocks[(x, x)] = r.matrix(rows, cols)
    ham = conv(TotalBlock(0, 0, hblocks))
    ops = {}
    for i in slots:
        _, up = read_total_block(r, sectors)
        read_total_block(r, sectors)          # elec_down
        _, S = read_total_block(r, sectors)
        _, n = read_total_block(r, sectors)
        ops[i] = {"c": conv(up), "S": conv(S), "n": conv(n)}
    if r.pos < len(r.data):
        extra = r.ints()
        for i in extra[1:1 + extra[0]]:
            ops[i] = {name: conv(read_total_block(r, sectors)[1]) for name in NAMES}
    return TotalModel(list(sites), sectors, ham, ops)


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
