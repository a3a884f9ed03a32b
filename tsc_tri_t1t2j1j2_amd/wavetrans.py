"""Wave-function guess between sweep steps (wave_trans_LR / wave_trans_RL,
hyx15903/TSC_Tri_t1t2J1J2 DMRG_code/dmrg_sub.f90:1471-1633, 1637-1789).

After a step has truncated one side, the ground state psi of that step is
carried into the basis of the next superblock, where the block has grown by
one site and the other block has shrunk by one:

* left to right: the sys side is truncated, U_sys^H psi; the env side is
  expanded, psi U_env^T (the env truncation that built the old env block),
  and the site between moves from the env to the sys side;
* right to left: the mirror image, U_sys psi on the sys side and psi
  conj(U_env) on the env side.

Each (old sector, site sub-block, new total spin) gives two GEMMs and a
recoupling factor of the moved site, so the whole transform is one grouped
two-stage plan: stage 1 applies the sys-side matrix to a window of the old
psi (shared by the new spins), stage 2 applies the env-side matrix with the
recoupling factor and writes a window of the new psi. The result is
normalised afterwards (wave_normalize).

Conventions: env operators act as psi . O^T here, and env truncation
matrices are the conjugates of the reference's (den_mat.f90:89), so the
reference's U_env^H (LR) and U_env (RL) become U_env^T and conj(U_env).
"""
import math

import numpy as np

from . import capi
from .blocks import SITE_SECTORS, SuperBasis, TotalBasis
from .planner import Operand, PlanBuilder


def _neg(e):
    """(-1)**e for an integer e."""
    return -1.0 if e % 2 else 1.0


def coef_lr(j12, j3, j123, j4, j34):
    """Recoupling factor of wave_trans_LR (dmrg_sub.f90:1583-1593) for a
    singlet superblock: j12 old sys spin, j3 site, j123 new sys spin, j4 new
    env spin, j34 old env spin (all doubled)."""
    c = _neg((j12 + j3 + j123) // 2) / math.sqrt((1.0 + j12) * (1.0 + j123))
    c *= math.sqrt((1.0 + j123) * (1.0 + j34))
    if ((j12 + j3 + j4) // 2) % 2:
        c = -c
    if ((j3 + j4 - j34) // 2) % 2:
        c = -c
    return c


def coef_rl(j12, j3, j123, j4, j34):
    """Recoupling factor of wave_trans_RL (dmrg_sub.f90:1734-1743): j12 old
    env spin, j3 site, j123 new env spin, j4 new sys spin, j34 old sys spin."""
    c = _neg((j34 + j4 + j3) // 2) / math.sqrt((1.0 + j12) * (1.0 + j4))
    c *= math.sqrt((1.0 + j123) * (1.0 + j34))
    if ((j4 + j3 + j12) // 2) % 2:
        c = -c
    if ((j12 + j3 - j123) // 2) % 2:
        c = -c
    return c


class Trunc:
    """A truncation that built a block: U[x] (d_x x kept_x) for the enlarged
    sectors x of `basis` that were kept, in the block's sector order `order`."""

    def __init__(self, basis, U, order):
        self.basis = basis
        self.U = {x: np.asarray(U[x], np.complex128) for x in order}
        self.order = list(order)
        # block sector index -> enlarged sector
        self.of_block = {a: x for a, x in enumerate(self.order)}


def plan_lr(old_super, old_sys_bs, old_env_bs, sys_trun, env_trun, new_super, new_sys_bs, new_env_bs):
    """PlanBuilder of wave_trans_LR. old_sys_bs / old_env_bs: enlarged bases
    of the old superblock; sys_trun: truncation of old_sys_bs that built the
    new sys block (the new sys_bs is TotalBasis of that block); env_trun:
    truncation of new_env_bs that built the old env block."""
    pb = PlanBuilder()
    uref = {}

    def umat(side, x, t):
        key = (side, x)
        if key not in uref:
            uref[key] = pb.op(key, t.U[x])
        return uref[key]

    new_sys_block_index = {x: a for a, x in enumerate(sys_trun.order)}
    for i, (xo, yo, dxo, dyo) in enumerate(old_super.sub):
        if xo not in sys_trun.U:
            continue
        u_sys = sys_trun.U[xo]
        knew = u_sys.shape[1]
        a_new = new_sys_block_index[xo]
        j12 = old_sys_bs.sectors[xo][1]
        n_sys_old = old_sys_bs.sectors[xo][0]
        j34 = old_env_bs.sectors[yo][1]
        off_i = old_super.offset[i]
        for (a_env, s, spos, sdim) in old_env_bs.sectors[yo][3]:
            # a_env: sector of the old env block = env_trun.order[a_env]
            xe = env_trun.of_block.get(a_env)
            if xe is None:
                continue
            u_env = env_trun.U[xe]
            n_st, j3, _ = SITE_SECTORS[s]
            j4 = new_env_bs.sectors[xe][1]
            uo, ul = umat("s", xo, sys_trun)
            vo, vl = umat("e", xe, env_trun)
            # stage 1: Z = U_sys[xo]^H psi_i[:, spos:spos+sdim]  (knew x sdim)
            z = pb.intermediate(("lr", i, spos), Operand(capi.BUF_OPS, uo, ul, capi.OP_C),
                                Operand(capi.BUF_IN, off_i + spos * dxo, max(dxo, 1), capi.OP_N), knew, sdim, dxo)
            for j123 in range(abs(j12 - j3), j12 + j3 + 1, 2):
                xs = new_sys_bs.index.get((n_sys_old + n_st, j123))
                if xs is None:
                    continue
                npos = None
                for (a, s2, p, d) in new_sys_bs.sectors[xs][3]:
                    if a == a_new and s2 == s:
                        npos, nsd = p, d
                        break
                if npos is None:
                    continue
                inew = new_super.index.get((xs, xe))
                if inew is None:
                    continue
                dxs = new_sys_bs.sectors[xs][2]
                dxe = new_env_bs.sectors[xe][2]
                assert nsd == knew and u_env.shape == (dxe, sdim)
                # the site's fermion moves past the env block's (env basis
                # vectors here are (block ops)(site op)|0>)
                c = coef_lr(j12, j3, j123, j4, j34) * _neg(n_st * new_env_bs.sectors[xe][0])
                wid = pb.out_window(("lr", inew, npos), new_super.offset[inew] + npos, dxs, knew, dxe)
                # stage 2: psi_new[s_new][npos:npos+knew, :] += c Z U_env[xe]^T
                pb.add(wid, z, Operand(capi.BUF_OPS, vo, vl, capi.OP_T), sdim, c)
    return pb


def plan_rl(old_super, old_sys_bs, old_env_bs, sys_trun, env_trun, new_super, new_sys_bs, new_env_bs):
    """PlanBuilder of wave_trans_RL. env_trun: truncation of old_env_bs that
    built the new env block (the new env_bs is TotalBasis of that block);
    sys_trun: truncation of new_sys_bs that built the old sys block."""
    pb = PlanBuilder()
    uref = {}

    def umat(side, x, t):
        key = (side, x)
        if key not in uref:
            uref[key] = pb.op(key, t.U[x])
        return uref[key]

    new_env_block_index = {x: a for a, x in enumerate(env_trun.order)}
    for i, (xo, yo, dxo, dyo) in enumerate(old_super.sub):
        if yo not in env_trun.U:
            continue
        u_env = env_trun.U[yo]
        knew = u_env.shape[1]
        a_new = new_env_block_index[yo]
        j12 = old_env_bs.sectors[yo][1]
        n_env_old = old_env_bs.sectors[yo][0]
        j34 = old_sys_bs.sectors[xo][1]
        off_i = old_super.offset[i]
        for (a_sys, s, spos, sdim) in old_sys_bs.sectors[xo][3]:
            xs = sys_trun.of_block.get(a_sys)
            if xs is None:
                continue
            u_sys = sys_trun.U[xs]
            n_st, j3, _ = SITE_SECTORS[s]
            j4 = new_sys_bs.sectors[xs][1]
            uo, ul = umat("s", xs, sys_trun)
            vo, vl = umat("e", yo, env_trun)
            dxs = new_sys_bs.sectors[xs][2]
            # stage 1: Z = U_sys[xs] psi_i[spos:spos+sdim, :]  (dxs x dyo)
            z = pb.intermediate(("rl", i, spos), Operand(capi.BUF_OPS, uo, ul, capi.OP_N),
                                Operand(capi.BUF_IN, off_i + spos, max(dxo, 1), capi.OP_N), dxs, dyo, sdim)
            for j123 in range(abs(j12 - j3), j12 + j3 + 1, 2):
                ye = new_env_bs.index.get((n_env_old + n_st, j123))
                if ye is None:
                    continue
                npos = None
                for (a, s2, p, d) in new_env_bs.sectors[ye][3]:
                    if a == a_new and s2 == s:
                        npos, nsd = p, d
                        break
                if npos is None:
                    continue
                inew = new_super.index.get((xs, ye))
                if inew is None:
                    continue
                assert nsd == knew and u_sys.shape == (dxs, sdim)
                # the site's fermion moves past those of the old env sector
                c = coef_rl(j12, j3, j123, j4, j34) * _neg(n_st * n_env_old)
                wid = pb.out_window(("rl", inew, npos), new_super.offset[inew] + npos * dxs, dxs, dxs, knew)
                # stage 2: psi_new[s_new][:, npos:npos+knew] += c Z conj(U_env[yo])
                pb.add(wid, z, Operand(capi.BUF_OPS, vo, vl, capi.OP_CONJ), dyo, c)
    return pb


def transform(ctx, direction, psi_old, old_sb, new_sb, sys_trun, env_trun, prec=capi.PREC_F64):
    """The normalised guess for new_sb (a device vector) from psi_old of
    old_sb. direction 'LR': sys_trun built new_sb's sys block from old_sb's
    sys_bs, env_trun built old_sb's env block from new_sb's env_bs; 'RL':
    the mirror. Returns None when no sector of the new superblock is reached
    (the caller then starts from a random vector, as the reference does)."""
    mk = plan_lr if direction == "LR" else plan_rl
    pb = mk(old_sb.super, old_sb.sys_bs, old_sb.env_bs, sys_trun, env_trun, new_sb.super, new_sb.sys_bs,
            new_sb.env_bs)
    if not pb.out_list:
        return None
    plan = pb.build(ctx, old_sb.super.dim, new_sb.super.dim)
    out = ctx.vec(new_sb.super.dim)
    plan.apply(psi_old, out, prec)
    ctx.sync()
    plan.close()
    nrm = out.norm()
    if not nrm > 0.0:
        out.close()
        return None
    out.axpby(0.0, out, 1.0 / nrm)
    return out


__all__ = ["Trunc", "coef_lr", "coef_rl", "plan_lr", "plan_rl", "transform", "TotalBasis", "SuperBasis"]
