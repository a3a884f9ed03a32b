"""Lattice and couplings of the extended t-J model on the YC triangular
cylinder (oracle restatement of hyx15903/TSC_Tri_t1t2J1J2
DMRG_code/dmrg_sub.f90:334-644 `dmrg_setup` for triangle=1, and the
parameters of DMRG_code/main.f90:54-99).

Sites are numbered i = iy + (ix-1)*Ny (1-based, column by column), y is
periodic (openy=0), x is open (open=1). Bond types: 1/2 y-neighbour,
3/4 x-neighbour, 5/6 (x+1, y-1), 7/8 = (y+1 then x+1), 9/10 = (x+1 then
type 5) [only if Ny != 3], 11/12 = (y-1 then type 5); types 1-6 carry
(t1, J1), 7-12 carry (t2, J2). Third neighbours (13-18) are pruned since
their couplings vanish.

H = sum_<ij> [ -t_ij sum_s (c+_is c_js + h.c.) + J_ij (S_i.S_j - n_i n_j / 4) ]
(DMRG_code/common.f90:1-4), no double occupancy.
"""


def neighbour_table(Lx, Ny, ncols=None):
    """neib(i, typ) of dmrg_setup (dmrg_sub.f90:442-575, triangle = 1) as
    {(i, typ): j} for bond types 1-12 (types 13-18 are dropped by the
    reference itself, their couplings being zero, dmrg_sub.f90:634-643)."""
    nxc = Lx if ncols is None else ncols
    neib = {}

    def add(ii, jj, typ):
        neib[(ii, typ)] = jj
        neib[(jj, typ + 1)] = ii

    for ix in range(1, nxc + 1):
        for iy in range(1, Ny + 1):
            ii = iy + (ix - 1) * Ny
            jy = ii - iy + (iy % Ny) + 1
            if iy != Ny or Ny != 2:
                add(ii, jy, 1)
            jx = ii + Ny
            if jx > Ny * nxc:
                jx = 0
            if jx:
                add(ii, jx, 3)
                j5 = jx - 1
                if iy == 1:
                    j5 += Ny
                add(ii, j5, 5)
    for ix in range(1, nxc + 1):
        for iy in range(1, Ny + 1):
            ii = iy + (ix - 1) * Ny
            j1 = neib.get((ii, 1), 0)
            j2 = neib.get((j1, 3), 0) if j1 else 0
            if j1 and j2:
                add(ii, j2, 7)
            if Ny != 3:
                j1 = neib.get((ii, 3), 0)
                j2 = neib.get((j1, 5), 0) if j1 else 0
                if j1 and j2:
                    add(ii, j2, 9)
            j1 = neib.get((ii, 2), 0)
            j2 = neib.get((j1, 5), 0) if j1 else 0
            if j1 and j2:
                add(ii, j2, 11)
    return neib


def bonds(Lx, Ny, t1, t2, J1, J2, ncols=None, present=None):
    """List of (i, j, t, J) with i < j (1-based chain indices).

    ncols: number of columns of the cylinder the neighbour table is built on
    (defaults to Lx); present: optional set of site indices kept (the
    warm-up superblock keeps the first and last sites of a shorter
    cylinder, dmrg_sub.f90:334-370)."""
    neib = neighbour_table(Lx, Ny, ncols)
    out = {}
    for (ii, typ), jj in neib.items():
        if typ % 2 == 0:
            continue
        a, b = min(ii, jj), max(ii, jj)
        if present is not None and (a not in present or b not in present):
            continue
        if typ <= 5:
            t, J = t1, J1
        else:
            t, J = t2, J2
        if t == 0 and J == 0:
            continue
        # a pair can appear through two types only on very narrow cylinders;
        # the reference keeps one neighbour slot per pair (neiba)
        out[(a, b)] = (t, J)
    return sorted((a, b, t, J) for (a, b), (t, J) in out.items())
