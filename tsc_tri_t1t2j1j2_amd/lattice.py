"""YC triangular cylinder and t1-t2-J1-J2 couplings (the neighbour tables
of hyx15903/TSC_Tri_t1t2J1J2 DMRG_code/dmrg_sub.f90:334-644, triangle=1,
open x, periodic y; parameters of DMRG_code/main.f90:54-99).

Sites i = iy + (ix-1)*Ny (1-based). Nearest-neighbour displacements (dx,dy)
= (0,1), (1,0), (1,-1); next-nearest (1,1), (2,-1), (1,-2), with (2,-1)
omitted on Ny = 3 cylinders as in the reference. One coupling per site
pair. The DMRG chain order is the site order.
"""
NN = ((0, 1), (1, 0), (1, -1))
NNN = ((1, 1), (2, -1), (1, -2))


def site(ix, iy, Ny):
    return iy + (ix - 1) * Ny


def bonds(Lx, Ny, t1, t2, J1, J2, ncols=None, present=None):
    """[(i, j, t, J)] with i < j. ncols: columns of the cylinder the table
    is built on (warm-up superblocks use a shorter one); present: subset of
    sites kept."""
    nc = Lx if ncols is None else ncols
    pairs = {}
    for shell, (t, J) in ((NN, (t1, J1)), (NNN, (t2, J2))):
        if t == 0 and J == 0:
            continue
        for (dx, dy) in shell:
            if Ny == 3 and (dx, dy) == (2, -1):
                continue
            if Ny == 2 and (dx, dy) == (0, 1):
                # a 2-leg cylinder keeps a single y bond per rung
                pass
            for ix in range(1, nc + 1):
                for iy in range(1, Ny + 1):
                    jx = ix + dx
                    if jx < 1 or jx > nc:
                        continue
                    jy = (iy - 1 + dy) % Ny + 1
                    a, b = site(ix, iy, Ny), site(jx, jy, Ny)
                    if a == b:
                        continue
                    key = (min(a, b), max(a, b))
                    if key not in pairs:
                        pairs[key] = (t, J)
    out = []
    for (a, b), (t, J) in sorted(pairs.items()):
        if present is not None and (a not in present or b not in present):
            continue
        out.append((a, b, t, J))
    return out


def target_electrons(N, delta=12):
    """qn0(N,1) of main.f90:93-97: electrons for a chain of N sites at hole
    doping 1/delta, rounded to an even number."""
    q = N // 2 - (N + 2 * delta - 1) // (2 * delta)
    return 2 * q


def neighbour(i, typ, Lx, Ny):
    """neib(i, typ) of the reference for the odd bond types (dmrg_sub.f90:
    334-404): 1 the y neighbour (iy+1, periodic; none from the last leg of a
    2-leg cylinder), 3 the x neighbour, 5 the (x+1, y-1) neighbour; 0 when
    the cylinder ends."""
    ix, iy = (i - 1) // Ny + 1, (i - 1) % Ny + 1
    if typ == 1:
        if Ny == 2 and iy == Ny:
            return 0
        return iy % Ny + 1 + (ix - 1) * Ny
    if ix + 1 > Lx:
        return 0
    if typ == 3:
        return iy + ix * Ny
    if typ == 5:
        return (iy - 2) % Ny + 1 + ix * Ny
    raise ValueError(f"bond type {typ} is not one of 1, 3, 5")


def op_slots(bond_list, L):
    """Operator slots of a block of L sites (INB1 / nleg11 of dmrg_setup0,
    dmrg_sub.f90:304-331): the block's sites, numbered 1..L from the end of
    the chain it grows from, that have a neighbour numbered L or higher, in
    increasing order; slot k (1-based) of the block's model file holds the
    operators of the k-th of them. bond_list: bonds() of the full cylinder
    (its couplings decide which neighbour shells exist). The same table
    serves the env block, whose site numbered II from the far end is chain
    site N + 1 - II (the cylinder is symmetric under that map)."""
    reach = {}
    for (a, b, _t, _J) in bond_list:
        reach[a] = max(reach.get(a, 0), b)
        reach[b] = max(reach.get(b, 0), a)
    return [i for i in range(1, L + 1) if reach.get(i, 0) >= L]
