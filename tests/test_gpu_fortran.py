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


EXE2 = os.path.join(os.path.dirname(capi.LIB_PATH), "fortran", "test_dmrg_boundary")


@pytest.mark.skipif(not os.path.exists(EXE2), reason="Fortran drop-in program not built (no flang at build time)")
def test_fortran_drop_in_subroutines_reproduce_ed(tmp_path):
    """A Fortran host program over the drop-in subroutines (reference names
    and argument lists): blocks grown from single sites with
    Get_Hamiltonian_Sys/_Env, the superblock solved by lanczos_memo_new,
    <psi|H psi> from model_wave, update_trun_dia with the identity, and a
    second superblock after Get_density_matrix_sys / Get_truncation_operator
    (every state kept) / Get_Hamiltonian_Sys. YC4 2x4 cluster (8 sites,
    6 electrons, S=0): both energies equal exact diagonalisation."""
    from oracle import tj_ed, tj_model
    Lx, Ny, Ne = 2, 4, 6
    N = Lx * Ny
    b = tj_model.bonds(Lx, Ny, 1.0, 0.2, 0.4, 0.016)
    e_ed = tj_ed.ground_energy(N, Ne, b)
    inp = tmp_path / "bonds.txt"
    inp.write_text(f"{N} {Ne} {len(b)}\n" + "".join(f"{i} {j} {t!r} {J!r}\n" for (i, j, t, J) in b))
    dump = tmp_path / "blocks.txt"
    out = subprocess.run([EXE2, str(inp), str(dump)], capture_output=True, text=True, timeout=120)
    print(out.stdout, out.stderr)
    assert out.returncode == 0
    vals = {}
    for line in out.stdout.splitlines():
        parts = line.split()
        if parts:
            vals[parts[0]] = [float(x) for x in parts[1:]]
    assert vals["LENS"][:2] == [N // 2 - 1, N // 2 - 1]
    assert vals["UOU_IDENTITY"][0] < 1e-13
    e1, e2 = vals["E1"][1], vals["E2"][1]
    assert abs(e1 - e_ed) / N < 1e-10, (e1, e_ed)
    assert abs(vals["RQ"][0] - e1) < 1e-9 and vals["RQ"][1] < 1e-12
    assert abs(e2 - e_ed) / N < 1e-10, (e2, e_ed)
    _blocks_are_the_references(dump.read_text(), N, b)


def _blocks_are_the_references(text, N, bond_list):
    """The untruncated 3-site sys and env blocks the drop-in subroutines
    built hold the reference's numbers: each sub-block of the Hamiltonian and
    of every carried site's c+ (elec_up), S (spin_sd), n (num_sn) equals the
    reduced matrix of the physical operator between the reference's block
    states (oracle/ref_blocks.py: block (x) site coupling, chain fermion
    order, identity truncations), conjugated on the env side as the
    reference stores env operators."""
    import numpy as np
    pytest.importorskip("sympy")
    from oracle import ref_blocks as RB
    subs = {"sys": [], "env": []}
    for line in text.splitlines():
        f = line.split()
        if f and f[0] == "SUB":
            side, name, site, up = f[1], f[2], int(f[3]), int(f[4])
            nu, nd, dd, rows, cols = (int(v) for v in f[5:10])
            vals = np.array([float(v) for v in f[10:]])
            m = (vals[0::2] + 1j * vals[1::2]).reshape((rows, cols), order="F")
            subs[side].append((name, site, up, nu, nd, dd, m))
    seen = 0
    for side in ("sys", "env"):
        env = side == "env"
        st = RB.BlockStates.site()
        for _ in range(2):
            eb = st.enlarge(env)
            st = eb.truncate([(x, np.eye(d)) for x, (_, _, d) in enumerate(eb.sectors)])
        sites = [1, 2, 3] if not env else [N - 2, N - 1, N]
        local = {s: k for k, s in enumerate(sites)}
        F, ops = RB.fock_ops(3)
        inner = [(local[a], local[c], t, J) for (a, c, t, J) in bond_list if a in local and c in local]
        want_of = {"H": (0, 0, RB.block_hamiltonian(F, inner))}
        for (name, site, up, nu, nd, dd, m) in subs[side]:
            if name == "H":
                k, dN, comps = want_of["H"]
            else:
                fock, dN, k = {"c": ("c+", 1, 1), "S": ("S", 0, 2), "n": ("n", 0, 0)}[name]
                comps = ops[local[site]][fock][1]
            assert up == dN
            want, dev = RB.reduced(st, st, k, comps, dN)
            assert dev < 1e-12
            key = (st.index[(nu + up, nd + dd)], st.index[(nu, nd)])
            w = want.get(key, np.zeros_like(m))
            if env:
                w = np.conj(w)
            assert np.abs(m - w).max() < 1e-12, (side, name, site, nu, nd, dd)
            seen += 1
    assert seen > 40
