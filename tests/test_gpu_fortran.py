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
    out = subprocess.run([EXE2, str(inp)], capture_output=True, text=True, timeout=120)
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
