"""The DMRG driver's step logic on the CPU (device objects replaced by the
oracle through tests/fake_device.py), against exact diagonalisation."""
import numpy as np
import pytest

from oracle import tj_ed, tj_model
from tests import fake_device
from tsc_tri_t1t2j1j2_amd.dmrg import DMRG


@pytest.mark.parametrize("kmax", [4000, 40])
def test_driver_on_fake_device(monkeypatch, kmax):
    ctx = fake_device.install(monkeypatch)
    Lx, Ny, Ne = 3, 3, 8
    b = tj_model.bonds(Lx, Ny, 1.0, 0.2, 0.4, 0.016)
    e_ed = tj_ed.ground_energy(Lx * Ny, Ne, b)
    log = []
    d = DMRG(ctx, Lx, Ny, Ne=Ne, kept_min=min(20, kmax), kept_max=kmax, sweeps=2, lanzero=1e-10, trun_err=1e-12,
             log=log.append, native_planner=False)
    e = d.run()
    print("\n".join(log))
    assert e >= e_ed - 1e-9
    assert (e - e_ed) / (Lx * Ny) < (1e-9 if kmax > 1000 else 1e-4), (e, e_ed)
