"""Wave-function guess between sweep steps (wave_trans_LR/RL,
DMRG_code/dmrg_sub.f90:1471-1789) on the fake device (numpy plans).

Without truncation the guess is the previous ground state written in the new
superblock's basis, so it must be the new superblock's ground state: the
overlap with the state the Lanczos then converges to is 1 (this pins every
recoupling factor and fermion sign of the transform; a wrong sign in any
channel lowers it: the reference's factors alone, without the fermion sign
of the site moving past the other block in this build's basis order, gave
guesses with energies near -3.8 on 3x3 where the ground energy is -6.40).
With truncation it stays close to 1."""
import numpy as np
import pytest

import fake_device
from oracle import tj_ed, tj_model
from tsc_tri_t1t2j1j2_amd.dmrg import DMRG


def _run(monkeypatch, Lx, Ny, Ne, kmax, sweeps):
    ctx = fake_device.install(monkeypatch)
    starts = []
    orig_initial = DMRG._initial

    def initial(self, sb, guess):
        v = orig_initial(self, sb, guess)
        starts.append((guess is not None, v.x.copy()))
        return v

    monkeypatch.setattr(DMRG, "_initial", initial)
    finals = []
    orig_step = DMRG.step

    def step(self, *a, **k):
        r = orig_step(self, *a, **k)
        finals.append(r["psi"].x.copy())
        return r

    monkeypatch.setattr(DMRG, "step", step)
    d = DMRG(ctx, Lx, Ny, Ne=Ne, kept_min=20, kept_max=kmax, sweeps=sweeps, lanzero=1e-12, trun_err=1e-14,
             native_planner=False)
    E = d.run()
    assert len(starts) == len(finals)
    ovs = [abs(np.vdot(f, v)) / (np.linalg.norm(f) * np.linalg.norm(v))
           for (g, v), f in zip(starts, finals) if g]
    return d, E, ovs


@pytest.mark.parametrize("Lx,Ny,Ne", [(3, 3, 8), (2, 4, 6)])
def test_guess_is_the_next_ground_state_without_truncation(monkeypatch, Lx, Ny, Ne):
    d, E, ovs = _run(monkeypatch, Lx, Ny, Ne, 4000, 1)
    assert d.guessed == len(ovs) > 8
    assert min(ovs) > 1 - 1e-9, ovs
    e_ed = tj_ed.ground_energy(Lx * Ny, Ne, tj_model.bonds(Lx, Ny, 1.0, 0.2, 0.4, 0.016))
    assert abs(E - e_ed) < 1e-9


def test_guess_with_truncation_stays_close(monkeypatch):
    d, E, ovs = _run(monkeypatch, 3, 4, 10, 30, 1)
    assert d.guessed == len(ovs) > 8
    # the first half-sweep starts from warm-up blocks made for other targets
    assert min(ovs) > 0.9 and np.median(ovs) > 0.999, ovs


def test_guess_cuts_lanczos_work(monkeypatch):
    ctx = fake_device.install(monkeypatch)
    counts = {}
    for wg in (False, True):
        d = DMRG(ctx, 3, 3, Ne=8, kept_min=20, kept_max=4000, sweeps=1, lanzero=1e-11, trun_err=1e-12,
                 native_planner=False, wave_guess=wg)
        log = []
        orig = d.step

        def step(*a, **k):
            r = orig(*a, **k)
            log.append(r["matvecs"])
            return r

        d.step = step
        d.run()
        counts[wg] = sum(log)
    assert counts[True] < 0.6 * counts[False], counts
