"""Pins of the Monte-Carlo oracle (CPU only): the jump time of a decaying
qubit in closed form, the reference draw order, and the ensemble mean
against the master equation within 3 sigma."""
import numpy as np
import pytest
import scipy.sparse as sp

from oracle import exact, mc, models, tableaus


def test_qubit_single_jump_closed_form():
    """H = 0, c = sqrt(g) sigma^-, psi0 = up: ||psi(t)||^2 = exp(-g t), so the
    jump happens at t* = -ln(r0) / g with r0 the first draw."""
    g = 0.7
    H = sp.csr_matrix((2, 2), dtype=complex)
    c = np.sqrt(g) * models.sigmam()
    tlist = np.linspace(0, 20, 41)
    for seed in range(5):
        r0 = np.random.default_rng(seed).random()
        tr = mc.Trajectory(H, [c], tableaus.vern7, rtol=1e-10, atol=1e-12)
        tr.run(models.basis(2, 0), tlist, np.random.default_rng(seed))
        t_star = -np.log(r0) / g
        if t_star < tlist[-1]:
            assert len(tr.collapses) == 1
            # secant stops at |norm^2 - r0| < 1e-4 r0  ->  |t - t*| < ~1e-4 / g
            assert abs(tr.collapses[0][0] - t_star) < 2e-4 / g
            assert tr.collapses[0][1] == 0
            assert tr.draws_used == 2  # initial target + next target (one channel)
        else:
            assert len(tr.collapses) == 0


def test_channel_draw_order_two_channels():
    """With two channels a jump consumes a channel draw and a target draw."""
    g = 0.5
    H = sp.csr_matrix((4, 4), dtype=complex)
    c1 = np.sqrt(g) * sp.kron(models.sigmam(), models.qeye(2), format="csr")
    c2 = np.sqrt(g) * sp.kron(models.qeye(2), models.sigmam(), format="csr")
    tr = mc.Trajectory(H, [c1, c2], tableaus.vern7, rtol=1e-10, atol=1e-12)
    tr.run(np.kron(models.basis(2, 0), models.basis(2, 0)), np.linspace(0, 40, 21),
           np.random.default_rng(3))
    assert tr.draws_used == 1 + 2 * len(tr.collapses)
    assert sorted(w for _, w in tr.collapses) == [0, 1]  # each qubit decays once


@pytest.mark.slow
def test_ensemble_mean_within_3_sigma_of_master_equation():
    m = models.ising(3, 0.2, tf=4.0, npts=9)
    tlist = m["tlist"]
    ntraj = 150
    runs = []
    for seed in range(ntraj):
        tr = mc.Trajectory(m["H"], m["c_ops"], tableaus.vern7)
        runs.append(tr.run(m["psi0"], tlist, np.random.default_rng(seed), m["e_ops"]).real)
    runs = np.array(runs)
    mean, std = runs.mean(0), runs.std(0)
    ex = exact.mesolve_expect(m["H"], m["c_ops"], m["psi0"], tlist, m["e_ops"])
    sig = std / np.sqrt(ntraj)
    assert np.all(np.abs(mean - ex.T) <= 3 * sig + 1e-9)
