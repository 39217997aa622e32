"""numpy restatement of QuTiP's Monte-Carlo trajectory loop (test oracle).

Follows qutip/solver/mcsolve.py:
* MCIntegrator.set_state            :245-281 (target = first draw)
* MCIntegrator.integrate            :286-302
* _prob_func / _norm_func           :311-319 (||psi||^2, ||psi||)
* _find_collapse_time               :321-369 (log-ratio secant, norm_steps,
                                              norm_tol, norm_t_tol, norm_min_step)
* _do_collapse                      :371-406 (channel by <c^dag c>, mc_corr_eps)
* H_eff = -iH - 1/2 sum c^dag c     :491-496
with the RK integrator of oracle/rk.py used through integrate(t, step=True)
(qutip_integrator.py:84-87).  Random numbers: ``numpy.random.default_rng(seed)``
per trajectory (multitraj.py:382-393), consumed in the reference order.

Pinned by: the closed-form decay of a single qubit (exactly one jump, at
the time the norm reaches the drawn target; qutip/tests/solver/
test_qubit_evolution.py:80-101 for the ensemble) and by the ensemble mean
against the master-equation solution within 3 sigma.
"""
import numpy as np
import scipy.sparse as sp

from . import rk as _rk

DEFAULTS = dict(norm_steps=25, norm_t_tol=1e-6, norm_tol=1e-4, norm_min_step=0.1,
                mc_corr_eps=1e-10)


class Trajectory:
    def __init__(self, H, c_ops, tableau, rtol=1e-6, atol=1e-8, nsteps=1000, **mc):
        self.opts = {**DEFAULTS, **mc}
        H = sp.csr_matrix(H)
        self.c_ops = [sp.csr_matrix(c) for c in c_ops]
        self.n_ops = [(c.conj().T @ c).tocsr() for c in self.c_ops]
        Heff = -1j * H
        for n in self.n_ops:
            Heff = Heff - 0.5 * n
        self.Heff = sp.csr_matrix(Heff)
        self.tableau = tableau
        self.rk_opts = dict(rtol=rtol, atol=atol, nsteps=nsteps)

    def _new_rk(self, t, psi):
        rk = _rk.ExplicitRK(lambda t, y: self.Heff @ y, self.tableau, **self.rk_opts)
        rk.set_initial_value(psi, t)
        self.rk = rk

    def _mcstep(self, t):
        self.rk.integrate(t, step=True)
        if not self.rk.successful():
            raise RuntimeError(_rk.STATUS_MESSAGES[self.rk.status])
        return self.rk.t, self.rk.y

    @staticmethod
    def _prob(psi):
        return float(np.vdot(psi, psi).real)

    def run(self, psi0, tlist, rng, e_ops=()):
        self.collapses = []
        self.draws_used = 0

        def draw():
            self.draws_used += 1
            return rng.random()

        self.target = draw()
        self._new_rk(tlist[0], np.asarray(psi0, dtype=np.complex128))
        out = [[self._expect(O, psi0) for O in e_ops]]
        for t in tlist[1:]:
            t_old, y_old = self.rk.t, self.rk.y
            norm_old = self._prob(y_old)
            while t_old < t:
                t_step, state = self._mcstep(t)
                nrm = self._prob(state)
                if nrm <= self.target:
                    t_col, state = self._find_collapse_time(norm_old, nrm, t_old, t_step)
                    self._do_collapse(t_col, state, draw)
                    t_old, y_old = self.rk.t, self.rk.y
                    norm_old = 1.0
                else:
                    t_old, y_old = t_step, state
                    norm_old = nrm
            psi = y_old / np.sqrt(self._prob(y_old))
            out.append([self._expect(O, psi) for O in e_ops])
        return np.array(out)

    @staticmethod
    def _expect(O, psi):
        return complex(np.vdot(psi, O @ psi) / np.vdot(psi, psi))

    def _find_collapse_time(self, norm_old, norm, t_prev, t_final):
        o = self.opts
        tries = 0
        ratio_cutoff = o["norm_min_step"]
        while tries < o["norm_steps"]:
            tries += 1
            if (t_final - t_prev) < o["norm_t_tol"]:
                t_guess = t_final
                state = self.rk.y
                break
            dt = t_final - t_prev
            ratio = np.log(norm_old / self.target) / np.log(norm_old / norm)
            if ratio < ratio_cutoff:
                ratio = ratio_cutoff
            if ratio > (1 - ratio_cutoff):
                ratio = 1 - ratio_cutoff
            t_guess = t_prev + dt * ratio
            if (t_guess - t_prev) < o["norm_t_tol"]:
                t_guess = t_prev + o["norm_t_tol"]
            _, state = self._mcstep(t_guess)
            norm2_guess = self._prob(state)
            if abs(self.target - norm2_guess) < o["norm_tol"] * self.target:
                break
            elif norm2_guess < self.target:
                t_final = t_guess
                norm = norm2_guess
            else:
                t_prev = t_guess
                norm_old = norm2_guess
        if tries >= o["norm_steps"]:
            raise RuntimeError("Could not find the collapse time within desired tolerance.")
        return t_guess, state

    def _do_collapse(self, t, state, draw):
        num_ops = len(self.n_ops)
        if num_ops == 1:
            which = 0
        else:
            probs = [float(np.vdot(state, n @ state).real) for n in self.n_ops]
            target = sum(probs) * draw() - probs[0]
            which = 0
            while target > 0 and which <= num_ops:
                which += 1
                target -= probs[which]
        new = self.c_ops[which] @ state
        nn = np.sqrt(self._prob(new))
        if nn < self.opts["mc_corr_eps"]:
            new = state / np.sqrt(self._prob(state))
        else:
            new = new / nn
            self.collapses.append((t, which))
            self.target = draw()
        self._new_rk(t, new)
