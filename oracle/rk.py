"""numpy restatement of QuTiP's explicit Runge-Kutta integrator (test oracle).

Follows qutip/solver/integrator/explicit_rk.pyx line by line in behaviour:

* set_initial_value            explicit_rk.pyx:204-230
* _estimate_first_step         explicit_rk.pyx:232-276
* integrate (step / no step)   explicit_rk.pyx:278-331
* _step_in_err                 explicit_rk.pyx:333-356
* _compute_step                explicit_rk.pyx:358-389
* _error / wrmn_error          explicit_rk.pyx:391-397, core/data/ode.pyx:27-52
* _prep_dense_out              explicit_rk.pyx:399-410
* _interpolate_step            explicit_rk.pyx:412-430
* _accumulate (skips zero dt*factor, as iadd_data does at :34-44)
                               explicit_rk.pyx:432-437
* _get_timestep                explicit_rk.pyx:439-449
* _recompute_safe_step         explicit_rk.pyx:451-467
* status codes                 explicit_rk.pxd:5-12

The derivative is any ``f(t, y) -> dy`` on complex128 numpy vectors.
"""
import numpy as np

AT_FRONT = 2
INTERPOLATED = 1
NORMAL = 0
TOO_MUCH_WORK = -1
DT_UNDERFLOW = -2
OUTSIDE_RANGE = -3
NOT_INITIATED = -4

STATUS_MESSAGES = {
    AT_FRONT: "Internal state at the desired time.",
    INTERPOLATED: ("Internal state past the desired time and "
                   "interpolation to step done."),
    NORMAL: "No work done.",
    TOO_MUCH_WORK: ("Too much work done in one call. Try to increase "
                    "the nsteps parameter or increasing the tolerance."),
    DT_UNDERFLOW: "Step size becomes too small. Try increasing tolerance.",
    OUTSIDE_RANGE: "Step outside available range.",
    NOT_INITIATED: "Not initialized.",
}


def wrmn_error(diff, state, atol, rtol):
    """Weighted RMS norm, qutip/core/data/ode.pyx:27-52."""
    diff = np.asarray(diff).ravel()
    state = np.asarray(state).ravel()
    s = np.sum((np.abs(diff) / (atol + rtol * np.abs(state))) ** 2)
    return float(np.sqrt(s / diff.size))


class ExplicitRK:
    def __init__(self, deriv, tableau, rtol=1e-6, atol=1e-8, nsteps=1000,
                 first_step=0.0, min_step=0.0, max_step=0.0, interpolate=True):
        self.deriv = deriv
        self.atol = atol
        self.rtol = rtol
        self._dt_safe = atol
        self.max_numsteps = nsteps
        self.first_step = first_step
        self.min_step = min_step or 1e-15
        self.max_step = max_step
        self.order = tableau["order"]
        self.a = np.asarray(tableau["a"], dtype=np.float64)
        self.b = np.asarray(tableau["b"], dtype=np.float64)
        self.c = np.asarray(tableau["c"], dtype=np.float64)
        self.e = tableau.get("e")
        self.bi = tableau.get("bi")
        self.rk_step = self.b.shape[0]
        self.rk_extra_step = self.c.shape[0]
        self.adaptative_step = self.e is not None
        self.interpolate = self.bi is not None and interpolate
        if self.interpolate:
            self.denseout_order = self.bi.shape[1]
        self.first_same_as_last = np.allclose(
            self.a[self.rk_step - 1, :self.rk_step], self.b, atol=1e-14)
        self._y_prev = None
        self._status = NOT_INITIATED
        # counters (not in the reference; used by the tests)
        self.n_rhs = 0
        self.n_attempts = 0

    # -- helpers -----------------------------------------------------------
    def _f(self, t, y):
        self.n_rhs += 1
        return np.asarray(self.deriv(t, y), dtype=np.complex128)

    def _accumulate(self, target, factors, dt, size):
        for i in range(size):
            f = dt * factors[i]
            if f != 0:
                target = target + f * self.k[i]
        return target

    # -- reference methods -------------------------------------------------
    def set_initial_value(self, y0, t):
        self._t = t
        self._t_prev = t
        self._t_front = t
        self._dt_int = 0.0
        self._y = np.array(y0, dtype=np.complex128)
        n_k = self.rk_extra_step if self.interpolate else self.rk_step
        self.k = [np.zeros_like(self._y) for _ in range(n_k)]
        self._y_front = self._y.copy()
        self._y_prev = self._y.copy()
        if self.first_same_as_last:
            self._k_fsal = self._f(t, self._y)
        if not self.first_step:
            self._dt_safe = self._estimate_first_step(t, self._y)
        else:
            self._dt_safe = self.first_step
        self._status = NORMAL

    def _estimate_first_step(self, t, y0):
        if not self.adaptative_step:
            return 0.0
        norm = float(np.linalg.norm(y0))
        tol = self.atol + norm * self.rtol
        if self.first_same_as_last:
            self.k[0] = self._k_fsal.copy()
        else:
            self.k[0] = self._f(t, y0)
        if norm <= self.atol:
            norm = 1
        tmp_norm = float(np.linalg.norm(self.k[0]))
        factorial = 1.0
        for i in range(1, self.order + 1):
            factorial *= i
        if tmp_norm >= self.atol * 1e-6:
            dt1 = (tol * factorial * norm ** self.order) ** (1 / (self.order + 1)) / tmp_norm
        else:
            dt1 = (tol * factorial) ** (1 / (self.order + 1)) * norm * 0.5
        t1 = t + dt1 / 100
        y_temp = y0 + (dt1 / 100) * self.k[0]
        self.k[1] = self._f(t1, y_temp)
        tmp_norm = float(np.linalg.norm(self.k[1]))
        if tmp_norm >= self.atol * 1e-6:
            dt2 = (tol * factorial * norm ** self.order) ** (1 / (self.order + 1)) / tmp_norm
        else:
            dt2 = dt1
        dt = min(dt1, dt2)
        if self.max_step:
            dt = min(self.max_step, dt)
        if self.min_step:
            dt = max(self.min_step, dt)
        return dt

    def integrate(self, t, step=False):
        nsteps_left = self.max_numsteps
        if self._y_prev is None:
            self._status = NOT_INITIATED
            return
        if t == self._t:
            return
        if t < self._t_prev:
            self._status = OUTSIDE_RANGE
            return
        if self.interpolate and t < self._t_front:
            if self._status != INTERPOLATED:
                self._prep_dense_out()
                self._status = INTERPOLATED
            self._y = self._interpolate_step(t)
            self._t = t
            return
        self._status = NORMAL
        if step and self._t < self._t_front and t > self._t_front:
            t = self._t_front
        while self._t_front < t and self._status >= 0:
            self._y_prev = self._y_front.copy()
            self._t_prev = self._t_front
            nsteps_left -= self._step_in_err(t, nsteps_left)
            if step:
                break
        if self._status < 0:
            return
        if self._t_front > t:
            self._prep_dense_out()
            self._status = INTERPOLATED
            self._t = t
            self._y = self._interpolate_step(t)
        else:
            self._status = AT_FRONT
            self._t = self._t_front
            self._y = self._y_front.copy()

    def _step_in_err(self, t, max_step):
        error = 1.0
        nsteps = 0
        while error >= 1:
            dt = self._get_timestep(t)
            error = self._compute_step(dt)
            self._dt_int = dt
            self._recompute_safe_step(error, dt)
            if dt == self.min_step and error > 1:
                self._status = DT_UNDERFLOW
                break
            nsteps += 1
            if nsteps > max_step:
                self._status = TOO_MUCH_WORK
                break
        if self.first_same_as_last:
            self._k_fsal = self.k[self.rk_step - 1].copy()
        return nsteps

    def _compute_step(self, dt):
        self.n_attempts += 1
        if self.first_same_as_last:
            self.k[0] = self._k_fsal.copy()
        else:
            self.k[0] = self._f(self._t_prev, self._y_prev)
        for i in range(1, self.rk_step):
            y_temp = self._accumulate(self._y_prev.copy(), self.a[i, :], dt, i)
            self.k[i] = self._f(self._t_prev + self.c[i] * dt, y_temp)
        self._y_front = self._accumulate(self._y_prev.copy(), self.b, dt, self.rk_step)
        self._t_front = self._t_prev + dt
        return self._error(self._y_front, dt)

    def _error(self, y_new, dt):
        if not self.adaptative_step:
            return 0.0
        err_vec = self._accumulate(np.zeros_like(y_new), self.e, dt, self.rk_step)
        return wrmn_error(err_vec, y_new, self.atol, self.rtol)

    def _prep_dense_out(self):
        dt = self._dt_int
        for i in range(self.rk_step, self.rk_extra_step):
            y_temp = self._accumulate(self._y_prev.copy(), self.a[i, :], dt, i)
            self.k[i] = self._f(self._t_prev + self.c[i] * dt, y_temp)

    def interp_factors(self, t):
        """b_factor of _interpolate_step (explicit_rk.pyx:419-423)."""
        tau = (t - self._t_prev) / self._dt_int
        bf = np.zeros(self.rk_extra_step)
        for i in range(self.rk_extra_step):
            v = 0.0
            for j in range(self.denseout_order - 1, -1, -1):
                v += self.bi[i, j]
                v *= tau
            bf[i] = v
        return bf

    def _interpolate_step(self, t):
        bf = self.interp_factors(t)
        return self._accumulate(self._y_prev.copy(), bf, self._dt_int, self.rk_extra_step)

    def _get_timestep(self, t):
        dt_needed = t - self._t_prev
        if not self.adaptative_step:
            return dt_needed
        if self.interpolate:
            return self._dt_safe
        if dt_needed <= self._dt_safe:
            return dt_needed
        return dt_needed / (int(dt_needed / self._dt_safe) + 1)

    def _recompute_safe_step(self, err, dt):
        if not self.adaptative_step:
            return
        if err == 0:
            factor = 10.0
        else:
            factor = 0.9 * err ** (-1 / (self.order + 1))
            factor = min(10.0, factor)
            factor = max(0.2, factor)
        self._dt_safe = dt * factor
        if self.max_step:
            self._dt_safe = min(self.max_step, self._dt_safe)
        if self.min_step:
            self._dt_safe = max(self.min_step, self._dt_safe)

    # -- accessors ---------------------------------------------------------
    def successful(self):
        return self._status >= 0

    @property
    def status(self):
        return self._status

    @property
    def y(self):
        return self._y

    @property
    def t(self):
        return self._t

    @property
    def t_prev(self):
        return self._t_prev

    @property
    def t_front(self):
        return self._t_front


def solve(deriv, y0, tlist, tableau, **opts):
    """Run the integrator over ``tlist`` like Integrator.run + Solver.run
    (qutip/solver/integrator/integrator.py:197-212, solver_base.py:159-226):
    returns the list of states at every time of ``tlist``."""
    rk = ExplicitRK(deriv, tableau, **opts)
    rk.set_initial_value(y0, tlist[0])
    out = [np.array(y0, dtype=np.complex128)]
    for t in tlist[1:]:
        rk.integrate(t, step=False)
        if not rk.successful():
            raise RuntimeError(STATUS_MESSAGES[rk.status])
        out.append(rk.y.copy())
    return out, rk
