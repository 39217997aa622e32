"""The Integrator protocol of QuTiP 5 (qutip/solver/integrator/integrator.py:
23-240), restated so a solver can drive ``hip_rk`` without qutip installed.

Same class attributes (``integrator_options``, ``rhs_format``, ``name``,
``method``), same methods (``set_state``, ``integrate``, ``mcstep``,
``get_state``, ``run``, ``reset``, ``options``) and the same exception type
name, ``IntegratorException``.  When qutip is importable,
qutip_amd.qutip_plugin subclasses qutip's own Integrator instead.
"""
import numpy as np

__all__ = ["Integrator", "IntegratorException"]


class IntegratorException(Exception):
    """Error from the ODE solver being unable to integrate with the given
    parameters (integrator.py:11-20)."""


class Integrator:
    integrator_options = {}
    _options = None
    rhs_format = "callable"
    name = None
    method = ""

    def __init__(self, derivative, options):
        if not callable(derivative):
            raise TypeError("The derivative must be a callable.")
        self.derivative = derivative
        self._is_set = False
        self._back = (np.inf, None)
        self._options = self.integrator_options.copy()
        self.options = options
        self._prepare()

    def _prepare(self):
        raise NotImplementedError

    def set_state(self, t, state):
        raise NotImplementedError

    def integrate(self, t, copy=True):
        raise NotImplementedError

    def mcstep(self, t, copy=True):
        # default implementation of integrator.py:132-157
        t_last, state = self.get_state()
        if t > t_last:
            self._back = t_last, state
        elif t > self._back[0]:
            self.set_state(*self._back)
        else:
            raise IntegratorException(
                "`t` is outside the integration range: "
                f"{self._back[0]}..{t_last}.")
        return self.integrate(t, copy)

    def get_state(self, copy=True):
        raise NotImplementedError

    def run(self, tlist):
        for t in tlist[1:]:
            yield self.integrate(t, False)

    def reset(self, hard=False):
        if self._is_set:
            state = self.get_state()
        if hard:
            self._prepare()
        if self._is_set:
            self.set_state(*state)

    @property
    def options(self):
        return self._options

    @options.setter
    def options(self, new_options):
        self._options = {
            **self._options,
            **{key: new_options[key] for key in self.integrator_options.keys()
               if key in new_options and new_options[key] is not None},
        }
