"""Result of a solver run (qutip/solver/result.py:94-340, the parts the
solvers here use): ``times``, ``expect`` (one array per e_op), ``e_data``
(dict keyed like e_ops), ``states``, ``final_state``, ``stats``."""
import numpy as np

__all__ = ["Result"]


class Result:
    def __init__(self, e_ops, options, solver="", stats=None):
        self.solver = solver
        self.stats = stats if stats is not None else {}
        self.options = dict(options or {})
        if e_ops is None:
            e_ops = []
        if isinstance(e_ops, dict):
            self._keys = list(e_ops.keys())
            self._ops = list(e_ops.values())
        else:
            if not isinstance(e_ops, (list, tuple)):
                e_ops = [e_ops]
            self._keys = list(range(len(e_ops)))
            self._ops = list(e_ops)
        self._dict_input = isinstance(e_ops, dict)
        self.times = []
        self.states = []
        self.final_state = None
        self._raw = [[] for _ in self._ops]

    @property
    def e_ops(self):
        return dict(zip(self._keys, self._ops))

    @property
    def e_data(self):
        return {k: self._cast(v) for k, v in zip(self._keys, self._raw)}

    @property
    def expect(self):
        return [self._cast(v) for v in self._raw]

    @staticmethod
    def _cast(vals):
        return np.asarray(vals)

    def add_expect_values(self, t, values):
        self.times.append(t)
        for k, v in enumerate(values):
            self._raw[k].append(v)

    def add_expect_block(self, times, per_op):
        """add_expect_values for many times at once: per_op[k] holds the
        values of e_op k at every time of ``times``."""
        self.times.extend(times)
        for k, v in enumerate(per_op):
            self._raw[k].extend(v)

    def add_state(self, state, store_states, store_final):
        if store_states:
            self.states.append(state)
        if store_final:
            self.final_state = state

    def __repr__(self):
        return (f"<Result solver={self.solver!r} ntimes={len(self.times)} "
                f"n_e_ops={len(self._ops)}>")
