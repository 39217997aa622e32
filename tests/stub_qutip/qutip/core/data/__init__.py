"""Data-layer registry stand-in: Data base class (base.pyx:23-93), the
conversion registry ``to`` (convert.pyx:208-404) and Dispatchers with the
reference signatures (dispatch.pyx:267-330)."""
import inspect as _inspect

P = _inspect.Parameter
POS, KW = P.POSITIONAL_ONLY, P.POSITIONAL_OR_KEYWORD


class Data:
    def __init__(self, shape):
        self.shape = tuple(shape)


class Dense(Data):
    def __init__(self, arr):
        import numpy as np
        self._arr = np.asarray(arr, dtype=complex)
        super().__init__(self._arr.shape)

    def to_array(self):
        return self._arr


class CSR(Data):
    def __init__(self, m):
        import scipy.sparse as sp
        self._m = sp.csr_matrix(m, dtype=complex)
        super().__init__(self._m.shape)

    def as_scipy(self):
        return self._m


class Dia(Data):
    def __init__(self, m):
        import scipy.sparse as sp
        self._m = sp.dia_matrix(m, dtype=complex)
        super().__init__(self._m.shape)

    def as_scipy(self):
        return self._m


class _To:
    """convert.pyx:208-404: add_conversions takes (to_type, from_type,
    converter[, weight]) with weight > 0; aliases and groups by name."""

    def __init__(self):
        self.dtypes = {Data, Dense, CSR, Dia}
        self.conversions = {}
        self.aliases = {}
        self.groups = {}

    def add_conversions(self, conversions):
        for conv in conversions:
            to_t, from_t, fn = conv[:3]
            w = conv[3] if len(conv) > 3 else 1
            if not w > 0:
                raise ValueError("conversion weights must be positive")
            if not (isinstance(to_t, type) and issubclass(to_t, Data)):
                raise TypeError("conversion target must be a Data type")
            self.dtypes.update((to_t, from_t))
            self.conversions[(to_t, from_t)] = (fn, w)

    def register_aliases(self, aliases, layer_type):
        if layer_type not in self.dtypes:
            raise ValueError("alias for an unknown type")
        for a in aliases:
            self.aliases[a] = layer_type

    def register_group(self, aliases, dense, sparse):
        for a in aliases:
            self.groups[a] = (dense, sparse)

    def __call__(self, to_type, data):
        if isinstance(data, to_type):
            return data
        fn, _ = self.conversions[(to_type, type(data))]
        return fn(data)


to = _To()


class Dispatcher:
    """The reference Dispatcher's contract for specialisations: one type per
    input (plus the output type when ``out``) and a callable with exactly the
    dispatcher's signature (dispatch.pyx:287-330)."""

    def __init__(self, sig, name, inputs, out=False):
        self.__signature__ = sig
        self.name = name
        self.inputs = inputs
        self.output = out
        self.specialisations = {}

    def add_specialisations(self, specs):
        n = len(self.inputs) + int(self.output)
        for spec in specs:
            spec = tuple(spec)
            if len(spec) != n + 1:
                raise ValueError(f"{self.name}: specialisation {spec} has the wrong length")
            for t in spec[:n]:
                if t not in to.dtypes:
                    raise ValueError(f"{self.name}: unknown data type {t}")
            got = _inspect.signature(spec[-1]).parameters
            want = self.__signature__.parameters
            if [(p.name, p.default) for p in got.values()] != \
                    [(p.name, p.default) for p in want.values()]:
                raise TypeError(f"{self.name}: specialisation signature {list(got)} "
                                f"differs from {list(want)}")
            self.specialisations[spec[:n]] = spec[-1]


def _sig(*params):
    return _inspect.Signature([P(name, kind, **({"default": d} if d is not P.empty else {}))
                               for name, kind, d in params])


E = P.empty
matmul = Dispatcher(_sig(("left", POS, E), ("right", POS, E), ("scale", KW, 1)),
                    "matmul", ("left", "right"), True)            # matmul.pyx:1018-1029
expect = Dispatcher(_sig(("op", POS, E), ("state", POS, E)),
                    "expect", ("op", "state"))                      # expect.pyx:391-400
inner = Dispatcher(_sig(("left", POS, E), ("right", POS, E), ("scalar_is_ket", KW, False)),
                   "inner", ("left", "right"))                      # inner.pyx:323-334
add = Dispatcher(_sig(("left", POS, E), ("right", POS, E), ("scale", KW, 1)),
                 "add", ("left", "right"), True)                    # add.pyx:324-335
iadd = Dispatcher(_sig(("left", POS, E), ("right", POS, E), ("scale", KW, 1)),
                  "iadd", ("left", "right"), True)                  # add.pyx:349-360
mul = Dispatcher(_sig(("matrix", POS, E), ("value", KW, E)), "mul", ("matrix",), True)
imul = Dispatcher(_sig(("matrix", POS, E), ("value", KW, E)), "imul", ("matrix",), True)
zeros = Dispatcher(_sig(("rows", KW, E), ("cols", KW, E)), "zeros", (), True)  # constant.py:18
trace = Dispatcher(_sig(("matrix", POS, E),), "trace", ("matrix",))           # trace.pyx:110
column_stack = Dispatcher(_sig(("matrix", POS, E),), "column_stack", ("matrix",), True)
column_unstack = Dispatcher(_sig(("matrix", POS, E), ("rows", KW, E)), "column_unstack",
                            ("matrix",), True)                      # reshape.pyx:208-259


class norm:  # norm.pyx:180-205
    l2 = Dispatcher(_sig(("vector", POS, E),), "l2", ("vector",))
    frobenius = Dispatcher(_sig(("matrix", POS, E),), "frobenius", ("matrix",))


class ode:  # ode.pyx:167-178
    wrmn_error = Dispatcher(_sig(("diff", POS, E), ("state", POS, E), ("atol", KW, E),
                                 ("rtol", KW, E)), "wrmn_error", ("diff", "state"))
