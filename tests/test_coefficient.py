"""Coefficients of the list form on the host (no GPU): the compiled stack
programs evaluated instruction by instruction as the device runs them agree
with the string's own evaluation in the reference's environment
(StrFunctionCoefficient, qutip/core/cy/coefficient.pyx:292-410); the
interpolated arrays agree with scipy's spline they are built from
(InterCoefficient, coefficient.pyx:412-560)."""
import numpy as np
import pytest
from scipy.interpolate import make_interp_spline

from qutip_amd import coefficient as qc

EXPRESSIONS = [
    ("cos(w * t)", {"w": 1.3}),
    ("exp(-1j * w * t)", {"w": 0.7}),
    ("A * exp(-(t - t0) ** 2 / (2 * s ** 2))", {"A": 2.0, "t0": 1.5, "s": 0.4}),
    ("sin(t) ** 2 + cos(t) ** 3 - tan(0.1 * t)", {}),
    ("sqrt(1 + t) * log(2 + t) / log10(3 + t)", {}),
    ("abs(sin(t)) + norm(exp(1j * t) + 0.5) - real(exp(1j * t)) + imag(conj(exp(1j * t)))", {}),
    ("arg(-1 + 1j * t) + sinh(0.2 * t) * cosh(0.1 * t) - tanh(t)", {}),
    ("asin(0.3 * sin(t)) + acos(0.2) + atan(t) + asinh(t) + acosh(2 + t) + atanh(0.5)", {}),
    ("erf(t - 1)", {}),
    ("np.cos(np.pi * t) + np.exp(-t)", {}),
    ("(t < 2) * 1.0 + (t >= 3) * 0.5", {}),
    ("1.0 if t < 1.5 else -2.0 * t", {}),
    ("(1 + 1j * t) ** 3 - (2 + t) ** -2 + t ** 0.5", {}),
    ("-t / (1 + 1j) + pi", {}),
]


@pytest.mark.parametrize("expr,args", EXPRESSIONS)
def test_string_program_matches_eval(expr, args):
    c = qc.coefficient(expr, args=args)
    prog = c.device_program()
    assert prog is not None, expr
    for t in np.linspace(0.0, 4.0, 41):
        want = c(t)
        got = prog.evaluate(t)
        assert abs(got - want) <= 1e-13 * max(1.0, abs(want)), (expr, t, got, want)


def test_replace_arguments_recompiles():
    c = qc.coefficient("cos(w * t)", args={"w": 1.0})
    c2 = c.replace_arguments(w=2.0)
    assert c2(0.5) == pytest.approx(np.cos(1.0))
    assert c2.device_program().evaluate(0.5) == pytest.approx(np.cos(1.0), abs=1e-15)
    assert c(0.5) == pytest.approx(np.cos(0.5))


def test_not_compilable_falls_back_to_host():
    c = qc.coefficient("spe.jv(0, t)")  # a Bessel function: host only
    assert c.device_program() is None
    f = qc.coefficient(lambda t, w: np.cos(w * t), args={"w": 2.0})
    assert f.device_program() is None and f(1.0) == pytest.approx(np.cos(2.0))
    g = qc.coefficient(lambda t, args: args["a"] * t, args={"a": 3.0})
    assert g(2.0) == pytest.approx(6.0)


def test_invalid_string_raises():
    with pytest.raises(Exception):
        qc.coefficient("cos(w * t)")  # w undefined


@pytest.mark.parametrize("order", [0, 1, 2, 3])
@pytest.mark.parametrize("uniform", [True, False])
def test_interpolation_matches_scipy(order, uniform):
    rng = np.random.default_rng(order)
    tl = np.linspace(0, 3, 13) if uniform else np.sort(rng.uniform(0, 3, 13))
    y = np.sin(2 * tl) + 1j * np.cos(tl)
    c = qc.coefficient(y, tlist=tl, order=order)
    prog = c.device_program()
    ts = np.linspace(tl[0] - 0.2, tl[-1] + 0.2, 57)
    for t in ts:
        got = prog.evaluate(t)
        assert got == c(t)
        if t <= tl[0]:
            want = y[0]
        elif t >= tl[-1]:
            want = y[-1]
        elif order == 0:
            want = y[np.searchsorted(tl, t, side="right") - 1]
        elif order == 1:
            want = np.interp(t, tl, y.real) + 1j * np.interp(t, tl, y.imag)
        else:
            want = make_interp_spline(tl, y, k=order)(t)
        assert abs(got - want) < 1e-12, (order, t, got, want)


def test_sum_product_conj_programs():
    a = qc.coefficient("cos(t)")
    b = qc.coefficient(np.linspace(0, 1, 5) ** 2, tlist=np.linspace(0, 1, 5), order=1)
    for c in (a + b, a * b, (a * 2.0).conj(), a + 1j):
        p = c.device_program()
        for t in (0.0, 0.3, 0.77, 1.2):
            assert abs(p.evaluate(t) - c(t)) < 1e-14


def test_program_set_layout():
    a = qc.coefficient("cos(w * t)", args={"w": 2.0}).device_program()
    b = qc.coefficient(np.arange(4.0), tlist=np.arange(4.0), order=0).device_program()
    s = qc.DeviceProgram.concat([a, b])
    start, code, consts = s.arrays()
    assert list(start) == [0, len(a.code), len(a.code) + len(b.code)]
    assert s.evaluate(0.4, 0) == pytest.approx(np.cos(0.8))
    assert s.evaluate(2.5, 1) == pytest.approx(2.0)
