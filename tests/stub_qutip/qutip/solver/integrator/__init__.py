"""Integrator base stand-in (qutip/solver/integrator/integrator.py:23-240)."""


class IntegratorException(Exception):
    pass


class Integrator:
    integrator_options = {}
    rhs_format = "callable"
