"""Solver registries stand-in: per-class integrator dicts filled by
add_integrator (solver_base.py:462-492, sesolve.py:144, mesolve.py:205,
multitraj.py:65), and MCSolver's constructor signature (mcsolve.py:469)."""
from .integrator import Integrator


class Solver:
    _avail_integrators = {}

    @classmethod
    def add_integrator(cls, integrator, key):
        if not issubclass(integrator, Integrator):
            raise TypeError("the integrator must be a subclass of Integrator")
        cls._avail_integrators = {**cls.__dict__.get("_avail_integrators", {}), key: integrator}


class SESolver(Solver):
    _avail_integrators = {}


class MESolver(Solver):
    _avail_integrators = {}


class MCSolver(Solver):
    _avail_integrators = {}
    solver_options = {"method": "vern7", "keep_runs_results": False, "norm_steps": 25,
                      "norm_t_tol": 1e-6, "norm_tol": 1e-4, "norm_min_step": 0.1,
                      "mc_corr_eps": 1e-10, "improved_sampling": False, "atol": 1e-8,
                      "rtol": 1e-6, "nsteps": 1000}

    def __init__(self, H, c_ops, *, options=None):
        extra = set(options or {}) - set(self.solver_options)
        if extra:
            raise KeyError(f"Options {sorted(extra)} are not supported.")
        self.options = {**self.solver_options, **(options or {})}
