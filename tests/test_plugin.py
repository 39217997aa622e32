"""The drop-in boundary on the CPU: qutip_plugin.register() against a
stand-in of QuTiP 5's registration API (tests/stub_qutip: dispatcher
signatures, the ``to`` registry and add_integrator restated from the
reference files they cite), and the integrator's handling of QuTiP's
LindbladMatrixForm, whose call raises (lindblad_matrix_form.pyx:217-222)
while ``H_nh`` / ``c_ops`` are public (lindblad_matrix_form.pxd:7-8)."""
import json
import pathlib
import subprocess
import sys
import types

import numpy as np
import pytest

from qutip_amd.solver import hip_rk
from qutip_amd.solver.solver_base import Solver
import qutip_amd as qa

ROOT = pathlib.Path(__file__).resolve().parents[1]

_REGISTER = r"""
import json, sys
sys.path.insert(0, sys.argv[1]); sys.path.insert(0, sys.argv[2])
import numpy as np, scipy.sparse as sp
import qutip, qutip.core.data as qd
from qutip.solver import SESolver, MESolver, MCSolver
from qutip_amd import qutip_plugin
qutip_plugin.register()
qutip_plugin.register()  # idempotent
out = {}
out["aliases"] = sorted(qd.to.aliases)
out["groups"] = sorted(qd.to.groups)
out["conversions"] = len(qd.to.conversions)
out["integrators"] = {c.__name__: sorted(c._avail_integrators) for c in (SESolver, MESolver, MCSolver)}
spec = {}
for name in ("matmul", "expect", "inner", "add", "iadd", "mul", "imul", "zeros", "trace",
             "column_stack", "column_unstack"):
    spec[name] = len(getattr(qd, name).specialisations)
spec["l2"] = len(qd.norm.l2.specialisations)
spec["frobenius"] = len(qd.norm.frobenius.specialisations)
spec["wrmn_error"] = len(qd.ode.wrmn_error.specialisations)
out["spec"] = spec
out["host_matmul"] = sorted(t[0].__name__ for t in qd.matmul.specialisations)
integ = SESolver._avail_integrators["hip_rk"]
out["run_overridden"] = "run" in {k for c in integ.__mro__[:2] for k in vars(c)} or \
    integ.run is not __import__("qutip.solver.integrator", fromlist=["Integrator"]).Integrator.__dict__.get("run")
HMC = qutip_plugin.HipMCSolver
out["mc_subclass"] = issubclass(HMC, MCSolver)

class Q:  # a constant operator as qutip hands it over: dims, to('CSR'), full()
    def __init__(self, m):
        self._m = sp.csr_matrix(m, dtype=complex)
        self.dims = [[m.shape[0]], [m.shape[1]]]
    def to(self, kind):
        return types.SimpleNamespace(data=qd.CSR(self._m))
    def full(self):
        return self._m.toarray()
import types
sm = np.array([[0, 0], [1, 0]], dtype=complex)
sz = np.diag([1.0, -1.0]).astype(complex)
s = HMC(Q(sz), [Q(sm)], options={"precision": "fp32", "max_jumps": 8})
out["batch_opts"] = [s._batch.options["precision"], s._batch.options["max_jumps"],
                     s._batch.options["method"]]
out["heff"] = np.round(s._batch.H_eff.toarray(), 12).tolist().__repr__()
print("JSON" + json.dumps(out))
"""


def test_register_against_reference_api():
    r = subprocess.run([sys.executable, "-c", _REGISTER, str(ROOT / "tests" / "stub_qutip"),
                        str(ROOT)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-3000:]
    out = json.loads(r.stdout.split("JSON", 1)[1])
    assert {"hip", "hipdense", "hipcsr"} <= set(out["aliases"])
    assert out["groups"] == ["hip"]
    assert out["conversions"] == 4
    for cls in ("SESolver", "MESolver", "MCSolver"):
        assert out["integrators"][cls] == ["hip_rk"]
    # every dispatcher got its specialisations with the exact signature
    # (the stand-in raises TypeError otherwise)
    assert out["spec"]["matmul"] == 3 and out["spec"]["expect"] == 3
    assert all(v >= 1 for v in out["spec"].values())
    assert "CSR" in out["host_matmul"] and "Dia" in out["host_matmul"]
    assert out["run_overridden"]
    assert out["mc_subclass"]
    assert out["batch_opts"] == ["fp32", 8, "vern7"]
    # H_eff = -i H - 1/2 c^dag c (mcsolve.py:491-496)
    heff = np.array(eval(out["heff"]))
    assert np.allclose(heff, -1j * np.diag([1, -1]) - 0.5 * np.diag([1, 0]))


class _ReferenceLikeMatrixForm:
    """Shaped like QuTiP's LindbladMatrixForm: constant, issuper, public
    H_nh (QobjEvo) and c_ops (list of QobjEvo), and a __call__ that raises."""
    isconstant = True
    issuper = True

    def __init__(self, Hnh, cops):
        self.H_nh = qa.QobjEvo(qa.Qobj(Hnh))
        self.c_ops = [qa.QobjEvo(qa.Qobj(c)) for c in cops]

    def __call__(self, t, _args=None, **kw):
        raise NotImplementedError("LindbladMatrixForm cannot be called to produce an operator.")

    def matmul_data(self, t, rho, out=None):
        raise AssertionError("not called on the CPU")


def test_matrix_form_owner_is_read_not_called():
    rng = np.random.default_rng(3)
    H = rng.standard_normal((6, 6)) + 1j * rng.standard_normal((6, 6))
    c = rng.standard_normal((6, 6)) + 0j
    owner = _ReferenceLikeMatrixForm(H, [c])
    assert hip_rk.constant_operator(owner.matmul_data) is None
    Hn, cs = hip_rk.matrix_form_operators(owner)
    assert np.array_equal(Hn, H) and len(cs) == 1 and np.array_equal(cs[0], c)
    # the product's own matrix form exposes the same attributes
    own = qa.LindbladMatrixForm(qa.Qobj(H + H.conj().T), [qa.Qobj(c)])
    Hn2, cs2 = hip_rk.matrix_form_operators(own)
    assert np.allclose(Hn2, H + H.conj().T - 0.5j * c.conj().T @ c)
    assert own.issuper and own.isconstant
    with pytest.raises(NotImplementedError):
        own(0.0)
    # an owner whose call raises and that is not a matrix form: plain form
    bad = types.SimpleNamespace(isconstant=True)
    bad_call = types.SimpleNamespace(__self__=bad)
    assert hip_rk.constant_operator(bad_call) is None


def test_eops_dims_are_checked():
    ok = [qa.sigmaz(), lambda t, s: t]
    Solver._check_eops(ok, [2])
    with pytest.raises(ValueError, match="incompatible dimensions"):
        Solver._check_eops([qa.num(3)], [2])
    with pytest.raises(ValueError):
        Solver._check_eops([qa.tensor(qa.sigmaz(), qa.qeye(2))], [2])
    with pytest.raises(TypeError):
        Solver._check_eops([3.0], [2])
