"""The continuous-batching solve server on the GPU: every result is the same
alone and in the mix (bit for bit: requests share no buffers and their
kernels' reduction order depends only on their own sizes), and the
sesolve / mesolve results agree with the FP64 matrix exponential."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import qutip_amd as qa  # noqa: E402
from qutip_amd.serve import SolveServer  # noqa: E402
from oracle import exact  # noqa: E402
from tests import systems  # noqa: E402


def _trace():
    return systems.serving_trace(14, dims=(16, 32, 64), ntraj=(16, 48), seed=11)


def test_alone_equals_in_the_mix():
    print("serve: mix of 14 requests", flush=True)
    mix = SolveServer(max_slots=16).run(_trace(), realtime=False)
    mix = {r.id: r for r in mix}
    for req in _trace():
        alone = SolveServer(max_slots=1).run([req], realtime=False)[0]
        m = mix[req.id]
        assert alone.error is None and m.error is None, (alone.error, m.error)
        diff = np.abs(np.asarray(alone.result) - np.asarray(m.result)).max()
        print(f"  request {req.id} kind={req.kind} d={req.d}: max |alone - mix| = {diff:.1e}",
              flush=True)
        assert diff <= 1e-12


def test_results_match_fp64():
    reqs = [r for r in _trace() if r.kind in ("se", "me")][:6]
    done = SolveServer().run(reqs, realtime=False)
    for r in done:
        H = r.H.to_csr()
        psi0 = r.psi0.full().ravel()
        ops = [O.to_csr() for O in r.e_ops]
        if r.kind == "se":
            ref = exact.sesolve_expect(H, psi0, r.tlist, ops)
        else:
            ref = exact.mesolve_expect(H, [c.to_csr() for c in r.c_ops], psi0, r.tlist, ops)
        err = np.abs(np.asarray(r.result).real - ref).max() / max(1.0, np.abs(ref).max())
        print(f"  request {r.id} kind={r.kind} d={r.d}: rel err vs FP64 {err:.2e}", flush=True)
        # default integrator tolerances (rtol 1e-6, atol 1e-8); FP32 MFMA for mesolve
        assert err < 1e-4
