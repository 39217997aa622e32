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


def _serving_reference(golden):
    g = golden("serving_seed7.npz")
    return int(g["nreq"]), {int(i): e for i, e in zip(g["ids"], g["expect"])}


def test_seed7_trace_at_the_metric_bar(golden):
    """The serving workload itself: the first 30 requests of the Poisson
    trace (200 req/s, seed 7, d up to 512, ntraj 64..512) at rtol 1e-8 /
    atol 1e-10, served together.  Every result equals the same request
    served alone bit for bit; sesolve <O>(t) is within 1e-6 of the matrix
    exponential, mesolve within 1e-6 of the FP64 fixture (the oracle's
    vern7 on the matrix-form RHS at rtol 1e-10), and each mcsolve ensemble
    mean within 4 standard errors of that master equation (plus 1e-6).

    The collapse operators of this workload are sqrt(gamma) rand_dm(d), so
    a trajectory jumps with probability p(t) = 1 - ||exp(H_eff t) psi0||^2
    of order gamma t / d^2: most requests see no jump at all and their
    sample standard error is 0.  The standard error used is therefore the
    larger of the sample one and the bound 2 ||O|| sqrt(p (1 - p) / ntraj)
    (a trajectory's <O> leaves the no-jump value only after a jump, by at
    most 2 ||O||), with p from the matrix exponential of H_eff."""
    from oracle import exact
    nreq, ref = _serving_reference(golden)
    print(f"serve: the first {nreq} requests of the seed-7 trace", flush=True)
    mix = {r.id: r for r in SolveServer(max_slots=64).run(systems.serving_trace(nreq),
                                                          realtime=False)}
    worst = {"se": 0.0, "me": 0.0, "mc": 0.0}
    for req in systems.serving_trace(nreq):
        m = mix[req.id]
        alone = SolveServer(max_slots=1).run([req], realtime=False)[0]
        assert alone.error is None and m.error is None, (req.id, alone.error, m.error)
        same = np.array_equal(np.asarray(alone.result), np.asarray(m.result))
        got = np.asarray(m.result).real
        if req.kind == "se":
            want = exact.sesolve_expect(req.H.to_csr(), req.psi0.full().ravel(), req.tlist,
                                        [O.to_csr() for O in req.e_ops])
            err = np.abs(got - want).max()
            ok = err < 1e-6
        elif req.kind == "me":
            err = np.abs(got - ref[req.id]).max()
            ok = err < 1e-6
        else:
            import scipy.sparse.linalg as spla
            H = req.H.full()
            heff = -1j * H - 0.5 * sum(c.full().conj().T @ c.full() for c in req.c_ops)
            psis = spla.expm_multiply(heff, req.psi0.full().ravel().astype(complex), start=0.0,
                                      stop=req.tlist[-1], num=len(req.tlist), endpoint=True)
            p = np.clip(1 - np.sum(np.abs(psis) ** 2, axis=1), 0.0, 1.0)
            onorm = np.array([1.0, np.linalg.norm(H, 2)])[:, None]
            se = np.maximum(np.asarray(m.stderr), 2 * onorm * np.sqrt(p * (1 - p) / req.ntraj))
            dev = np.abs(got - ref[req.id])
            err = float((dev / (4 * se + 1e-6)).max())
            ok = err <= 1.0
        worst[req.kind] = max(worst[req.kind], err)
        print(f"  request {req.id} {req.kind} d={req.d} nc={len(req.c_ops)} ntraj={req.ntraj}: "
              f"{'|err|' if req.kind != 'mc' else 'dev/(4 se)'} {err:.2e}, alone==mix {same}",
              flush=True)
        assert same, req.id
        assert ok, (req.id, req.kind, err)
    print(f"  worst: {worst}", flush=True)
