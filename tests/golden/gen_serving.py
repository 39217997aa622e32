"""Write tests/golden/serving_seed7.npz: FP64 <O>(t) of the mesolve and
mcsolve requests among the first NREQ requests of the serving trace
(Poisson 200 req/s, seed 7: tests/systems.py::serving_trace), from the
oracle's restated vern7 integrator on the matrix-form right-hand side at
rtol 1e-10 / atol 1e-12 (oracle.exact.mesolve_expect_rk).  An mcsolve
request's reference is the master equation of the same system (its
ensemble mean must agree within 3 sigma).  sesolve requests are not stored:
their reference, the matrix exponential, is cheap enough for the test to
compute.

Usage: python tests/golden/gen_serving.py   (about 20 minutes on 8 cores)
"""
import pathlib
import sys
import time

import numpy as np

ROOT = pathlib.Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from oracle import exact  # noqa: E402
from tests import systems  # noqa: E402

NREQ = 30
OUT = ROOT / "tests" / "golden" / "serving_seed7.npz"


def main():
    t0 = time.time()
    reqs = systems.serving_trace(NREQ)
    ids, expect = [], []
    for r in reqs:
        if r.kind == "se":
            continue
        ts = time.time()
        H = r.H.full()
        cs = [c.full() for c in r.c_ops]
        ops = [O.to_csr() for O in r.e_ops]
        ex = exact.mesolve_expect_rk(H, cs, r.psi0.full().ravel(), r.tlist, ops)
        ids.append(r.id)
        expect.append(ex)
        print(f"request {r.id} {r.kind} d={r.d} nc={len(cs)}: {time.time() - ts:.1f} s",
              flush=True)
    np.savez_compressed(OUT, ids=np.asarray(ids, np.int64), expect=np.asarray(expect),
                        tlist=reqs[0].tlist, nreq=NREQ)
    print(f"wrote {OUT} ({len(ids)} requests) in {time.time() - t0:.0f} s", flush=True)


if __name__ == "__main__":
    main()
