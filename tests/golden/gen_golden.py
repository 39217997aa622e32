"""Write the golden fixtures of the parity tests (run in the build container).

Every fixture is produced by the oracle's FP64 matrix exponential
(oracle/exact.py, scipy.sparse.linalg.expm_multiply) on the benchmark systems
restated in oracle/models.py -- no qutip import is involved (the reference
is not importable here).  Files:

  jc_mesolve.npz      config "JC N=10 x qubit" mesolve: <a^dag a>, <sigma_z>
  ising8_mesolve.npz  dissipative TFIM L=8 mesolve: <sum sz/L>, <sz1 sz2>
  ising10_mesolve.npz dissipative TFIM L=10 (gamma=0.05, t in [0,20], 100 pts):
                      the master-equation target of the mcsolve ensemble
  kerr_mesolve.npz    Kerr N=256 (first 40 output times, t <= 7.8): <n>, parity
  kerr_full_mesolve.npz  the same over t in [0, 30], 150 points (--kerr-full)
  ising6_sesolve.npz  closed TFIM L=6 sesolve: <sum sz/L>, <sz1 sz2>

Usage: python tests/golden/gen_golden.py [--all | --ising10 | --kerr-full]
(--all adds the L=10 case, which builds a 1M x 1M Liouvillian and takes
several minutes; --ising10 / --kerr-full write only that one).
"""
import pathlib
import sys
import time

import numpy as np

ROOT = pathlib.Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from oracle import exact, models  # noqa: E402

OUT = pathlib.Path(__file__).resolve().parent


def save(name, **kw):
    np.savez_compressed(OUT / name, **kw)
    print("wrote", OUT / name, {k: np.shape(v) for k, v in kw.items()}, flush=True)


def main(argv):
    t0 = time.time()
    if "--kerr-full" in argv:
        m = models.kerr(256)
        ex = exact.mesolve_expect(m["H"], m["c_ops"], m["psi0"], m["tlist"], m["e_ops"])
        save("kerr_full_mesolve.npz", tlist=m["tlist"], expect=ex)
        return
    if "--ising10" in argv:
        m = models.ising(10, 0.05, tf=20.0, npts=100)
        ex = exact.mesolve_expect(m["H"], m["c_ops"], m["psi0"], m["tlist"], m["e_ops"])
        save("ising10_mesolve.npz", tlist=m["tlist"], expect=ex, h=m["h"])
        return
    m = models.jaynes_cummings()
    ex = exact.mesolve_expect(m["H"], m["c_ops"], m["psi0"], m["tlist"], m["e_ops"])
    save("jc_mesolve.npz", tlist=m["tlist"], expect=ex)

    m = models.ising(8, 0.1)
    ex = exact.mesolve_expect(m["H"], m["c_ops"], m["psi0"], m["tlist"], m["e_ops"])
    save("ising8_mesolve.npz", tlist=m["tlist"], expect=ex, h=m["h"])

    m = models.ising(6, 0.1)
    tl = np.linspace(0, 5, 51)
    ex = exact.sesolve_expect(m["H"], m["psi0"], tl, m["e_ops"])
    save("ising6_sesolve.npz", tlist=tl, expect=ex, h=m["h"])

    m = models.kerr(256)
    tl = m["tlist"][:40]
    ex = exact.mesolve_expect(m["H"], m["c_ops"], m["psi0"], tl, m["e_ops"])
    save("kerr_mesolve.npz", tlist=tl, expect=ex)

    if "--all" in argv:
        m = models.ising(10, 0.05, tf=20.0, npts=100)
        ex = exact.mesolve_expect(m["H"], m["c_ops"], m["psi0"], m["tlist"], m["e_ops"])
        save("ising10_mesolve.npz", tlist=m["tlist"], expect=ex, h=m["h"])
    print(f"done in {time.time() - t0:.1f}s")


if __name__ == "__main__":
    main(sys.argv[1:])
