"""CPU oracle for the time-evolution hot path -- TEST INFRASTRUCTURE ONLY.

Nothing in the product (``qutip_amd``) imports this package.  Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` use it,
and only as the checker / the timed CPU restatement, never as a code path of
the product.

Contents and how each function is pinned (a function without a pin says so):

* ``tableaus``   published Verner-7 / Tsitouras-5 coefficients.  Pinned by the
                 order conditions and by measured convergence order
                 (tests/test_oracle.py).
* ``rk``         numpy restatement of qutip/solver/integrator/explicit_rk.pyx
                 (step control, error norm, dense output).  Pinned by closed
                 forms (dy/dt = lambda y, the analytic system of
                 qutip/tests/solver/test_integrator.py:23-56) and by the FP64
                 matrix exponential (``exact``) on the benchmark systems.
* ``models``     the benchmark operators and the Liouvillian restated from
                 qutip/core/superoperator.py:123-134 and
                 qutip/core/operators.py.  Pinned by invariants (trace and
                 Hermiticity preservation) and by an independent dense
                 construction on small sizes, plus the closed-form decay of
                 qutip/tests/solver/test_mesolve.py:84-95.
                 ``matrix_form_rhs`` (LindbladMatrixForm, the A + A^dag
                 form) is pinned against that Liouvillian on Hermitian
                 states of random systems.  ``td_rhs`` / ``str_coefficient``
                 (the list-form QobjEvo with string coefficients) are
                 pinned, through ``rk``, by closed forms: a qubit driven by
                 f(t) sigma_x (P = sin^2 F(t), F the integral of f), with
                 sigma_x dephasing (<sigma_z> = cos 2F e^{-2 gamma t}), and
                 by scipy's solve_ivp on a small driven chain.
* ``exact``      vec(rho(t)) = expm(L t) vec(rho0) through
                 scipy.sparse.linalg.expm_multiply -- a library routine, the
                 FP64 reference the parity tests compare against.
                 ``mesolve_expect_rk`` (``rk`` on ``matrix_form_rhs`` at
                 rtol 1e-10, for systems whose superoperator is too large)
                 is pinned against ``mesolve_expect`` on small random
                 systems.
* ``mc``         numpy restatement of the mcsolve jump loop
                 (qutip/solver/mcsolve.py:245-406).  Pinned statistically
                 (ensemble mean within 3 sigma of the ``exact`` master-equation
                 solution) and by the qubit-decay closed form
                 (qutip/tests/solver/test_qubit_evolution.py:80-101).
* ``c/``         C restatement of the CSR spmv (qutip/core/data/src/
                 matmul_csr_vector.cpp:107-149) and of the explicit RK loop;
                 the CPU baseline of bench.py.  Its spmv is pinned bit-exactly
                 against the reference's own C++ spmv built into
                 ``oracle/_ref`` (tests/test_oracle_c.py) and its solver
                 against ``rk``.
"""
