"""A stand-in for the parts of QuTiP 5's public API that
qutip_amd.qutip_plugin.register() touches, for the CPU test of the
registration (qutip itself is not importable where the tests run).

Each piece restates the interface of the reference file it names:
signatures, inputs and output flags of the data-layer dispatchers, the
``to`` conversion registry and the solvers' ``add_integrator``.  It holds no
arithmetic: the test checks what register() hands over, not results.
"""
__version__ = "5.4.0-stub"
