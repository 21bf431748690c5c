"""CPU oracle for the v4 RIME hot path.

TEST INFRASTRUCTURE ONLY. Nothing in the product path (``montblanc_amd``)
imports this package; only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` use it, and only as the checker.

Modules
-------
rime_np    numpy float64 restatement of the reference's v4 ``CPUSolver``
           ``compute_*`` chain (montblanc/impl/rime/v4/cpu/CPUSolver.py).
grad_np    hand-derived float64 reverse-mode gradient of chi-squared through
           that chain (checked against central finite differences of
           ``rime_np`` and against torch float64 autograd in the tests).
synth      seeded synthetic inputs for the workloads the tests and the bench
           use (numpy.random.default_rng).
"""
