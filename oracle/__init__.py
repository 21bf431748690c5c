"""CPU oracle for the v4 RIME hot path.

TEST INFRASTRUCTURE ONLY. Nothing in the product path (``montblanc_amd``)
imports this package; only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` use it, and only as the checker.

Modules
-------
rime_np     numpy float64 restatement of the reference's v4 ``CPUSolver``
            ``compute_*`` chain (montblanc/impl/rime/v4/cpu/CPUSolver.py).
rime_torch  torch complex128 twin of ``rime_np`` whose autograd gives the
            reference gradients (checked equal to ``rime_np`` and against
            Richardson-extrapolated central differences of its chi-squared
            in tests/test_oracle_grad.py).
synth       seeded synthetic inputs for the workloads the tests and the bench
            use (numpy.random.default_rng).
parallel    ``rime_np`` over source ranges in one worker process per core
            (the multi-core CPU baseline).

Every function is pinned by checks in tests/test_oracle.py that do not
retype its expression: library routines, Fourier transforms of image-plane
models computed on a grid, invariants, brute force and planted mistakes.
"""
