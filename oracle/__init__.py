"""CPU oracle for moraine's distributed-scatterer hot path -- TEST INFRASTRUCTURE ONLY.

Nothing in ``moraine_amd`` (the product) imports this package.  It is used by
``tests/`` as the parity checker, by ``__graft_entry__.smoke()`` as the checker of
one small GPU call, and by ``bench.py`` only to time the ``cpu_baseline`` leg.

Contents
--------
``moraine_ref``  numpy restatement of the reference numba kernels (exact semantics:
                 integer KS step count, f64 p series, c128 coherence accumulators,
                 c64 temporal-coherence accumulators) plus FP64 "truth" variants.
``port``         ctypes loader for ``port.c``: a C/OpenMP restatement of the same
                 numba kernels, used for larger parity cases and as the timed CPU
                 baseline (the reference itself needs numba, which is absent).

Pinning status of each function (how a plausible mistake would be caught) is
listed in the header of ``moraine_ref``.
"""
