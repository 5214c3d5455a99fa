"""CPU oracle for the BAFQMC Markov-chain sweep (test infrastructure only).

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this package, and only as the checker.  The product
path (``bafqmc_amd``) never imports it.
"""
