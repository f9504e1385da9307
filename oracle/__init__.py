"""CPU oracle for the RREA electron transport step -- test infrastructure.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` import this package.  The shipped HIP path never does.
"""
