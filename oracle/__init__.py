"""CPU oracle: a NumPy / C restatement of the reference arithmetic.

TEST INFRASTRUCTURE ONLY.  Nothing in the product path (``scico_amd``)
imports, links or executes anything under ``oracle/``.  The only
callers are ``tests/``, ``__graft_entry__.smoke()`` (as the checker) and
the ``cpu_baseline`` leg of ``bench.py`` (timed as the CPU baseline,
never as the measured GPU path).

Every function cites the reference file:line it restates (paths are
relative to the lanl/scico source tree).  Each module header lists how
its functions are pinned (golden vectors from the reference's own
tests, closed forms, brute force on tiny inputs, invariants).
"""
