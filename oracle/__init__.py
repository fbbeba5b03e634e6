"""CPU FP64 oracle for the plane-Couette KMM SBDF3 step — test infrastructure only.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this package, and only as the checker.  The shipped
package ``jaxfun_amd`` never imports it.
"""
