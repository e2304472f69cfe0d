"""CPU oracle for the NEGF transmission path (test infrastructure only).

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` import this package, and only as the checker.  The shipped
package ``negfpy_amd`` never imports it.
"""
