"""CPU oracle for the MountainSort front end (test infrastructure only).

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline
leg may import this package.  The product path (``mountainsort_amd``) never
imports it: it is the checker, not the thing measured or shipped.

``oracle.core`` holds the numpy restatements and the ctypes bindings of the C
restatement in ``oracle/msoracle.c``.
"""
