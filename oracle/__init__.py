"""CPU oracle for the split-RK2 step -- test infrastructure, never the product.

``oracle/mom6_oracle.c`` restates the reference serial path in C;
``oracle/ref/`` builds the reference's own serial sources with a golden-vector
harness.  Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s
CPU-baseline leg use this package, and only as the checker.
"""
