"""Parity oracle for the POM step (test infrastructure only).

Only ``tests/``, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` import this package.  The product (``neuralpom_amd``) never does.
"""
