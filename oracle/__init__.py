"""Oracle for the filtration front end — TEST INFRASTRUCTURE ONLY.

Two checkers live here:

* ``Restatement`` loads ``libttoracle.so``: the plain-C restatement of the
  reference's scalar front end (``ttoracle.c``).
* ``Reference`` loads ``_ref/libttref_nosimd.so`` (or the SIMD build): the
  reference's own headers compiled in place from its sources by
  ``oracle/Makefile``; absent wherever the reference sources were absent at
  build time.

Only tests/, ``__graft_entry__.smoke()`` and bench.py's cpu_baseline leg may
import this package, and only as the checker. The product never does.
"""
from .bindings import (  # noqa: F401
    METRIC_KINDS,
    Complex,
    index_values,
    index_words,
    Reference,
    Restatement,
    parse_metric,
    reference_available,
)
