"""acdalm -- Python view of the MI355X acd-alm library.

The product is the C library lib/libacdalm.so (headers under include/);
this package only loads it (capi), generates seeded instances (instances)
and computes the KKT residual used to report convergence (kkt).
"""
from . import instances  # noqa: F401
from .capi import CLib, LIB_PATH, load  # noqa: F401
