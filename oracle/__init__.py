"""TEST INFRASTRUCTURE ONLY: the serial FP64 CPU restatement of the
reference's per-step path (lica_oracle.c) and its ctypes binding.

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg use
this package, as the checker; the product library never links or calls it.
"""
from .binding import Oracle, load, build, ORC_FIELDS  # noqa: F401
