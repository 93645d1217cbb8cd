"""MI355X (gfx950) implementation of lica_boiling1d's per-step hot path.

The compute runs in ``liblica_mi355x.so`` (hand-written HIP kernels behind
the C ABI declared in ``include/lica_mi355x.h``); this package is the host
mirror of the reference's call sequence on top of that ABI.  There is no
CPU implementation: without the library or a GPU every entry point raises.
"""
from ._abi import (CaseDesc, LicaError, FIELDS, SCALARS, POISSON_THOMAS, POISSON_DIRECT,
                   POISSON_MG64, POISSON_MIXED, POISSON_LEARNED, POISSON_PROJECTED, load, status_string)
from .solver import Batch, Context, ReferenceSeams, ref_array, parse_deck, poisson, trdiag, trdiag_periodic, ptdiag1a
from .workloads import deck_case, single_large_case, batched_cases
from .learned import Net, collect_pairs

__all__ = ["CaseDesc", "LicaError", "FIELDS", "SCALARS", "POISSON_THOMAS", "POISSON_DIRECT",
           "POISSON_MG64", "POISSON_MIXED", "POISSON_LEARNED", "POISSON_PROJECTED", "load", "status_string",
           "Batch", "Context", "ReferenceSeams", "ref_array", "parse_deck", "poisson", "trdiag", "trdiag_periodic", "ptdiag1a", "deck_case", "single_large_case",
           "batched_cases", "Net", "collect_pairs"]
