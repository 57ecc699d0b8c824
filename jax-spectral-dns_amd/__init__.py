"""MI355X-native velocity-vorticity channel-flow step (forward, perturbation
and dual) with the Domain / Field / Equation surface of dakling/jax-spectral-dns.

Compute runs only through the HIP library ``lib/libjsd.so`` (C ABI in
``include/jsd.h``); the Python layer builds operator tables, owns device
memory through torch, and mirrors the reference's classes.
"""

__version__ = "0.1.0"

from . import domain_math, initial  # noqa: F401
from ._lib import JsdError, KIND_DUAL, KIND_NS, KIND_PERTURBATION  # noqa: F401
