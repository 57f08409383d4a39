"""hacs_amd -- the MI355X (gfx950) device backend of HACS's exact matrix engine.

Everything computed here runs in libhacs_mi355x.so (hand-written HIP kernels for
gfx950, C-ABI in include/hacs_mi355x.h).  Importing the package loads that library
and fails if it is missing: there is no CPU fallback.
"""
import torch  # noqa: F401  (first: the library then binds to the HIP runtime torch has loaded, one per process)

from ._lib import HxError, available, lib  # noqa: F401

lib()   # load now: a missing library is an import error, never a silent fallback

__version__ = "0.1.0"
