"""MI355X LSM compaction engine for the xpsair/luda LevelDB fork.

The product is the C-ABI library libluda.so (include/luda.h, sources under
luda_amd/csrc/).  This Python package only declares that ABI for ctypes
callers (ffi) and holds job plumbing used by the tests and bench.py (jobs).
"""
from . import ffi  # noqa: F401
