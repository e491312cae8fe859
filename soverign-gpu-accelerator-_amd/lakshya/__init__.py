"""MI355X-native Mehrotra interior-point path (LAKSHYA device path).

`lakshya.native` binds liblakshya.so (C ABI in include/lakshya.h); `lakshya.gen` builds the
seeded synthetic LP families.  Importing this package does not touch the GPU.
"""
from . import gen  # noqa: F401
from . import native  # noqa: F401

__all__ = ["gen", "native"]
