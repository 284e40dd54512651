"""isai_kernel_amd -- MI355X-native block-sparse ILU(0)/ISAI preconditioning.

The product is the C-ABI shared library ``lib/libisai_amd.so`` (HIP kernels for
gfx950 behind the entry points declared in ``include/*.h``).  This package
exposes that library to Python through ctypes for tests, benchmarks and the
autograd harness; it adds no numerics of its own.
"""
from ._ffi import LIB_PATH, DeviceArray, IsaiError, SetupOpts, call, check, lib, ptr  # noqa: F401
from . import problems  # noqa: F401

__all__ = ["LIB_PATH", "DeviceArray", "IsaiError", "SetupOpts", "call", "check", "lib", "ptr", "problems"]
