"""cpanel_amd: MI355X (gfx950) implementation of the CPanel solve hot path.

Native library: cpanel_amd/libcpgpu.so (C ABI in include/cpgpu.h).
"""
from .abi import CP_FP32_REFINE, CP_FP32_REFINE64, CP_FP64, lib  # noqa: F401
from .case import Context, CPanelError, Geometry  # noqa: F401
