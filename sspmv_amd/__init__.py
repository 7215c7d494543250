"""MI355X-native (gfx950) SpMV hot path of LeSpMV.

`sspmv_amd.lib` binds the C ABI (include/lespmv_hip.h) of the HIP library
`sspmv_amd/lib/libsspmv_hip.so`; `sspmv_amd.lespmv` mirrors the reference's
C++ interface (LeSpMV_<fmt>, csr_to_<fmt>, test_<fmt>_matrix_kernels) in
Python for the SmartAdpter side.
"""
from . import lib  # noqa: F401

__all__ = ["lib"]
