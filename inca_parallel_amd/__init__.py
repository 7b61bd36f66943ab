"""INCA intracule engine for MI355X (gfx950).

The product is libinca_hip.so (HIP kernels behind a C-ABI, include/inca_hip.h)
and the Fortran host `inca` that binds it with ISO_C_BINDING.  `capi` is the
Python binding of the same C-ABI.
"""
