"""setsm_amd — MI355X-native VerticalLineLocus engine for SETSM.

The product is the C-ABI library libsmx.so (include/smx_vll.h) built from
setsm_amd/csrc; setsm_amd.smx is its ctypes host binding, named after the
reference entry points (InitializeVoxel, VerticalLineLocus, AWNCC).
"""
