"""MI355X (gfx950) implementation of COMER2's libmycu search path.

The compute path is the HIP library ``libcomer2_hip.so`` behind the C ABI in
``include/comer2_hip.h``; ``comer2_amd.host`` is a thin Python mirror of the
reference's ``CuBatchProcessing`` interface over that ABI.
"""
__all__ = ["tables"]
