"""MI355X-native backend for the Stwo Circle-STARK prover.

The compute path is libstwo_hip.so (hand-written HIP kernels for gfx950 behind
the C-ABI in include/stwo_hip.h); ``proving_amd.backend`` mirrors the
reference's backend trait family on top of it.
"""
from ._lib import HipError, load  # noqa: F401
