"""cuopus_amd -- MI355X-native (gfx950) batch Opus decoder.

Drop-in for the decode path of the reference cuopus: the C ABI in
include/cuopus.h (libcuopus.so) and the Python binding in
``cuopus_amd.cuopus_native``.
"""
from . import cuopus_native  # noqa: F401
from .cuopus_native import (CuopusError, decode_ogg_multi, decode_ogg_to_cuda_tensor,  # noqa: F401
                            decode_ogg_lossy, decode_ogg_to_s16_tensor, probe, release_workspaces,
                            warmup, workspace_bytes, last_arena_overflow)
