// stereo_launch.h -- host entry points of the stereo-capable kernel build
// (stereo.hip), called by the runtime in libcuopus.hip.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

// Bytes of one wavefront's HBM heap, of the kernel argument block and of the
// LDS state in the stereo build.
size_t cuo_stereo_heap_bytes();
size_t cuo_stereo_args_bytes();
size_t cuo_stereo_lds_bytes();
// Launch the stereo build's persistent (persistent != 0) or one-workgroup-
// per-chain decode kernel; args points to a DecodeArgs of the mono build's
// layout.
void cuo_stereo_launch(int persistent, uint32_t grid, hipStream_t st, const void *args);
