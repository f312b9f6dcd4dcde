// stereo.hip -- the stereo-capable build of the decode kernels.
//
// The decoder headers compiled with two coded channels and one or two output
// channels (CUO_MAXCH 2) in namespace cuo_st: the same kernels as the mono
// build (one wavefront per chain, persistent or one workgroup per chain), with
// a 13 KB LDS block (the SILK excitation buffers move to the heap) and a
// larger HBM heap per wavefront, three wavefronts per SIMD.  The host runtime
// (libcuopus.hip) launches them for launches holding a stream whose OpusHead
// says two channels or that has stereo packets; their argument block has the
// mono build's layout (pointers and counts only).
#define CUO_MAXCH 2
#define CUO_NS cuo_st
#include "kernels.h"
#include "stereo_launch.h"
#include <string.h>

size_t cuo_stereo_heap_bytes() { return sizeof(cuo_st::DecHeap); }
size_t cuo_stereo_args_bytes() { return sizeof(cuo_st::DecodeArgs); }
size_t cuo_stereo_lds_bytes() { return sizeof(cuo_st::WaveState); }

void cuo_stereo_launch(int persistent, uint32_t grid, hipStream_t st, const void *args)
{
    cuo_st::DecodeArgs a;
    memcpy(&a, args, sizeof(a));
    if (persistent)
        hipLaunchKernelGGL(cuo_st::k_decode_persistent, dim3(grid), dim3(64), 0, st, a);
    else
        hipLaunchKernelGGL(cuo_st::k_decode_chains, dim3(grid), dim3(64), 0, st, a);
}
