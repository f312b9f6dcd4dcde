// common.h -- qualifiers and the exact fixed-point arithmetic of the decoder.
//
// Every helper reproduces the integer semantics of libopus 1.5.2 built
// FIXED_POINT on a 64-bit host (OPUS_FAST_INT64 = 1): the CELT macro layer
// (vendor/celt/fixed_generic.h:40-190, arch.h:130-150), the SILK macro layer
// (vendor/silk/macros.h:40-110, SigProc_FIX.h) and the shared math helpers
// (vendor/celt/mathops.{c,h}).  Products that libopus forms in 64 bits are
// formed in 64 bits here; sums that libopus lets wrap (the *_ovflw macros) wrap
// here through unsigned arithmetic.  Results are returned as int32; callers
// assign to int16 exactly where libopus stores into an opus_val16/opus_int16.
//
// All functions are __host__ __device__ so the same source runs on gfx950 and,
// for the same-source CPU gate in tests/, on the host.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <type_traits>

// The wavefront-parallel versions of the SILK pulse, parameter and
// synthesis stages and of the CELT codeword decode exist on the device only;
// their serial twins are what the host build runs (and what the sanitizer
// build checks).  A device build with CUO_SERIAL_TWINS runs the serial twins
// on the GPU instead (test builds only: it lets a test compare every lane
// version with its twin on the same inputs).
#if defined(__HIP_DEVICE_COMPILE__) && !defined(CUO_SERIAL_TWINS)
#define CUO_LANE_PATHS 1
#else
#define CUO_LANE_PATHS 0
#endif

#define CUO_FN __host__ __device__ inline
#define CUO_NOINLINE __host__ __device__ __attribute__((noinline))
#define CUO_FI __host__ __device__ inline __attribute__((always_inline))

// The decoder compiles twice: a mono build (CUO_MAXCH 1, namespace cuo: the
// drop-in path's kernels, four wavefronts per SIMD) and a stereo-capable
// build (CUO_MAXCH 2, namespace cuo_st, stereo.hip: two coded channels, one
// or two output channels, more LDS per wavefront).  Channel loops run to a
// runtime count bounded by kMaxCh, so the mono build compiles them away.
#ifndef CUO_MAXCH
#define CUO_MAXCH 1
#endif
#ifndef CUO_NS
#define CUO_NS cuo
#endif

namespace CUO_NS {

constexpr int kMaxCh = CUO_MAXCH;
static_assert(kMaxCh == 1 || kMaxCh == 2, "one or two channels");

// Wave-uniform LDS pointer.  All decoder state, scratch and frame output of a
// wavefront live in one LDS block (g_wave_lds, one 64-lane wavefront per
// workgroup) and every lane runs the decoder on the same data.  On the
// device, lds_uniform(p) rebuilds a pointer into that block from the block's
// own address plus a scalar offset, so the compiler sees an LDS address that
// is the same in every lane: accesses become LDS instructions and the values
// loaded are uniform (scalar registers, scalar ALU).  The host build
// (same-source CPU gate) uses the pointer as it is.
constexpr size_t kWaveLdsBytes = kMaxCh == 1 ? 10 * 1024 : 13 * 1024;
__shared__ __attribute__((aligned(16))) char g_wave_lds[kWaveLdsBytes];
#if defined(__HIP_DEVICE_COMPILE__)

template <class T>
__device__ __forceinline__ T *lds_uniform(T *p)
{
    uint32_t off = (uint32_t)((const char *)p - (const char *)g_wave_lds);
    off = __builtin_amdgcn_readfirstlane(off);
    return (T *)(g_wave_lds + off);
}
#else
template <class T>
__host__ __device__ inline T *lds_uniform(T *p)
{
    return p;
}
#endif

// Wave-uniform value: on the device, moves a value every lane holds alike
// into a scalar register, so what is computed from it runs on the scalar ALU
// and table lookups indexed by it become scalar loads.  Identity on the host.
#if defined(__HIP_DEVICE_COMPILE__)
template <class T>
__device__ __forceinline__ T uniform(T v)
{
    static_assert(sizeof(T) <= 4, "32-bit values only");
    return (T)__builtin_amdgcn_readfirstlane((int)v);
}
#else
template <class T>
__host__ __device__ inline T uniform(T v)
{
    return v;
}
#endif

// Table and packet-byte reads with a wave-uniform address.  On the device the
// aligned 32-bit word holding the element is read through the constant
// address space, which the compiler turns into a scalar (K-cache) load; the
// element is then extracted with a shift.  Only for memory that no kernel
// writes while the decoder runs: the constant tables and the packet bytes
// (whose device buffer is padded so the word of the last byte is inside it).
#if defined(__HIP_DEVICE_COMPILE__)
__device__ __forceinline__ uint64_t uniform_u64(uint64_t v)
{
    uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)v);
    uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(v >> 32));
    return (uint64_t)hi << 32 | lo;
}
__device__ __forceinline__ uint32_t const_word(const void *p)
{
    uint64_t a = uniform_u64((uint64_t)(uintptr_t)p);
    typedef __attribute__((address_space(4))) const uint32_t cw_t;
    return *(cw_t *)(uintptr_t)(a & ~(uint64_t)3);
}
__device__ __forceinline__ uint32_t ld_u8(const uint8_t *t, int i)
{
    const uint8_t *p = t + i;
    return (const_word(p) >> (8 * ((uint32_t)(uintptr_t)p & 3))) & 0xFFu;
}
__device__ __forceinline__ int32_t ld_s16(const int16_t *t, int i)
{
    const int16_t *p = t + i;
    uint32_t w = const_word(p);
    return ((uint32_t)(uintptr_t)p & 2) ? (int32_t)w >> 16 : (int32_t)(int16_t)w;
}
#else
__host__ __device__ inline uint32_t ld_u8(const uint8_t *t, int i) { return t[i]; }
__host__ __device__ inline int32_t ld_s16(const int16_t *t, int i) { return t[i]; }
#endif

// Lanes of the wavefront.  The decoder's control flow is wave-uniform; loops
// whose iterations are independent spread their iterations over the 64 lanes
// (CUO_PFOR).  Each iteration's arithmetic is the serial loop's, so results
// are identical.  Writes made in such a loop are read afterwards by other
// lanes through LDS, which executes a wavefront's accesses in program order.
// On the host there is one lane.
#if defined(__HIP_DEVICE_COMPILE__)
constexpr int kLanes = 64;
// The lane index is formed where it is used (two VALU instructions): as a
// volatile asm it is not hoisted out of the chain loop, where the per-lane
// offsets derived from it would otherwise stay live across the whole inlined
// frame decoder and be spilled.
__device__ __forceinline__ int lane_id()
{
    int l;
    asm volatile("v_mbcnt_lo_u32_b32 %0, -1, 0\n\tv_mbcnt_hi_u32_b32 %0, -1, %0" : "=v"(l));
    return l;
}
#else
constexpr int kLanes = 1;
__host__ __device__ inline int lane_id() { return 0; }
#endif
#define CUO_PFOR(i, lo, hi) for (int i = (lo) + lane_id(); i < (hi); i += kLanes)

// Register-blocked serial recurrences: a block of kLanes consecutive elements
// is loaded one per lane, the recurrence walks the block reading element k
// with lane_get(v, k) (a scalar register) and collecting results with
// lane_set(v, k, x); the block is then stored one per lane.  The serial
// order of the arithmetic is unchanged.  k and x must be wave-uniform and
// every lane active.
//
// serial_lanes(nb, f) calls f(k) for k = 0 .. nb-1 in order.  On the device
// the steps come in unrolled chunks of 8 where k is a compile-time LaneK<k>,
// so lane_get / lane_set take immediate lane selects (lane_set is then one
// v_writelane_b32) and the reads of later elements can be issued ahead of
// the recurrence; the last nb % 8 steps take a runtime k.
template <int K>
using LaneK = std::integral_constant<int, K>;
#if defined(__HIP_DEVICE_COMPILE__)
__device__ __forceinline__ int32_t lane_get(int32_t v, int k) { return __builtin_amdgcn_readlane(v, k); }
// Mask of the lanes where p holds.
__device__ __forceinline__ uint64_t wave_ballot(bool p) { return __builtin_amdgcn_ballot_w64(p); }
// Mask of the lanes where `a Pred b` holds, straight from one v_cmp (a ballot
// of a combined condition first turns the compare masks into a per-lane 0/1
// and compares that again); combine masks with & and |.
enum { kCmpEq = 32, kCmpNe, kCmpUgt, kCmpUge, kCmpUlt, kCmpUle, kCmpSgt, kCmpSge, kCmpSlt, kCmpSle };
template <int Pred>
__device__ __forceinline__ uint64_t lanes_u(uint32_t a, uint32_t b) { return __builtin_amdgcn_uicmp(a, b, Pred); }
template <int Pred>
__device__ __forceinline__ uint64_t lanes_s(int32_t a, int32_t b) { return __builtin_amdgcn_sicmp(a, b, Pred); }
__device__ __forceinline__ int32_t lane_set(int32_t v, int k, int32_t x) { return lane_id() == k ? x : v; }
template <int K>
__device__ __forceinline__ int32_t lane_get(int32_t v, LaneK<K>)
{
    return __builtin_amdgcn_readlane(v, K);
}
template <int K>
__device__ __forceinline__ int32_t lane_set(int32_t v, LaneK<K>, int32_t x)
{
    asm("v_writelane_b32 %0, %1, %2" : "=v"(v) : "s"(__builtin_amdgcn_readfirstlane(x)), "i"(K), "0"(v));
    return v;
}
// lane_set for a value the compiler already holds in a scalar register (no
// readfirstlane in front of the write, which it would otherwise move into
// the value's computation and duplicate part of it).  A value it cannot
// prove uniform fails to compile here.
template <int K>
__device__ __forceinline__ int32_t lane_set_s(int32_t v, LaneK<K>, int32_t x)
{
    asm("v_writelane_b32 %0, %1, %2" : "=v"(v) : "s"(x), "i"(K), "0"(v));
    return v;
}
__device__ __forceinline__ int32_t lane_set_s(int32_t v, int k, int32_t x) { return lane_id() == k ? x : v; }
template <int K, int E, class F>
__device__ __forceinline__ void serial_unrolled(F &f)
{
    if constexpr (K < E) {
        f(LaneK<K>());
        serial_unrolled<K + 1, E>(f);
    }
}
template <int C, class F>
__device__ __forceinline__ int serial_chunks(int nb, F &f)
{
    if constexpr (C < 64) {
        if (C + 8 <= nb) {
            serial_unrolled<C, C + 8>(f);
            return serial_chunks<C + 8>(nb, f);
        }
    }
    return C;
}
template <class F>
__device__ __forceinline__ void serial_lanes(int nb, F &&f)
{
    for (int k = serial_chunks<0>(nb, f); k < nb; k++) f(k);
}
#else
__host__ __device__ inline int32_t lane_get(int32_t v, int) { return v; }
__host__ __device__ inline int32_t lane_set(int32_t, int, int32_t x) { return x; }
__host__ __device__ inline int32_t lane_set_s(int32_t, int, int32_t x) { return x; }
template <class F>
inline void serial_lanes(int nb, F &&f)
{
    for (int k = 0; k < nb; k++) f(k);
}
#endif

// 16-lane row primitives (DPP row shifts; device only).  row16_prefix_sum:
// inclusive wrapping sums within each row, so lane 15 holds the sum of lanes
// 0..15.  row16_shift_in: every row moves up by one lane and the row's first
// lane takes fill.
#if defined(__HIP_DEVICE_COMPILE__)
__device__ __forceinline__ int32_t row16_prefix_sum(int32_t v)
{
    v = (int32_t)((uint32_t)v + (uint32_t)__builtin_amdgcn_update_dpp(0, v, 0x111, 0xF, 0xF, false));
    v = (int32_t)((uint32_t)v + (uint32_t)__builtin_amdgcn_update_dpp(0, v, 0x112, 0xF, 0xF, false));
    v = (int32_t)((uint32_t)v + (uint32_t)__builtin_amdgcn_update_dpp(0, v, 0x114, 0xF, 0xF, false));
    v = (int32_t)((uint32_t)v + (uint32_t)__builtin_amdgcn_update_dpp(0, v, 0x118, 0xF, 0xF, false));
    return v;
}
__device__ __forceinline__ int32_t row16_shift_in(int32_t v, int32_t fill)
{
    return __builtin_amdgcn_update_dpp(fill, v, 0x111, 0xF, 0xF, false);
}
#endif

#if defined(__HIP_DEVICE_COMPILE__)
// A byte table held in one VGPR: lane w holds bytes 4w..4w+3 of a 4-byte
// aligned table of at most 256 bytes (the last word's bytes past the end read
// as zero).  One vector load per lane, then every lookup is a v_readlane.
__device__ __forceinline__ int32_t lane_table(const uint8_t *t, int nbytes)
{
    const int w = lane_id();
    if (4 * w + 4 <= nbytes) return (int32_t)((const uint32_t *)t)[w];
    uint32_t v = 0;
    for (int b = 0; b < 3; b++)
        if (4 * w + b < nbytes) v |= (uint32_t)t[4 * w + b] << (8 * b);
    return (int32_t)v;
}
// byte i (wave-uniform) of a lane-held table
__device__ __forceinline__ int lane_byte(int32_t tv, int i)
{
    return (int)(((uint32_t)lane_get(tv, i >> 2) >> (8 * (i & 3))) & 0xFFu);
}
#endif

// Reductions over the lanes of per-lane partial results (identity on the
// host).  Integer sums wrap like the serial loop's, so any order is exact.
// Call with every lane active (outside CUO_PFOR bodies).
#if defined(__HIP_DEVICE_COMPILE__)
// Each is an inclusive 16-lane DPP scan (row_shr 1, 2, 4, 8; lanes shifted in
// from outside the row read the operation's identity) whose row results, in
// lanes 15, 31, 47 and 63, are combined in scalar registers.
template <class Op>
__device__ __forceinline__ int32_t wave_reduce(int32_t v, int32_t ident, Op op)
{
    v = op(v, __builtin_amdgcn_update_dpp(ident, v, 0x111, 0xF, 0xF, false));
    v = op(v, __builtin_amdgcn_update_dpp(ident, v, 0x112, 0xF, 0xF, false));
    v = op(v, __builtin_amdgcn_update_dpp(ident, v, 0x114, 0xF, 0xF, false));
    v = op(v, __builtin_amdgcn_update_dpp(ident, v, 0x118, 0xF, 0xF, false));
    return op(op(__builtin_amdgcn_readlane(v, 15), __builtin_amdgcn_readlane(v, 31)),
              op(__builtin_amdgcn_readlane(v, 47), __builtin_amdgcn_readlane(v, 63)));
}
__device__ __forceinline__ int32_t wave_sum(int32_t v)
{
    return wave_reduce(v, 0, [](int32_t a, int32_t b) { return (int32_t)((uint32_t)a + (uint32_t)b); });
}
__device__ __forceinline__ int32_t wave_max(int32_t v)
{
    return wave_reduce(v, INT32_MIN, [](int32_t a, int32_t b) { return a > b ? a : b; });
}
__device__ __forceinline__ int32_t wave_min(int32_t v)
{
    return wave_reduce(v, INT32_MAX, [](int32_t a, int32_t b) { return a < b ? a : b; });
}
__device__ __forceinline__ uint32_t wave_or(uint32_t v)
{
    return (uint32_t)wave_reduce((int32_t)v, 0, [](int32_t a, int32_t b) { return a | b; });
}
#else
__host__ __device__ inline int32_t wave_sum(int32_t v) { return v; }
__host__ __device__ inline int32_t wave_max(int32_t v) { return v; }
__host__ __device__ inline int32_t wave_min(int32_t v) { return v; }
__host__ __device__ inline uint32_t wave_or(uint32_t v) { return v; }
#endif

// Optional cycle accounting (built with -DCUO_PROF only): a scope adds the
// shader clock cycles it spans to g_prof[id], once per wavefront.
constexpr int kProfSlots = 64;  // one per lane; scope ids must stay below
#if defined(CUO_PROF)
__device__ unsigned long long g_prof[kProfSlots];
#endif
#if defined(CUO_PROF) && defined(__HIP_DEVICE_COMPILE__)
// Per-wavefront sums in LDS (lane 0 adds; one wavefront per workgroup), added
// to g_prof once per chain by prof_flush().
__shared__ unsigned long long g_prof_lds[kProfSlots];
struct ProfScope {
    int id;
    unsigned long long t0;
    __device__ explicit ProfScope(int i) : id(i), t0(__builtin_readcyclecounter()) {}
    __device__ ~ProfScope()
    {
        unsigned long long dt = __builtin_readcyclecounter() - t0;
        if (lane_id() == 0) g_prof_lds[id] += dt;
    }
};
__device__ inline void prof_reset()
{
    g_prof_lds[lane_id()] = 0;
}
__device__ inline void prof_flush()
{
    if (g_prof_lds[lane_id()]) atomicAdd(&g_prof[lane_id()], g_prof_lds[lane_id()]);
}
#define CUO_PROF_CAT2(a, b) a##b
#define CUO_PROF_CAT(a, b) CUO_PROF_CAT2(a, b)
#define CUO_PROF_SCOPE(id) ProfScope CUO_PROF_CAT(prof_scope_, __LINE__)(id)
#else
#define CUO_PROF_SCOPE(id)
#define prof_reset()
#define prof_flush()
#endif
enum {
    kProfChain = 0, kProfFrame, kProfCelt, kProfCeltHead, kProfCeltBands, kProfCeltSynth, kProfCeltComb,
    kProfCeltDeemph, kProfSilk, kProfEmit, kProfCeltLost, kProfCeltEnergy, kProfCeltAlloc, kProfCeltTail,
    kProfPvq, kProfCwrsi, kProfExpRot, kProfTheta, kProfLeafFold, kProfBandPre, kProfBandPost, kProfEcUint,
    kProfAllocBisect, kProfAllocTail, kProfQpLeafBits, kProfBandsPrologue, kProfLowbandOut, kProfQpStage,
    kProfSilkIndices, kProfSilkPulses, kProfSilkParams, kProfSilkCore, kProfSilkResamp, kProfSilkPlc,
    kProfSilkNlsf, kProfSilkNlsf2a, kProfSilkCng, kProfSilkPulseCounts, kProfSilkShell, kProfSilkSigns,
    kProfQpLeaf, kProfCount
};
static_assert(kProfCount <= kProfSlots, "profiling scope ids exceed the per-lane slots");

// ---------------------------------------------------------------- basics
CUO_FN int imin(int a, int b) { return a < b ? a : b; }
CUO_FN int imax(int a, int b) { return a > b ? a : b; }
CUO_FN int32_t min32(int32_t a, int32_t b) { return a < b ? a : b; }
CUO_FN int32_t max32(int32_t a, int32_t b) { return a > b ? a : b; }
CUO_FN int16_t min16(int16_t a, int16_t b) { return a < b ? a : b; }
CUO_FN int16_t max16(int16_t a, int16_t b) { return a > b ? a : b; }
CUO_FN int32_t abs32(int32_t a) { return a < 0 ? -a : a; }

// number of significant bits of v (0 for v == 0): EC_ILOG
CUO_FN int ilog(uint32_t v) { return v ? 32 - __builtin_clz(v) : 0; }

// ---------------------------------------------------------------- CELT layer
CUO_FN int32_t shl32(int32_t a, int s) { return (int32_t)((uint32_t)a << s); }
CUO_FN int32_t shr32(int32_t a, int s) { return a >> s; }
CUO_FN int16_t shl16(int32_t a, int s) { return (int16_t)((uint16_t)a << s); }
CUO_FN int32_t pshr32(int32_t a, int s) { return (a + ((1 << s) >> 1)) >> s; }
CUO_FN int32_t vshr32(int32_t a, int s) { return s > 0 ? shr32(a, s) : shl32(a, -s); }
CUO_FN int32_t add32_ovf(int32_t a, int32_t b) { return (int32_t)((uint32_t)a + (uint32_t)b); }
CUO_FN int32_t sub32_ovf(int32_t a, int32_t b) { return (int32_t)((uint32_t)a - (uint32_t)b); }
CUO_FN int32_t neg32_ovf(int32_t a) { return (int32_t)(0u - (uint32_t)a); }
CUO_FN int32_t saturate(int32_t x, int32_t a) { return x > a ? a : (x < -a ? -a : x); }
// (int16_t)(x >> 15) as an int32: on the device one bit-field extract (bits
// 15..30 sign-extended), so a serial recurrence that feeds the value both to
// the next step and to a lane write forms it once.
CUO_FI int32_t shr15_to16(int32_t x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_sbfe(x, 15, 16);
#else
    return (int16_t)(x >> 15);
#endif
}
CUO_FN int16_t sat16(int32_t x) { return (int16_t)(x > 32767 ? 32767 : (x < -32768 ? -32768 : x)); }

CUO_FN int32_t mul16_16(int32_t a, int32_t b) { return (int32_t)(int16_t)a * (int32_t)(int16_t)b; }
CUO_FN int32_t mul16_16su(int32_t a, int32_t b) { return (int32_t)(int16_t)a * (int32_t)(uint16_t)b; }
CUO_FN int32_t mac16_16(int32_t c, int32_t a, int32_t b) { return c + mul16_16(a, b); }
CUO_FN int32_t mul16_16_q15(int32_t a, int32_t b) { return mul16_16(a, b) >> 15; }
CUO_FN int32_t mul16_16_q14(int32_t a, int32_t b) { return mul16_16(a, b) >> 14; }
CUO_FN int32_t mul16_16_q13(int32_t a, int32_t b) { return mul16_16(a, b) >> 13; }
CUO_FN int32_t mul16_16_p15(int32_t a, int32_t b) { return (16384 + mul16_16(a, b)) >> 15; }
CUO_FN int32_t mul16_16_p14(int32_t a, int32_t b) { return (8192 + mul16_16(a, b)) >> 14; }
CUO_FN int32_t mul16_32_q15(int32_t a, int32_t b) { return (int32_t)(((int64_t)(int16_t)a * (int64_t)b) >> 15); }
#if defined(__HIP_DEVICE_COMPILE__)
// high word of b * (a16 << 16) == floor(a16 * b / 2^16)
CUO_FN int32_t mul16_32_q16(int32_t a, int32_t b) { return __mulhi(b, (int32_t)((uint32_t)(uint16_t)a << 16)); }
#else
CUO_FN int32_t mul16_32_q16(int32_t a, int32_t b) { return (int32_t)(((int64_t)(int16_t)a * (int64_t)b) >> 16); }
#endif
CUO_FN int32_t mul16_32_p16(int32_t a, int32_t b) { return (int32_t)(((int64_t)(int16_t)a * (int64_t)b + 32768) >> 16); }
CUO_FN int32_t mul32_32_q31(int32_t a, int32_t b) { return (int32_t)(((int64_t)a * (int64_t)b) >> 31); }
CUO_FN int32_t mul32_32_q16(int32_t a, int32_t b) { return (int32_t)(((int64_t)a * (int64_t)b) >> 16); }
CUO_FN int32_t mac16_32_q16(int32_t c, int32_t a, int32_t b) { return c + mul16_32_q16(a, b); }
// FRAC_MUL16 (vendor/celt/mathops.h)
CUO_FN int32_t frac_mul16(int32_t a, int32_t b) { return (16384 + (int32_t)(int16_t)a * (int16_t)b) >> 15; }
CUO_FN int16_t round16(int32_t x, int a) { return (int16_t)pshr32(x, a); }
CUO_FN int16_t sround16(int32_t x, int a) { return (int16_t)saturate(pshr32(x, a), 32767); }

constexpr int kQ15One = 32767;
constexpr int kSigShift = 12;
constexpr int32_t kSigSat = 300000000;
constexpr int kDbShift = 10;
constexpr int kBitRes = 3;

// celt_ilog2 / celt_zlog2
CUO_FN int celt_ilog2(int32_t x) { return ilog((uint32_t)x) - 1; }
CUO_FN int celt_zlog2(int32_t x) { return x <= 0 ? 0 : celt_ilog2(x); }

// SIG2WORD16 (vendor/celt/fixed_generic.h:182-188)
CUO_FN int16_t sig2word16(int32_t x)
{
    x = pshr32(x, kSigShift);
    x = max32(x, -32768);
    x = min32(x, 32767);
    return (int16_t)x;
}

// isqrt32: floor(sqrt(v)), v > 0 (vendor/celt/mathops.c:43-63)
CUO_FN uint32_t isqrt32(uint32_t v)
{
    uint32_t g = 0;
    int bshift = (ilog(v) - 1) >> 1;
    uint32_t b = 1u << bshift;
    do {
        uint32_t t = ((g << 1) + b) << bshift;
        if (t <= v) { g += b; v -= t; }
        b >>= 1;
        bshift--;
    } while (bshift >= 0);
    return g;
}

// ADD16 / SUB16 exactly as the macros: operands narrowed to int16 first;
// ADD16 narrows the sum, SUB16 does not.
CUO_FN int16_t add16(int32_t a, int32_t b) { return (int16_t)((int16_t)a + (int16_t)b); }
CUO_FN int32_t sub16(int32_t a, int32_t b) { return (int32_t)(int16_t)a - (int32_t)(int16_t)b; }

// celt_rcp: Q15 reciprocal approximation (vendor/celt/mathops.c:181-205)
CUO_FN int32_t celt_rcp(int32_t x)
{
    int i = celt_ilog2(x);
    int16_t n = (int16_t)(vshr32(x, i - 15) - 32768);
    int16_t r = add16(30840, mul16_16_q15(-15420, n));
    r = (int16_t)sub16(r, mul16_16_q15(r, add16(mul16_16_q15(r, n), add16(r, -32768))));
    r = (int16_t)sub16(r, add16(1, mul16_16_q15(r, add16(mul16_16_q15(r, n), add16(r, -32768)))));
    return vshr32((int32_t)r, i - 16);
}

CUO_FN int32_t celt_div(int32_t a, int32_t b) { return mul32_32_q31(a, celt_rcp(b)); }

// frac_div32 (vendor/celt/mathops.c:70-89)
CUO_FN int32_t frac_div32(int32_t a, int32_t b)
{
    int shift = celt_ilog2(b) - 29;
    a = vshr32(a, shift);
    b = vshr32(b, shift);
    int16_t rcp = round16(celt_rcp(round16(b, 16)), 3);
    int32_t result = mul16_32_q15(rcp, a);
    int32_t rem = pshr32(a, 2) - mul32_32_q31(result, b);
    result = result + shl32(mul16_32_q15(rcp, rem), 2);
    if (result >= 536870912) return 2147483647;
    if (result <= -536870912) return -2147483647;
    return shl32(result, 2);
}

// celt_rsqrt_norm: Q14 1/sqrt(x) for Q16 x in [0.25,1) (vendor/celt/mathops.c:91-117)
CUO_FN int16_t celt_rsqrt_norm(int32_t x)
{
    int16_t n = (int16_t)(x - 32768);
    int16_t r = add16(23557, mul16_16_q15(n, add16(-13490, mul16_16_q15(n, 6713))));
    int16_t r2 = (int16_t)mul16_16_q15(r, r);
    int16_t y = shl16(sub16(add16(mul16_16_q15(r2, n), r2), 16384), 1);
    return add16(r, mul16_16_q15(r, mul16_16_q15(y, sub16(mul16_16_q15(y, 12288), 16384))));
}

// celt_sqrt (vendor/celt/mathops.c:119-140)
CUO_FN int32_t celt_sqrt(int32_t x)
{
    if (x == 0) return 0;
    if (x >= 1073741824) return 32767;
    int k = (celt_ilog2(x) >> 1) - 7;
    x = vshr32(x, 2 * k);
    int16_t n = (int16_t)(x - 32768);
    int32_t rt = add16(23175, mul16_16_q15(n, add16(11561, mul16_16_q15(n, add16(-3011,
                 mul16_16_q15(n, add16(1699, mul16_16_q15(n, -664))))))));
    return vshr32(rt, 7 - k);
}

// _celt_cos_pi_2 / celt_cos_norm (vendor/celt/mathops.c:142-179)
CUO_FN int16_t cos_pi_2(int16_t x)
{
    int16_t x2 = (int16_t)mul16_16_p15(x, x);
    int32_t v = sub16(32767, x2) + mul16_16_p15(x2, -7651 + mul16_16_p15(x2, 8277 + mul16_16_p15(-626, x2)));
    return add16(1, min32(32766, v));
}
CUO_FN int16_t celt_cos_norm(int32_t x)
{
    x = x & 0x0001ffff;
    if (x > (1 << 16)) x = (1 << 17) - x;
    if (x & 0x00007fff) {
        if (x < (1 << 15)) return cos_pi_2((int16_t)x);
        return (int16_t)-cos_pi_2((int16_t)(65536 - x));
    }
    if (x & 0x0000ffff) return 0;
    if (x & 0x0001ffff) return -32767;
    return 32767;
}

// celt_exp2_frac / celt_exp2 / celt_log2 (vendor/celt/mathops.h, FIXED_POINT)
CUO_FN int32_t celt_exp2_frac(int16_t x)
{
    int16_t f = shl16(x, 4);
    return add16(16383, mul16_16_q15(f, add16(22804, mul16_16_q15(f, add16(14819, mul16_16_q15(10204, f))))));
}
CUO_FN int32_t celt_exp2(int16_t x)
{
    int integer = x >> 10;
    if (integer > 14) return 0x7f000000;
    if (integer < -15) return 0;
    int16_t frac = (int16_t)celt_exp2_frac((int16_t)(x - shl16(integer, 10)));
    return vshr32((int32_t)frac, -integer - 2);
}
CUO_FN int16_t celt_log2(int32_t x)
{
    if (x == 0) return -32767;
    int i = celt_ilog2(x);
    int16_t n = (int16_t)(vshr32(x, i - 15) - 32768 - 16384);
    int16_t frac = add16(-6793, mul16_16_q15(n, add16(15746, mul16_16_q15(n, add16(-5217,
                   mul16_16_q15(n, add16(2545, mul16_16_q15(n, -1401))))))));
    return (int16_t)(shl16(i - 13, kDbShift) + (frac >> (14 - kDbShift)));
}

CUO_FN uint32_t lcg_rand(uint32_t seed) { return 1664525u * seed + 1013904223u; }

// n steps of lcg_rand at once: lcg_rand^n(s) = A_n s + C_n (mod 2^32) with
// A_n = A_{n-1} a, C_n = C_{n-1} a + c.  Lets lanes draw a band's noise in
// parallel with the exact serial sequence.
constexpr int kLcgJumpMax = 256;
struct LcgJumpTable {
    uint32_t a[kLcgJumpMax + 1], c[kLcgJumpMax + 1];
};
constexpr LcgJumpTable make_lcg_jump_table()
{
    LcgJumpTable t{};
    t.a[0] = 1u;
    t.c[0] = 0u;
    for (int n = 1; n <= kLcgJumpMax; n++) {
        t.a[n] = t.a[n - 1] * 1664525u;
        t.c[n] = t.c[n - 1] * 1664525u + 1013904223u;
    }
    return t;
}
alignas(16) static constexpr LcgJumpTable kLcgJump = make_lcg_jump_table();
CUO_FN uint32_t lcg_jump(uint32_t seed, int n)
{
    if (n <= kLcgJumpMax) return kLcgJump.a[n] * seed + kLcgJump.c[n];
    for (int i = 0; i < n; i++) seed = lcg_rand(seed);
    return seed;
}

// ---------------------------------------------------------------- SILK layer
#if defined(__HIP_DEVICE_COMPILE__)
// (a * b16) >> 16 is the high word of a * (b16 << 16): one multiply instead
// of a 64-bit product and shift.  The value is the same: floor(a*b16 / 2^16)
// fits in 32 bits for 32-bit a and 16-bit b16.
CUO_FN int32_t silk_smulwb(int32_t a, int32_t b) { return __mulhi(a, (int32_t)((uint32_t)(uint16_t)b << 16)); }
CUO_FN int32_t silk_smlawb(int32_t a, int32_t b, int32_t c)
{
    return (int32_t)((uint32_t)a + (uint32_t)__mulhi(b, (int32_t)((uint32_t)(uint16_t)c << 16)));
}
CUO_FN int32_t silk_smulwt(int32_t a, int32_t b) { return __mulhi(a, (int32_t)((uint32_t)b & 0xFFFF0000u)); }
CUO_FN int32_t silk_smlawt(int32_t a, int32_t b, int32_t c)
{
    return (int32_t)((uint32_t)a + (uint32_t)__mulhi(b, (int32_t)((uint32_t)c & 0xFFFF0000u)));
}
#else
CUO_FN int32_t silk_smulwb(int32_t a, int32_t b) { return (int32_t)(((int64_t)a * (int64_t)(int16_t)b) >> 16); }
CUO_FN int32_t silk_smlawb(int32_t a, int32_t b, int32_t c) { return (int32_t)(a + (((int64_t)b * (int64_t)(int16_t)c) >> 16)); }
CUO_FN int32_t silk_smulwt(int32_t a, int32_t b) { return (int32_t)(((int64_t)a * (int64_t)(b >> 16)) >> 16); }
CUO_FN int32_t silk_smlawt(int32_t a, int32_t b, int32_t c) { return (int32_t)(a + (((int64_t)b * (int64_t)(c >> 16)) >> 16)); }
#endif
CUO_FN int32_t silk_smulbb(int32_t a, int32_t b) { return (int32_t)(int16_t)a * (int32_t)(int16_t)b; }
CUO_FN int32_t silk_smlabb(int32_t a, int32_t b, int32_t c) { return a + (int32_t)(int16_t)b * (int32_t)(int16_t)c; }
CUO_FN int32_t silk_smulbt(int32_t a, int32_t b) { return (int32_t)(int16_t)a * (b >> 16); }
CUO_FN int32_t silk_smulww(int32_t a, int32_t b) { return (int32_t)(((int64_t)a * b) >> 16); }
CUO_FN int32_t silk_smlaww(int32_t a, int32_t b, int32_t c) { return (int32_t)(a + (((int64_t)b * c) >> 16)); }
CUO_FN int64_t silk_smull(int32_t a, int32_t b) { return (int64_t)a * (int64_t)b; }
CUO_FN int32_t silk_mul(int32_t a, int32_t b) { return (int32_t)((uint32_t)a * (uint32_t)b); }
CUO_FN int32_t silk_mla(int32_t a, int32_t b, int32_t c) { return (int32_t)((uint32_t)a + (uint32_t)b * (uint32_t)c); }
CUO_FN int32_t silk_rshift_round(int32_t a, int s) { return s == 1 ? (a >> 1) + (a & 1) : ((a >> (s - 1)) + 1) >> 1; }
CUO_FN int64_t silk_rshift_round64(int64_t a, int s) { return s == 1 ? (a >> 1) + (a & 1) : ((a >> (s - 1)) + 1) >> 1; }
CUO_FN int32_t silk_lshift(int32_t a, int s) { return (int32_t)((uint32_t)a << s); }
CUO_FN int32_t silk_add_lshift32(int32_t a, int32_t b, int s) { return (int32_t)((uint32_t)a + ((uint32_t)b << s)); }
// silk_ADD_SAT32: the wrapped sum unless both operands share a sign the sum
// lacks, then the limit of that sign (written without branches).
CUO_FN int32_t silk_add_sat32(int32_t a, int32_t b)
{
    const uint32_t s = (uint32_t)a + (uint32_t)b;
    const bool ov = (int32_t)(((uint32_t)a ^ s) & ((uint32_t)b ^ s)) < 0;
    const int32_t lim = (int32_t)((uint32_t)(a >> 31) ^ 0x7FFFFFFFu);
    return ov ? lim : (int32_t)s;
}
CUO_FN int32_t silk_sub_sat32(int32_t a, int32_t b)
{
    uint32_t s = (uint32_t)a - (uint32_t)b;
    if ((s & 0x80000000u) == 0) return ((uint32_t)a & ((uint32_t)b ^ 0x80000000u) & 0x80000000u) ? INT32_MIN : (int32_t)s;
    return (((uint32_t)a ^ 0x80000000u) & (uint32_t)b & 0x80000000u) ? INT32_MAX : (int32_t)s;
}
CUO_FN int32_t silk_sat16(int32_t a) { return a > 32767 ? 32767 : (a < -32768 ? -32768 : a); }
CUO_FN int32_t silk_limit(int32_t a, int32_t l1, int32_t l2)
{
    return l1 > l2 ? (a > l1 ? l1 : (a < l2 ? l2 : a)) : (a > l2 ? l2 : (a < l1 ? l1 : a));
}
CUO_FN int32_t silk_lshift_sat32(int32_t a, int s)
{
    // silk_LIMIT_32 with lower limit below upper: a clamp
    const int32_t lo = (int32_t)(INT32_MIN >> s), hi = (int32_t)(INT32_MAX >> s);
    return silk_lshift(a < lo ? lo : (a > hi ? hi : a), s);
}
CUO_FN int32_t silk_clz32(int32_t a) { return a == 0 ? 32 : __builtin_clz((uint32_t)a); }
CUO_FN int32_t silk_clz16(int16_t a) { return silk_clz32((int32_t)(uint16_t)a) - 16; }
CUO_FN int32_t silk_ror32(int32_t a, int rot)
{
    uint32_t x = (uint32_t)a;
    uint32_t r = (uint32_t)rot, m = (uint32_t)-rot;
    if (rot == 0) return a;
    if (rot < 0) return (int32_t)((x << m) | (x >> (32 - m)));
    return (int32_t)((x << (32 - r)) | (x >> r));
}
CUO_FN void silk_clz_frac(int32_t in, int32_t *lz, int32_t *frac_Q7)
{
    int32_t l = silk_clz32(in);
    *lz = l;
    *frac_Q7 = silk_ror32(in, 24 - l) & 0x7f;
}
// silk_SQRT_APPROX (vendor/silk/Inlines.h)
CUO_FN int32_t silk_sqrt_approx(int32_t x)
{
    if (x <= 0) return 0;
    int32_t lz, frac_Q7;
    silk_clz_frac(x, &lz, &frac_Q7);
    int32_t y = (lz & 1) ? 32768 : 46214;
    y >>= (lz >> 1);
    y = silk_smlawb(y, y, silk_smulbb(213, frac_Q7));
    return y;
}
// silk_DIV32_varQ (vendor/silk/Inlines.h)
CUO_FN int32_t silk_div32_varq(int32_t a32, int32_t b32, int Qres)
{
    int a_headrm = silk_clz32(a32 < 0 ? -a32 : a32) - 1;
    int32_t a32_nrm = silk_lshift(a32, a_headrm);
    int b_headrm = silk_clz32(b32 < 0 ? -b32 : b32) - 1;
    int32_t b32_nrm = silk_lshift(b32, b_headrm);
    int32_t b32_inv = (INT32_MAX >> 2) / (b32_nrm >> 16);
    int32_t result = silk_smulwb(a32_nrm, b32_inv);
    a32_nrm = sub32_ovf(a32_nrm, (int32_t)((uint32_t)(int32_t)(silk_smull(b32_nrm, result) >> 32) << 3));
    result = silk_smlawb(result, a32_nrm, b32_inv);
    int lshift = 29 + a_headrm - b_headrm - Qres;
    if (lshift < 0) return silk_lshift_sat32(result, -lshift);
    if (lshift < 32) return result >> lshift;
    return 0;
}
// silk_INVERSE32_varQ (vendor/silk/Inlines.h)
CUO_FN int32_t silk_inverse32_varq(int32_t b32, int Qres)
{
    int b_headrm = silk_clz32(b32 < 0 ? -b32 : b32) - 1;
    int32_t b32_nrm = silk_lshift(b32, b_headrm);
    int32_t b32_inv = (INT32_MAX >> 2) / (b32_nrm >> 16);
    int32_t result = silk_lshift(b32_inv, 16);
    int32_t err_Q32 = silk_lshift((1 << 29) - silk_smulwb(b32_nrm, b32_inv), 3);
    result = silk_smlaww(result, err_Q32, b32_inv);
    int lshift = 61 - b_headrm - Qres;
    if (lshift <= 0) return silk_lshift_sat32(result, -lshift);
    if (lshift < 32) return result >> lshift;
    return 0;
}

}  // namespace CUO_NS
