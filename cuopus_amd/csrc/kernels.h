// kernels.h -- device kernels of the batched decoder.
//
// k_decode_chains: one wavefront decodes one chain.  The 64 lanes run the
//   serial Opus decoder in lockstep on identical data (wave-uniform state, so
//   scalar values live in SGPRs), then split the fused epilogue: each decoded
//   packet's samples that fall inside the chain's keep window are copied by
//   all 64 lanes in parallel straight into the compacted 48 kHz output, so
//   there is no separate compaction pass (cuopus_kernels.cuh:386-433 fused).
//   Each lane only reads back PCM it wrote itself, so no barrier is needed.
//   Chains are taken from a device work queue (an atomic counter) in the
//   persistent variant, or one per workgroup in the plain variant.
// k_resample_48to16: 49-tap Kaiser polyphase decimator, f32 out
//   (cuopus_kernels.cuh:441-456; gpu_resample.py:20-49).  A workgroup stages
//   its int16 input tile plus the 24-sample halo on each side in LDS once;
//   each thread computes 4 consecutive outputs from a register window.
#pragma once
#include "dec/opus_dec.h"

namespace CUO_NS {

constexpr int kMaxPktSamples = 5760;
// Resident wavefronts of the persistent decoder: four per SIMD (10 KB of
// LDS and <= 128 VGPRs per wavefront), 16 per CU, 256 CUs.
constexpr uint32_t kDefaultWaves = 4096;
constexpr int kRsTaps = 49;
constexpr int kRsPad = 24;

struct ChainDesc {
    uint32_t first;   // first packet (global index)
    uint32_t npkts;
    uint32_t trim;    // samples of the chain's output dropped before the keep window
    uint32_t keep;    // samples kept
    uint64_t dst;     // offset of the keep window in the compacted output
    uint32_t rec_end; // per-packet records (gate): only packets below this index
    int16_t gain;     // output gain, Q8 dB (OpusHead; libopus OPUS_SET_GAIN semantics)
    uint8_t channels; // output channels of the chain's decoder (1, or 2 in the stereo build)
    uint8_t pad;
};

struct ChainResult {
    int32_t stop;      // 0 or the first negative decoder return
    uint32_t samples;  // samples produced by the chain's packets
    uint32_t t0, t1;   // constant 100 MHz clock (low 32 bits) when the chain started and ended
};

// Frame sink of a chain: each decoded frame's samples that fall inside the
// chain's keep window are copied by the 64 lanes in parallel from the LDS
// frame buffer straight into the compacted output.
struct ChainOut {
    int16_t *buf;
    int16_t *out;
    uint64_t dst, lo, hi, pos;  // in sample frames (one sample per output channel)
    int cc;                     // output channels, interleaved in buf and out
    __device__ void emit(int n)
    {
        CUO_PROF_SCOPE(kProfEmit);
        const uint64_t C = kMaxCh == 1 ? 1 : (uint64_t)cc;
        uint64_t a = pos > lo ? pos : lo;
        uint64_t b = pos + (uint64_t)n < hi ? pos + (uint64_t)n : hi;
        for (uint64_t e = C * a + (uint64_t)lane_id(); e < C * b; e += 64) out[C * (dst - lo) + e] = buf[e - C * pos];
        pos += (uint64_t)n;
    }
};

// Decoder memory of one wavefront.  The chain's state and per-call scratch
// (WaveState, under 10 KB) sit in LDS, so four wavefronts fit per SIMD; the
// CELT history and synthesis spectrum, the output, SILK, transition and
// redundancy frames, the concealment and comfort-noise buffers (DecHeap,
// 37 KB) sit in a
// per-wavefront HBM slot, where only lane-parallel loops and the rare
// concealment touch them.  Chains restart from a fresh decoder, so nothing is
// carried between chains.
// The chain loop's own state.  It lives in LDS and is read where it is used,
// so the inlined CELT frame decoder has every scalar register to itself.
struct ChainWalk {
    const uint8_t *data;
    const uint32_t *off, *len;
    const uint8_t *lost;
    int32_t *pkt_nsamp;
    uint32_t *pkt_rng;
    uint32_t first, npkts, rec_end, k;
};

struct WaveState {
    OpusDecState st;
    DecScratch sc;
    ChainOut out;
    ChainWalk walk;
};
static_assert(sizeof(WaveState) <= kWaveLdsBytes, "wavefront decoder memory exceeds its LDS block");
// Resident wavefronts per SIMD: four for the mono build (10 KB of LDS, <= 128
// VGPRs), three for the stereo one (13 KB, <= 168 VGPRs).
constexpr int kWavesPerSimd = kMaxCh == 1 ? 4 : 3;
static_assert(kWaveLdsBytes * 4 * kWavesPerSimd <= 160 * 1024, "LDS per wavefront exceeds the CU's share");

struct DecodeArgs {
    const uint8_t *data;
    const uint32_t *off;
    const uint32_t *len;
    const uint8_t *lost;        // nullable: 1 = packet lost
    const ChainDesc *chains;
    const uint32_t *order;      // nullable: work order (chain indices)
    uint32_t n_chains;
    DecHeap *heap;              // one per workgroup of the launch
    int16_t *out;               // compacted 48 kHz PCM
    ChainResult *res;
    int32_t *pkt_nsamp;         // nullable, per packet
    uint32_t *pkt_rng;          // nullable, per packet
    uint32_t *queue;            // work-queue counter (persistent variant)
};

__device__ __forceinline__ const uint8_t *uniform_ptr(const uint8_t *p)
{
    const uint64_t v = (uint64_t)p;
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)v);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(v >> 32));
    return (const uint8_t *)(((uint64_t)hi << 32) | lo);
}

// Every packet the plain CELT path does not take (lost packets, SILK and
// hybrid packets, mode transitions, stereo): the general packet decoder, out
// of line, so the chain loop below holds the inlined CELT frame decoder
// without the general path's registers.  data may be nullptr (concealment).
__device__ __noinline__ int decode_packet_general(WaveState *ws, const uint8_t *data, int len, int frame_size,
                                                  int fec)
{
    ws = lds_uniform(ws);
    return opus_decode_packet(&ws->st, uniform_ptr(data), uniform(len), ws->out, uniform(frame_size), uniform(fec),
                              &ws->sc);
}

// Decode one chain on the calling wavefront.  Plain CELT packets (the whole
// of a CELT-only stream without loss) decode in this loop with the CELT frame
// decoder inlined, so no call is made per frame; every other packet goes
// through decode_packet_general.  Kept out of line: the persistent loop
// around it then stays a small loop whose every iteration runs with all 64
// lanes together.
__device__ __noinline__ void decode_one_chain(const DecodeArgs &a, uint32_t c, WaveState *ws, DecHeap *heap)
{
    prof_reset();
    const uint32_t t_start = (uint32_t)__builtin_amdgcn_s_memrealtime();
    {
    CUO_PROF_SCOPE(kProfChain);
    const ChainDesc ch = a.chains[c];
    ws = lds_uniform(ws);
    OpusDecState *st = lds_uniform(&ws->st);
    DecScratch *sc = lds_uniform(&ws->sc);
    ChainOut *out = lds_uniform(&ws->out);
    ChainWalk *w = lds_uniform(&ws->walk);
    dec_bind(st, sc, heap);
    opus_dec_init(st, kMaxCh == 1 ? 1 : ch.channels);
    st->decode_gain = ch.gain;
    out->buf = sc->frame_out;
    out->out = a.out;
    out->dst = ch.dst;
    out->lo = ch.trim;
    out->hi = (uint64_t)ch.trim + ch.keep;
    out->pos = 0;
    out->cc = st->channels;
    w->data = a.data;
    w->off = a.off;
    w->len = a.len;
    w->lost = a.lost;
    w->pkt_nsamp = a.pkt_nsamp;
    w->pkt_rng = a.pkt_rng;
    w->first = ch.first;
    w->npkts = ch.npkts;
    w->rec_end = ch.rec_end;
    w->k = 0;
    int32_t stop = 0;
    while (w->k < w->npkts) {
        const uint32_t k = w->k;
        const uint32_t p = w->first + k;
        const uint8_t *data = w->data;
        const uint32_t *off = w->off, *len = w->len;
        const uint8_t *lost = w->lost;
        const uint8_t *pk = data + off[p];
        int r;
        if (lost && lost[p]) {
            // loss rule: FEC from the next packet of the chain when it carries
            // LBRR, else concealment of this packet's duration
            int dur = packet_nb_samples(pk, (int)len[p]);
            if (dur <= 0) { stop = dur ? dur : OPUS_INVALID_PACKET; break; }
            bool fec = false;
            if (k + 1 < w->npkts && !lost[p + 1])
                fec = packet_has_lbrr(data + off[p + 1], (int)len[p + 1]) > 0;
            if (fec) r = decode_packet_general(ws, data + off[p + 1], (int)len[p + 1], dur, 1);
            else r = decode_packet_general(ws, nullptr, 0, dur, 0);
        } else {
            r = celt_plain_packet(st, pk, (int)len[p], *out, sc);
            if (r == kNotPlain) r = decode_packet_general(ws, w->data + w->off[p], (int)w->len[p], kMaxPktSamples, 0);
        }
        if (r < 0) { stop = r; break; }
        // a packet held by several chains (preroll overlap) is recorded by the
        // last chain that holds it, which owns packets below the next chain's
        // first packet
        const uint32_t q = w->first + w->k;
        if (w->pkt_nsamp && lane_id() == 0 && q < w->rec_end) {
            w->pkt_nsamp[q] = r;
            w->pkt_rng[q] = st->rangeFinal;
        }
        w->k = w->k + 1;
    }
    const uint32_t t_end = (uint32_t)__builtin_amdgcn_s_memrealtime();
    if (lane_id() == 0) {
        a.res[c].stop = stop;
        a.res[c].samples = (uint32_t)out->pos;
        a.res[c].t0 = t_start;
        a.res[c].t1 = t_end;
    }
    }
    prof_flush();
}

// The launch's DecodeArgs, read in place from the kernel-argument segment (a
// decode kernel's only argument).  Passing the by-value argument on by
// reference would copy it to per-lane scratch memory, and a kernel that keeps
// values live across the out-of-line decoder needs more registers than three
// wavefronts per SIMD allow; the kernels below only forward this pointer.
__device__ __forceinline__ const DecodeArgs *launch_args()
{
    return (const DecodeArgs *)__builtin_amdgcn_kernarg_segment_ptr();
}

// Plain variant: workgroup g decodes chain order[g].
__device__ __noinline__ void plain_worker(const DecodeArgs *ap, uint32_t g, int lane)
{
    if (g >= ap->n_chains) return;
    const uint32_t c = ap->order ? ap->order[g] : g;
    decode_one_chain(*ap, c, (WaveState *)g_wave_lds, ap->heap + g);
}
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(kWavesPerSimd))) void k_decode_chains(DecodeArgs)
{
    plain_worker(launch_args(), blockIdx.x, (int)threadIdx.x);
}

// Persistent variant: the grid's wavefronts pull chains from an atomic queue
// until it is drained (every wavefront exits once the counter passes n_chains).
__device__ __noinline__ void persistent_worker(const DecodeArgs *ap, DecHeap *heap, int lane)
{
    WaveState *ws = (WaveState *)g_wave_lds;
    while (true) {
        // The fetched index is made wave-uniform with readfirstlane (not a
        // lane shuffle) so the loop exit is a scalar branch: with a shuffle the
        // compiler treats the exit as divergent and may let lanes run different
        // iterations, which a wavefront-per-chain decoder cannot allow.
        uint32_t w = 0;
        if (lane == 0) w = atomicAdd(ap->queue, 1u);
        w = (uint32_t)__builtin_amdgcn_readfirstlane((int)w);
        if (w >= ap->n_chains) break;
        const uint32_t c = ap->order ? ap->order[w] : w;
        decode_one_chain(*ap, c, ws, heap);
    }
}
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(kWavesPerSimd))) void k_decode_persistent(DecodeArgs)
{
    const DecodeArgs *ap = launch_args();
    persistent_worker(ap, ap->heap + blockIdx.x, (int)threadIdx.x);
}

// Zero a work-queue counter (a kernel rather than a memset so the sequence can
// be captured into a HIP graph and replayed).
__global__ void k_reset_counter(uint32_t *q) { if (threadIdx.x == 0) q[0] = 0; }

struct RsTaps {
    float t[kRsTaps];
};

struct RsFile {
    uint64_t in_off;   // offset of the file's compacted 48 kHz PCM
    uint64_t in_len;   // playable samples of the file
    uint64_t out_len;  // outputs of the file
    float *out;        // output pointer (device)
    uint64_t tile0;    // first global tile index of this file
};

// Resampler launch shapes (threads per workgroup x consecutive outputs per
// thread); the autotuner picks one per device.  Outputs per thread stay a
// multiple of 4 so each thread's window starts 16-byte aligned in LDS.
constexpr int kRsVariants = 5;
constexpr int kRsDefaultVariant = 0;
constexpr int kRsShape[kRsVariants][2] = {{256, 4}, {256, 8}, {128, 8}, {512, 4}, {128, 4}};
__host__ __device__ constexpr int rs_tile(int v) { return kRsShape[v][0] * kRsShape[v][1]; }

// Batched resampler: workgroup b handles tile b of the concatenated per-file
// tile list; out[i] = sum_t taps[t] * x[3 i - 24 + t] / 32768 with zeros
// outside [0, in_len), each output accumulated in tap order with fused
// multiply-adds.  The tile's input (3 * Threads * PerThread samples plus the
// 24-sample halo on each side) is staged in LDS as f32, with 16-byte loads
// where the tile lies inside the file and its start is 16-byte aligned;
// thread t then reads its window as float4s (lanes 12 * PerThread bytes
// apart) and produces PerThread consecutive outputs, stored as float4s when
// aligned.
template <int Threads, int PerThread>
__global__ __launch_bounds__(Threads) void k_resample_48to16(const int16_t *__restrict__ in,
                                                             const RsFile *__restrict__ files, int n_files,
                                                             RsTaps taps)
{
    constexpr int Tile = Threads * PerThread;
    constexpr int In = 3 * Tile + 2 * kRsPad;                      // staged input samples
    constexpr int Win = (3 * (PerThread - 1) + kRsTaps + 3) & ~3;  // one thread's window
    static_assert(PerThread % 4 == 0, "windows must start 16-byte aligned");
    static_assert(In % 8 == 0, "staging works in 16-byte chunks of 8 samples");
    static_assert(3 * PerThread * (Threads - 1) + Win <= In, "the last window stays inside the tile");
    __shared__ __align__(16) float xs[In];
    const uint64_t b = blockIdx.x;
    // find the file of this tile (files are few per launch; binary search)
    int lo = 0, hi = n_files - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (files[mid].tile0 <= b) lo = mid; else hi = mid - 1;
    }
    const RsFile f = files[lo];
    const uint64_t o0 = (b - f.tile0) * Tile;
    const int64_t base = (int64_t)(3 * o0) - kRsPad;
    const int16_t *x = in + f.in_off;
    const int tid = (int)threadIdx.x;
    const bool inside = base >= 0 && base + In <= (int64_t)f.in_len;
    if (inside && ((uintptr_t)(x + base) & 15) == 0) {
        const int4 *src = (const int4 *)(x + base);
        for (int c = tid; c < In / 8; c += Threads) {
            const int4 v = src[c];
            const int32_t w4[4] = {v.x, v.y, v.z, v.w};
            float y[8];
#pragma unroll
            for (int k = 0; k < 4; k++) {
                y[2 * k] = (float)(int16_t)(w4[k] & 0xFFFF) * (1.0f / 32768.0f);
                y[2 * k + 1] = (float)(int16_t)((uint32_t)w4[k] >> 16) * (1.0f / 32768.0f);
            }
            float4 *d = (float4 *)&xs[8 * c];
            d[0] = make_float4(y[0], y[1], y[2], y[3]);
            d[1] = make_float4(y[4], y[5], y[6], y[7]);
        }
    } else {
        for (int i = tid; i < In; i += Threads) {
            const int64_t gi = base + i;
            xs[i] = (gi >= 0 && gi < (int64_t)f.in_len) ? (float)x[gi] * (1.0f / 32768.0f) : 0.f;
        }
    }
    __syncthreads();
    const uint64_t o = o0 + (uint64_t)(PerThread * tid);
    if (o >= f.out_len) return;
    float w[Win];
    const float4 *wv = (const float4 *)&xs[3 * PerThread * tid];
#pragma unroll
    for (int m = 0; m < Win / 4; m++) {
        const float4 v = wv[m];
        w[4 * m] = v.x;
        w[4 * m + 1] = v.y;
        w[4 * m + 2] = v.z;
        w[4 * m + 3] = v.w;
    }
    float acc[PerThread];
#pragma unroll
    for (int j = 0; j < PerThread; j++) acc[j] = 0.f;
#pragma unroll
    for (int t = 0; t < kRsTaps; t++)
#pragma unroll
        for (int j = 0; j < PerThread; j++) acc[j] = __builtin_fmaf(taps.t[t], w[3 * j + t], acc[j]);
    float *dst = f.out + o;
    if (o + PerThread <= f.out_len && ((uintptr_t)dst & 15) == 0) {
#pragma unroll
        for (int j = 0; j < PerThread; j += 4)
            *(float4 *)(dst + j) = make_float4(acc[j], acc[j + 1], acc[j + 2], acc[j + 3]);
    } else {
#pragma unroll
        for (int j = 0; j < PerThread; j++)
            if (o + j < f.out_len) dst[j] = acc[j];
    }
}

}  // namespace CUO_NS
