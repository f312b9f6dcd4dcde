// demux.h -- host-side Ogg/Opus demuxer and chain builder.
//
// Default behaviour is the reference's cu_demux_mem / cu_build_chains2
// (cuopus_demux.h:79-269, 293-378): granule-exact packet index (TOC durations
// cross-checked against every page granule, RFC 7845 end-trim on the final
// data page), OpusHead channels/pre-skip, and the chain table that cuts the
// playable timeline into fixed windows decoded independently after a decoder
// preroll.  On streams the reference accepts the packet index and chain table
// are identical.  Beyond it (the reference's known limitation, README.md:
// 232-235, and its unimplemented CUOPUS_CHECK_CRC, BENCHMARKS.md:240):
//
//  * logical streams: pages of other serial numbers (multiplexed video or
//    metadata streams) are skipped; a new Opus BOS page after audio has
//    started opens a new link (chained stream, RFC 7845 section 3).  Each link
//    has its own pre-skip and timeline; chains never cross a link, so the
//    decoded 48 kHz PCM is the links' playable PCM back to back.
//  * reassembly follows the continuation flag and page sequence numbers: a
//    packet whose pages are not all present is dropped, never spliced.
//  * check_crc: page CRC-32 (polynomial 0x04C11DB7, RFC 3533) verified; a bad
//    page is an error under the strict policy and skipped (resync at the next
//    capture pattern) under the lenient one.
//  * lenient granule policy: mismatching page granules are counted as
//    warnings instead of failing the stream; the timeline re-bases at each
//    mismatch (so streams that start at a non-zero granule, or skip granules,
//    decode), end-trim is taken from the final page when it lies within that
//    page, and the playable length is the TOC total minus pre-skip and
//    end-trim.  The per-chain capacity checks still apply.
//
// Written as C++ over std::vector; errors return false with a message, never
// abort.
#pragma once
#include <stdint.h>
#include <string>
#include <memory>
#include <thread>
#include <utility>
#include <vector>

namespace cuo_host {

struct DemuxOpts {
    bool check_crc = false;   // CUOPUS_CHECK_CRC=1
    bool strict = true;       // CUOPUS_GRANULE=strict (default) | warn
    // Host threads for one stream (CUOPUS_HOST_THREADS; 0: the machine's
    // hardware threads, at most 16) and the stream size from which they are
    // used (CUOPUS_DEMUX_PARALLEL_MIN, bytes).  The result does not depend on
    // either.
    int threads = 0;
    size_t parallel_min_bytes = 4u << 20;
};

// Options from the environment (CUOPUS_CHECK_CRC, CUOPUS_GRANULE,
// CUOPUS_HOST_THREADS, CUOPUS_DEMUX_PARALLEL_MIN).
DemuxOpts demux_opts_from_env();

// Host threads the demuxer and the library's host loops use: opts.threads if
// set, else the machine's hardware threads, at most 16.
int host_threads(const DemuxOpts &opts);

// Run f(t) for t = 0 .. n-1 on n threads (t = 0 on the caller's thread).
template <class F>
void parallel_for(int n, F &&f);

// One logical Opus stream of a (possibly chained) physical stream.
struct Link {
    uint32_t first = 0, npkts = 0;  // audio packet range in Demuxed::off/len
    int channels = 0;
    int preskip = 0;
    int gain_q8 = 0;                // OpusHead output gain (RFC 7845 5.1), Q7.8 dB
    uint32_t serial = 0;
    int64_t toc_total = 0;          // sum of TOC durations of its packets
    int64_t last_granule = 0;
    int64_t end_trim = 0;
    int64_t play48 = 0;             // playable samples of this link
};

struct DemuxStats {
    uint32_t pages = 0;             // pages of the Opus stream(s) used
    uint32_t crc_bad = 0;           // pages failing the CRC (check_crc only)
    uint32_t foreign_pages = 0;     // pages of other logical streams
    uint32_t dropped_packets = 0;   // incomplete packets (missing pages, truncation)
    uint32_t granule_warnings = 0;  // mismatching granules accepted by the lenient policy
    uint64_t skipped_bytes = 0;     // bytes outside accepted pages (garbage, bad pages)
};

// std::vector storage whose resize leaves new elements uninitialised: the
// demuxer writes every element of its large tables once, so zero-filling
// them first would only add a pass over memory.
template <class T>
struct NoInitAlloc : std::allocator<T> {
    template <class U>
    struct rebind { using other = NoInitAlloc<U>; };
    NoInitAlloc() = default;
    template <class U>
    NoInitAlloc(const NoInitAlloc<U> &) noexcept {}
    template <class U>
    void construct(U *p) noexcept { ::new ((void *)p) U; }
    template <class U, class... A>
    void construct(U *p, A &&...a) { ::new ((void *)p) U(std::forward<A>(a)...); }
};
template <class T>
using RawVec = std::vector<T, NoInitAlloc<T>>;

struct Demuxed {
    RawVec<uint8_t> data;           // audio packets, contiguous
    RawVec<uint32_t> off, len;      // per audio packet
    RawVec<int64_t> sample_off;     // per packet, from the start of its link
    std::vector<Link> links;
    int64_t toc_total = 0;
    int channels = 0;
    int preskip = 0;                // of the first link
    int64_t last_granule = 0;       // of the first link
    int64_t pcm_samples_48k = 0;    // playable samples, all links
    uint32_t n_pages = 0;
    uint32_t granule_checked = 0, granule_match = 0;
    uint32_t stereo_packets = 0;    // packets whose TOC codes two channels
    DemuxStats stats;
};

// Whether decoding needs the stereo-capable decoder build: a stereo stream,
// or a mono stream that holds stereo packets (decoded to one channel).
inline bool needs_stereo_build(const Demuxed &d) { return d.channels == 2 || d.stereo_packets > 0; }

struct Chain {
    uint32_t first, npkts, trim, keep;
};

// Duration in 48 kHz samples of one Opus packet from its TOC (RFC 6716 3.1).
int pkt_samples48(const uint8_t *p, uint32_t len);

// Ogg page CRC-32: polynomial 0x04C11DB7, initial value 0, not reflected, no
// final xor (RFC 3533 section 6; ogg_index.py:22-35).
uint32_t ogg_crc(const uint8_t *p, size_t n, uint32_t crc = 0);

bool demux(const uint8_t *buf, size_t n, Demuxed *d, std::string *err, const DemuxOpts &opts = DemuxOpts());

bool build_chains(const Demuxed &d, double seconds, double preroll_ms, std::vector<Chain> *out,
                  std::string *err);

// f32 16 kHz output length for play48 playable 48 kHz samples (the CPU
// resampling path's 48-sample unflushed tail): max(0, (play48 - 48) / 3).
inline int64_t out16_len(int64_t play48)
{
    int64_t v = (play48 - 48) / 3;
    return v < 0 ? 0 : v;
}

template <class F>
void parallel_for(int n, F &&f)
{
    if (n <= 1) {
        if (n == 1) f(0);
        return;
    }
    std::vector<std::thread> th;
    th.reserve((size_t)n - 1);
    for (int t = 1; t < n; t++) th.emplace_back([&f, t] { f(t); });
    f(0);
    for (auto &x : th) x.join();
}

}  // namespace cuo_host
