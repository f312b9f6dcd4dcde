// demux.h -- host-side Ogg/Opus demuxer and chain builder.
//
// Same contract as the reference's cu_demux_mem / cu_build_chains2
// (cuopus_demux.h:79-269, 293-378): granule-exact packet index (TOC durations
// cross-checked against every page granule, RFC 7845 end-trim on the final
// data page), OpusHead channels/pre-skip, and the chain table that cuts the
// playable timeline into fixed windows decoded independently after a decoder
// preroll.  Written as C++ over std::vector; errors return false with a
// message, never abort.
#pragma once
#include <stdint.h>
#include <string>
#include <vector>

namespace cuo_host {

struct Demuxed {
    std::vector<uint8_t> data;      // audio packets, contiguous
    std::vector<uint32_t> off, len; // per audio packet
    std::vector<int64_t> sample_off;
    int64_t toc_total = 0;
    int channels = 0;
    int preskip = 0;
    int64_t last_granule = 0;
    int64_t pcm_samples_48k = 0;    // playable samples
    uint32_t n_pages = 0;
    uint32_t granule_checked = 0, granule_match = 0;
};

struct Chain {
    uint32_t first, npkts, trim, keep;
};

// Duration in 48 kHz samples of one Opus packet from its TOC (RFC 6716 3.1).
int pkt_samples48(const uint8_t *p, uint32_t len);

bool demux(const uint8_t *buf, size_t n, Demuxed *d, std::string *err);

bool build_chains(const Demuxed &d, double seconds, double preroll_ms, std::vector<Chain> *out,
                  std::string *err);

// f32 16 kHz output length for play48 playable 48 kHz samples (the CPU
// resampling path's 48-sample unflushed tail): max(0, (play48 - 48) / 3).
inline int64_t out16_len(int64_t play48)
{
    int64_t v = (play48 - 48) / 3;
    return v < 0 ? 0 : v;
}

}  // namespace cuo_host
