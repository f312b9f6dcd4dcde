// demux.cpp -- see demux.h.  Behaviour follows cuopus_demux.h:79-378.
#include "demux.h"

#include <string.h>

#include <algorithm>

namespace cuo_host {

static const int kQuarterMs[32] = {40, 80, 160, 240, 40, 80, 160, 240, 40, 80, 160, 240, 40, 80, 40, 80,
                                   10, 20, 40, 80, 10, 20, 40, 80, 10, 20, 40, 80, 10, 20, 40, 80};

int pkt_samples48(const uint8_t *p, uint32_t len)
{
    if (!len) return 0;
    int config = p[0] >> 3, code = p[0] & 3, nf;
    if (code == 0) nf = 1;
    else if (code < 3) nf = 2;
    else {
        if (len < 2) return 0;
        nf = p[1] & 0x3F;
    }
    return kQuarterMs[config] * 12 * nf;
}

namespace {
struct Span {
    uint64_t off;
    uint32_t len;
    uint32_t page;
    bool done;  // ends in a lacing value < 255 (packet completes here)
};
}  // namespace

bool demux(const uint8_t *buf, size_t n, Demuxed *d, std::string *err)
{
    *d = Demuxed();
    if (!buf || n == 0) { *err = "empty stream"; return false; }
    std::vector<Span> spans;
    std::vector<int64_t> page_granule;
    size_t pos = 0;
    static const uint8_t kCapture[4] = {'O', 'g', 'g', 'S'};
    while (pos < n) {
        const uint8_t *hit = std::search(buf + pos, buf + n, kCapture, kCapture + 4);
        if (hit == buf + n) break;
        size_t idx = (size_t)(hit - buf);
        if (idx + 27 > n) break;
        int64_t granule;
        memcpy(&granule, buf + idx + 6, 8);
        int nsegs = buf[idx + 26];
        const uint8_t *seg = buf + idx + 27;
        size_t body = idx + 27 + (size_t)nsegs;
        if (body > n) break;
        uint32_t body_len = 0;
        for (int i = 0; i < nsegs; i++) body_len += seg[i];
        if (body + body_len > n) break;
        uint32_t pg = (uint32_t)page_granule.size();
        page_granule.push_back(granule);
        uint64_t run_start = body;
        uint32_t run_len = 0;
        int last = 0;
        for (int i = 0; i < nsegs; i++) {
            run_len += seg[i];
            last = seg[i];
            if (seg[i] < 255) {
                spans.push_back({run_start, run_len, pg, true});
                run_start += run_len;
                run_len = 0;
            }
        }
        if (run_len > 0 || (nsegs > 0 && last == 255)) spans.push_back({run_start, run_len, pg, false});
        pos = body + body_len;
    }
    if (page_granule.empty()) { *err = "no OGG pages found"; return false; }
    d->n_pages = (uint32_t)page_granule.size();

    // reassemble packets (a trailing partial packet is dropped)
    std::vector<uint64_t> pk_begin_span;
    std::vector<uint32_t> pk_len, pk_page;
    {
        uint64_t acc = 0;
        size_t first_span = 0;
        for (size_t i = 0; i < spans.size(); i++) {
            acc += spans[i].len;
            if (spans[i].done) {
                pk_begin_span.push_back(first_span);
                pk_len.push_back((uint32_t)acc);
                pk_page.push_back(spans[i].page);
                acc = 0;
                first_span = i + 1;
            }
        }
    }
    size_t n_packets = pk_len.size();
    if (n_packets < 2) { *err = "not an OGG Opus stream (fewer than 2 packets)"; return false; }
    // the OpusHead header fits in its first span for any valid stream
    const Span &h0 = spans[pk_begin_span[0]];
    if (pk_len[0] < 19 || h0.len < 19 || memcmp(buf + h0.off, "OpusHead", 8) != 0) {
        *err = "not an OGG Opus stream (no OpusHead)";
        return false;
    }
    d->channels = buf[h0.off + 9];
    d->preskip = buf[h0.off + 10] | (buf[h0.off + 11] << 8);

    size_t n_audio = n_packets - 2;
    uint64_t audio_bytes = 0;
    for (size_t i = 2; i < n_packets; i++) audio_bytes += pk_len[i];
    if (audio_bytes >= (1ull << 32)) { *err = "stream too large (>= 4 GiB of packets)"; return false; }
    d->data.resize(audio_bytes ? audio_bytes : 1);
    d->off.resize(n_audio);
    d->len.resize(n_audio);
    d->sample_off.resize(n_audio);
    std::vector<uint32_t> pkt_page(n_audio);
    uint64_t w = 0;
    for (size_t k = 0; k < n_audio; k++) {
        size_t p = k + 2;
        d->off[k] = (uint32_t)w;
        d->len[k] = pk_len[p];
        pkt_page[k] = pk_page[p];
        uint64_t remaining = pk_len[p];
        for (size_t s = pk_begin_span[p]; remaining > 0; s++) {
            memcpy(d->data.data() + w, buf + spans[s].off, spans[s].len);
            w += spans[s].len;
            remaining -= spans[s].len;
        }
    }

    // per-packet sample offsets and the per-page granule cross-check
    int64_t cum = 0;
    for (size_t k = 0; k < n_audio; k++) {
        d->sample_off[k] = cum;
        cum += pkt_samples48(d->data.data() + d->off[k], d->len[k]);
    }
    d->toc_total = cum;
    std::vector<int64_t> page_end(page_granule.size(), -1);
    int64_t final_data_page = -1;
    for (size_t k = 0; k < n_audio; k++) {
        uint32_t pg = pkt_page[k];
        page_end[pg] = d->sample_off[k] + pkt_samples48(d->data.data() + d->off[k], d->len[k]);
        final_data_page = std::max<int64_t>(final_data_page, pg);
    }
    int last_dur = n_audio ? pkt_samples48(d->data.data() + d->off[n_audio - 1], d->len[n_audio - 1]) : 0;
    uint32_t checked = 0, match = 0;
    int64_t last_granule = 0;
    for (size_t pg = 0; pg < page_granule.size(); pg++) {
        int64_t g = page_granule[pg];
        if (g != -1 && g > last_granule) last_granule = g;
        if (g == -1 || page_end[pg] < 0) continue;
        checked++;
        if (g == page_end[pg]) match++;
        else if ((int64_t)pg == final_data_page && page_end[pg] - g >= 0 && page_end[pg] - g < last_dur) match++;
    }
    d->granule_checked = checked;
    d->granule_match = match;
    d->last_granule = last_granule;
    d->pcm_samples_48k = last_granule - d->preskip;
    return true;
}

bool build_chains(const Demuxed &d, double seconds, double preroll_ms, std::vector<Chain> *out, std::string *err)
{
    out->clear();
    const int64_t step = (int64_t)(seconds * 48000);
    const int64_t preroll = (int64_t)(preroll_ms * 48);
    const int64_t total_play = d.pcm_samples_48k;
    const size_t n_pk = d.off.size();
    if (total_play <= 0) { *err = "no playable samples"; return false; }
    if (step <= 0) { *err = "bad segment length"; return false; }
    if (n_pk == 0) { *err = "no audio packets"; return false; }
    std::vector<int64_t> cum(n_pk + 1, 0);
    for (size_t k = 0; k < n_pk; k++) cum[k + 1] = cum[k] + pkt_samples48(d.data.data() + d.off[k], d.len[k]);
    const std::vector<int64_t> &so = d.sample_off;
    int64_t kept = 0;
    for (int64_t k = 0; k * step < total_play; k++) {
        int64_t play_start = k * step;
        int64_t play_end = std::min((k + 1) * step, total_play);
        int64_t toc_start = play_start + d.preskip;
        int64_t toc_end = play_end + d.preskip;
        int64_t start_target = std::max<int64_t>(0, toc_start - preroll);
        // first packet whose span reaches past the preroll start
        int64_t first = (int64_t)(std::upper_bound(so.begin(), so.end(), start_target) - so.begin()) - 1;
        if (first < 0) first = 0;
        if (cum[first + 1] <= start_target) first++;
        // last packet starting before the window end
        int64_t last = (int64_t)(std::lower_bound(so.begin(), so.end(), toc_end) - so.begin()) - 1;
        if (last < first) last = first;
        if (first >= (int64_t)n_pk) { *err = "chain starts past the last packet"; return false; }
        int64_t trim = toc_start - so[first];
        int64_t keep = play_end - play_start;
        if (trim < 0) { *err = "negative trim"; return false; }
        int64_t capn = cum[last + 1] - cum[first];
        if (capn < trim + keep) {
            *err = "chain " + std::to_string(k) + ": capacity " + std::to_string(capn) + " < trim+keep";
            return false;
        }
        out->push_back({(uint32_t)first, (uint32_t)(last - first + 1), (uint32_t)trim, (uint32_t)keep});
        kept += keep;
    }
    if (kept != total_play) { *err = "internal: sum(keep) != playable"; return false; }
    return true;
}

}  // namespace cuo_host
