// demux.cpp -- see demux.h.  Default behaviour follows cuopus_demux.h:79-378.
#include "demux.h"

#include <stdlib.h>
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
struct CrcTable {
    uint32_t t[256];
    CrcTable()
    {
        for (uint32_t i = 0; i < 256; i++) {
            uint32_t r = i << 24;
            for (int k = 0; k < 8; k++) r = (r & 0x80000000u) ? (r << 1) ^ 0x04C11DB7u : (r << 1);
            t[i] = r;
        }
    }
};
const CrcTable kCrc;
}  // namespace

uint32_t ogg_crc(const uint8_t *p, size_t n, uint32_t crc)
{
    for (size_t i = 0; i < n; i++) crc = (crc << 8) ^ kCrc.t[((crc >> 24) ^ p[i]) & 0xFF];
    return crc;
}

DemuxOpts demux_opts_from_env()
{
    DemuxOpts o;
    const char *c = getenv("CUOPUS_CHECK_CRC");
    o.check_crc = c && *c && strcmp(c, "0") != 0;
    const char *g = getenv("CUOPUS_GRANULE");
    o.strict = !(g && (!strcmp(g, "warn") || !strcmp(g, "lenient")));
    return o;
}

namespace {
// A lacing run of one page: consecutive segments up to and including the
// first one shorter than 255 bytes (a packet ends there), or the page's
// trailing run of 255-byte segments (the packet continues on the next page).
struct Span {
    uint64_t off;
    uint32_t len;
    bool done;
};

struct Page {
    int64_t granule;
    uint32_t serial, seq;
    uint8_t flags;       // 1 continued, 2 BOS, 4 EOS
};

// A packet assembled from spans of consecutive pages of one stream: its first
// piece inline, further pieces (a packet crossing pages) in the link's
// `more` list at [more0, more0 + nparts - 1).
struct Packet {
    uint64_t off0 = 0;
    uint32_t len0 = 0;
    uint32_t more0 = 0, nparts = 0;
    uint32_t len = 0;
    uint32_t page = 0;  // page on which it completes
};

struct LinkBuild {
    uint32_t serial = 0;
    bool eos = false;
    bool have_seq = false;
    uint32_t last_seq = 0;
    std::vector<Packet> packets;  // header packets included
    std::vector<std::pair<uint64_t, uint32_t>> more;  // (file offset, length) of continuation pieces
    Packet partial;
    bool has_partial = false;
    std::vector<uint32_t> pages;  // page indices of this link
};
}  // namespace

bool demux(const uint8_t *buf, size_t n, Demuxed *d, std::string *err, const DemuxOpts &opts)
{
    *d = Demuxed();
    if (!buf || n == 0) { *err = "empty stream"; return false; }
    DemuxStats &st = d->stats;

    // ---- pages, and the packets of the Opus links, in one pass
    std::vector<Page> pages;
    pages.reserve(n / 4000 + 16);
    std::vector<LinkBuild> links;
    Span spans[255];
    size_t done_spans = 0;  // spans ending a packet on any accepted page (error messages only)
    size_t pos = 0;
    static const uint8_t kCapture[4] = {'O', 'g', 'g', 'S'};
    while (pos < n) {
        const uint8_t *hit = pos + 4 <= n && memcmp(buf + pos, kCapture, 4) == 0
                                 ? buf + pos
                                 : std::search(buf + pos, buf + n, kCapture, kCapture + 4);
        if (hit == buf + n) { st.skipped_bytes += n - pos; break; }
        size_t idx = (size_t)(hit - buf);
        st.skipped_bytes += idx - pos;
        if (idx + 27 > n) { st.skipped_bytes += n - idx; break; }
        int nsegs = buf[idx + 26];
        const uint8_t *seg = buf + idx + 27;
        size_t body = idx + 27 + (size_t)nsegs;
        if (body > n) { st.skipped_bytes += n - idx; break; }
        uint32_t body_len = 0;
        for (int i = 0; i < nsegs; i++) body_len += seg[i];
        if (body + body_len > n) { st.skipped_bytes += n - idx; break; }
        Page pg;
        memcpy(&pg.granule, buf + idx + 6, 8);
        memcpy(&pg.serial, buf + idx + 14, 4);
        memcpy(&pg.seq, buf + idx + 18, 4);
        pg.flags = buf[idx + 5];
        if (opts.check_crc) {
            uint32_t want;
            memcpy(&want, buf + idx + 22, 4);
            static const uint8_t kZero[4] = {0, 0, 0, 0};
            uint32_t c = ogg_crc(buf + idx, 22);
            c = ogg_crc(kZero, 4, c);
            c = ogg_crc(buf + idx + 26, 1 + (size_t)nsegs + body_len, c);
            if (c != want || buf[idx + 4] != 0) {
                st.crc_bad++;
                if (opts.strict) {
                    *err = "page at byte " + std::to_string(idx) + ": CRC mismatch";
                    return false;
                }
                st.skipped_bytes += 4;
                pos = idx + 4;  // resync at the next capture pattern
                continue;
            }
        }
        // this page's spans
        int nspan = 0;
        {
            uint64_t run_start = body;
            uint32_t run_len = 0;
            int last = 0;
            for (int i = 0; i < nsegs; i++) {
                run_len += seg[i];
                last = seg[i];
                if (seg[i] < 255) {
                    spans[nspan++] = {run_start, run_len, true};
                    run_start += run_len;
                    run_len = 0;
                }
            }
            if (run_len > 0 || (nsegs > 0 && last == 255)) spans[nspan++] = {run_start, run_len, false};
        }
        const uint32_t pi = (uint32_t)pages.size();
        pages.push_back(pg);
        pos = body + body_len;
        for (int i = 0; i < nspan; i++) done_spans += spans[i].done;

        // ---- link of this page (first Opus stream, a new chained link, or foreign)
        const bool head = !(pg.flags & 1) && nspan > 0 && spans[0].len >= 19 &&
                          memcmp(buf + spans[0].off, "OpusHead", 8) == 0;
        LinkBuild *L = links.empty() ? nullptr : &links.back();
        if (head && (!L || ((pg.flags & 2) && (L->eos || L->packets.size() > 2)))) {
            if (L && L->has_partial) st.dropped_packets++;
            links.emplace_back();
            L = &links.back();
            L->serial = pg.serial;
            if (links.size() == 1) L->packets.reserve(n / 32 + 16);  // about one packet per 32 bytes
        } else if (!L || pg.serial != L->serial) {
            st.foreign_pages++;
            continue;
        }
        L->pages.push_back(pi);
        st.pages++;
        if (L->has_partial && (!(pg.flags & 1) || (L->have_seq && pg.seq != L->last_seq + 1))) {
            L->has_partial = false;  // its continuation is missing
            if (L->partial.nparts > 1) L->more.resize(L->partial.more0);
            L->partial = Packet();
            st.dropped_packets++;
        }
        L->have_seq = true;
        L->last_seq = pg.seq;
        int s = 0;
        if ((pg.flags & 1) && !L->has_partial && s < nspan) {
            // continuation of a packet whose start is missing
            while (s < nspan && !spans[s].done) s++;
            if (s < nspan) { s++; st.dropped_packets++; }
        }
        for (; s < nspan; s++) {
            Packet &p = L->partial;
            if (!L->has_partial) {
                p.off0 = spans[s].off;
                p.len0 = spans[s].len;
                p.nparts = 1;
            } else {
                if (p.nparts == 1) p.more0 = (uint32_t)L->more.size();
                L->more.push_back({spans[s].off, spans[s].len});
                p.nparts++;
            }
            p.len += spans[s].len;
            if (spans[s].done) {
                p.page = pi;
                L->packets.push_back(p);
                L->partial = Packet();
                L->has_partial = false;
            } else {
                L->has_partial = true;
            }
        }
        if (pg.flags & 4) L->eos = true;
    }
    if (pages.empty()) { *err = "no OGG pages found"; return false; }
    for (auto &L : links)
        if (L.has_partial) st.dropped_packets++;  // trailing partial packet
    d->n_pages = (uint32_t)pages.size();
    if (links.empty()) {
        // mirror the reference's messages for streams without an Opus head
        *err = done_spans < 2 ? "not an OGG Opus stream (fewer than 2 packets)" : "not an OGG Opus stream (no OpusHead)";
        return false;
    }
    if (links[0].packets.size() < 2) { *err = "not an OGG Opus stream (fewer than 2 packets)"; return false; }

    // ---- packet table
    uint64_t audio_bytes = 0;
    size_t n_audio = 0;
    for (auto &L : links)
        for (size_t i = 2; i < L.packets.size(); i++) {
            audio_bytes += L.packets[i].len;
            n_audio++;
        }
    if (audio_bytes >= (1ull << 32)) { *err = "stream too large (>= 4 GiB of packets)"; return false; }
    d->data.resize(audio_bytes ? audio_bytes : 1);
    d->off.resize(n_audio);
    d->len.resize(n_audio);
    d->sample_off.resize(n_audio);
    uint64_t w = 0;
    size_t k = 0;
    for (size_t li = 0; li < links.size(); li++) {
        LinkBuild &LB = links[li];
        Link lk;
        lk.serial = LB.serial;
        const Packet &h = LB.packets[0];
        if (LB.packets.size() < 2 || h.nparts == 0 || h.len0 < 19) {
            // a link cut short before its headers completed carries no audio
            if (li == 0) { *err = "not an OGG Opus stream (no OpusHead)"; return false; }
            st.dropped_packets += (uint32_t)LB.packets.size();
            continue;
        }
        const uint8_t *hp = buf + h.off0;
        lk.channels = hp[9];
        lk.preskip = hp[10] | (hp[11] << 8);
        lk.gain_q8 = (int16_t)(uint16_t)(hp[16] | (hp[17] << 8));
        lk.first = (uint32_t)k;
        // packets, their TOC timeline and, per page, the end (link timeline) of
        // the last packet completing on it: granule cross-check below
        std::vector<std::pair<uint32_t, int64_t>> page_end;  // (page, end) in page order
        int64_t page_begin_final = 0;
        int64_t cum = 0;
        int last_dur = 0;
        for (size_t i = 2; i < LB.packets.size(); i++, k++) {
            const Packet &p = LB.packets[i];
            uint8_t *dst = d->data.data() + w;
            d->off[k] = (uint32_t)w;
            d->len[k] = p.len;
            memcpy(dst, buf + p.off0, p.len0);
            w += p.len0;
            for (uint32_t j = 0; j + 1 < p.nparts; j++) {
                const auto &piece = LB.more[p.more0 + j];
                memcpy(d->data.data() + w, buf + piece.first, piece.second);
                w += piece.second;
            }
            const int dur = pkt_samples48(dst, p.len);
            d->sample_off[k] = cum;
            if (page_end.empty() || page_end.back().first != p.page) {
                page_begin_final = cum;
                page_end.push_back({p.page, cum + dur});
            } else {
                page_end.back().second = cum + dur;
            }
            cum += dur;
            last_dur = dur;
            if (p.len > 0 && (dst[0] & 0x4)) d->stereo_packets++;
        }
        lk.npkts = (uint32_t)k - lk.first;
        lk.toc_total = cum;
        if (!lk.npkts) last_dur = 0;

        // ---- granule cross-check over this link's pages
        const int final_page = page_end.empty() ? -1 : (int)page_end.back().first;
        int64_t last_granule = 0;
        for (uint32_t pgi : LB.pages)
            if (pages[pgi].granule != -1 && pages[pgi].granule > last_granule) last_granule = pages[pgi].granule;
        lk.last_granule = last_granule;
        uint32_t checked = 0, match = 0;
        int64_t base = 0;  // lenient: granule minus TOC time, re-based at each mismatch
        int64_t end_trim = 0;
        for (auto &pe : page_end) {
            int64_t g = pages[pe.first].granule;
            if (g == -1) continue;
            checked++;
            const bool is_final = (int)pe.first == final_page;
            if (g == pe.second) match++;
            else if (is_final && pe.second - g >= 0 && pe.second - g < last_dur) match++;
            if (opts.strict) continue;
            int64_t want = pe.second + base;
            if (g == want) continue;
            if (is_final && g < want && want - g <= pe.second - page_begin_final) {
                end_trim = want - g;
                continue;
            }
            st.granule_warnings++;
            if (!is_final) base = g - pe.second;
        }
        d->granule_checked += checked;
        d->granule_match += match;
        if (opts.strict) {
            lk.play48 = last_granule - lk.preskip;
        } else {
            lk.end_trim = end_trim;
            lk.play48 = lk.toc_total - end_trim - lk.preskip;
            if (lk.play48 < 0) lk.play48 = 0;
        }
        d->links.push_back(lk);
    }
    if (opts.strict && d->granule_match != d->granule_checked) {
        *err = "granule cross-check failed (" + std::to_string(d->granule_match) + "/" +
               std::to_string(d->granule_checked) + ")";
        return false;
    }
    const Link &l0 = d->links[0];
    d->channels = l0.channels;
    d->preskip = l0.preskip;
    d->last_granule = l0.last_granule;
    d->toc_total = 0;
    d->pcm_samples_48k = 0;
    for (const Link &lk : d->links) {
        if (lk.channels != d->channels) { *err = "channel count changes between chained links"; return false; }
        d->toc_total += lk.toc_total;
        d->pcm_samples_48k += lk.play48 > 0 ? lk.play48 : 0;
    }
    if (d->links.size() == 1) d->pcm_samples_48k = l0.play48;  // may be <= 0 (nothing playable)
    return true;
}

// Chains of one link: packets [lk.first, lk.first + lk.npkts), link timeline.
static bool link_chains(const Demuxed &d, const Link &lk, int64_t step, int64_t preroll, std::vector<Chain> *out,
                        std::string *err)
{
    const int64_t total_play = lk.play48;
    if (total_play <= 0) return true;
    const size_t n_pk = lk.npkts;
    if (n_pk == 0) { *err = "no audio packets"; return false; }
    const int64_t *so = d.sample_off.data() + lk.first;
    // cum(k): TOC time at the start of packet k (the link's total past the last)
    auto cum = [&](int64_t k) -> int64_t { return k < (int64_t)n_pk ? so[k] : lk.toc_total; };
    int64_t kept = 0;
    for (int64_t k = 0; k * step < total_play; k++) {
        int64_t play_start = k * step;
        int64_t play_end = std::min((k + 1) * step, total_play);
        int64_t toc_start = play_start + lk.preskip;
        int64_t toc_end = play_end + lk.preskip;
        int64_t start_target = std::max<int64_t>(0, toc_start - preroll);
        // first packet whose span reaches past the preroll start
        int64_t first = (int64_t)(std::upper_bound(so, so + n_pk, start_target) - so) - 1;
        if (first < 0) first = 0;
        if (cum(first + 1) <= start_target) first++;
        // last packet starting before the window end
        int64_t last = (int64_t)(std::lower_bound(so, so + n_pk, toc_end) - so) - 1;
        if (last < first) last = first;
        if (first >= (int64_t)n_pk) { *err = "chain starts past the last packet"; return false; }
        int64_t trim = toc_start - so[first];
        int64_t keep = play_end - play_start;
        if (trim < 0) { *err = "negative trim"; return false; }
        int64_t capn = cum(last + 1) - cum(first);
        if (capn < trim + keep) {
            *err = "chain " + std::to_string(out->size()) + ": capacity " + std::to_string(capn) + " < trim+keep";
            return false;
        }
        out->push_back({(uint32_t)(lk.first + first), (uint32_t)(last - first + 1), (uint32_t)trim,
                        (uint32_t)keep});
        kept += keep;
    }
    if (kept != total_play) { *err = "internal: sum(keep) != playable"; return false; }
    return true;
}

bool build_chains(const Demuxed &d, double seconds, double preroll_ms, std::vector<Chain> *out, std::string *err)
{
    out->clear();
    const int64_t step = (int64_t)(seconds * 48000);
    const int64_t preroll = (int64_t)(preroll_ms * 48);
    if (d.pcm_samples_48k <= 0) { *err = "no playable samples"; return false; }
    if (step <= 0) { *err = "bad segment length"; return false; }
    if (d.off.empty()) { *err = "no audio packets"; return false; }
    for (const Link &lk : d.links)
        if (!link_chains(d, lk, step, preroll, out, err)) return false;
    return true;
}

}  // namespace cuo_host
