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
    const char *t = getenv("CUOPUS_HOST_THREADS");
    if (t && *t) o.threads = atoi(t);
    const char *m = getenv("CUOPUS_DEMUX_PARALLEL_MIN");
    if (m && *m) o.parallel_min_bytes = (size_t)strtoull(m, nullptr, 10);
    return o;
}

int host_threads(const DemuxOpts &opts)
{
    if (opts.threads > 0) return opts.threads;
    const unsigned hw = std::thread::hardware_concurrency();
    return hw == 0 ? 1 : (int)std::min(16u, hw);
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
// (No member initialisers: the parallel scan sizes its packet table without
// a zero-filling pass; Packet() / Packet{} still zero every field.)
struct Packet {
    uint64_t off0;
    uint32_t len0;
    uint32_t more0, nparts;
    uint32_t len;
    uint32_t page;  // page on which it completes
};

struct LinkBuild {
    uint32_t serial = 0;
    bool eos = false;
    bool have_seq = false;
    uint32_t last_seq = 0;
    RawVec<Packet> packets;  // header packets included
    std::vector<std::pair<uint64_t, uint32_t>> more;  // (file offset, length) of continuation pieces
    Packet partial{};
    bool has_partial = false;
    std::vector<uint32_t> pages;  // page indices of this link
};
}  // namespace

namespace {
// Pass 1 over the whole stream, one page at a time: pages, links and the
// packets of each link (the general path: any framing, foreign streams,
// chained links, missing pages, CRC policy).  False with *err set on a strict
// error.
bool scan_sequential(const uint8_t *buf, size_t n, const DemuxOpts &opts, std::vector<Page> &pages,
                     std::vector<LinkBuild> &links, DemuxStats &st, size_t &done_spans, std::string *err)
{
    // ---- pages, and the packets of the Opus links, in one pass
    pages.reserve(n / 4000 + 16);
    Span spans[255];
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
    return true;
}

// Pass 1 for a well-formed single-stream file, in parallel page ranges.
// Eligible streams (checked first, sequentially and cheaply): pages back to
// back from byte 0, one serial number, the first page a BOS page opening with
// OpusHead, no other BOS page, EOS only on the last page, consecutive
// sequence numbers, every page's continuation flag equal to whether the
// previous page ended inside a packet, the last page ending on a packet
// boundary, and (with check_crc) every CRC and version byte good.  Then the
// sequential scan would take every page into one link and drop nothing, and
// each packet is the spans from its start to its first done span: thread t
// builds the packets starting on its pages (following the last one onto the
// next pages), written at their final index.  Returns false (nothing
// written) when the stream is not eligible.
bool scan_parallel(const uint8_t *buf, size_t n, const DemuxOpts &opts, int nthreads, std::vector<Page> &pages,
                   std::vector<LinkBuild> &links, DemuxStats &st, size_t &done_spans)
{
    struct PageAt {
        uint64_t body;
        uint32_t body_len;
        uint32_t starts;  // packets starting on the page
        uint32_t done;    // spans ending a packet
        int nsegs;
    };
    std::vector<PageAt> at;
    at.reserve(n / 4000 + 16);
    pages.clear();
    pages.reserve(n / 4000 + 16);
    size_t pos = 0;
    bool prev_partial = false;
    while (pos < n) {
        if (pos + 27 > n || memcmp(buf + pos, "OggS", 4) != 0) return false;
        const int nsegs = buf[pos + 26];
        if (nsegs == 0) return false;
        const uint8_t *seg = buf + pos + 27;
        const size_t body = pos + 27 + (size_t)nsegs;
        if (body > n) return false;
        uint32_t body_len = 0, done = 0;
        for (int i = 0; i < nsegs; i++) {
            body_len += seg[i];
            done += seg[i] < 255;
        }
        if (body + body_len > n) return false;
        Page pg;
        memcpy(&pg.granule, buf + pos + 6, 8);
        memcpy(&pg.serial, buf + pos + 14, 4);
        memcpy(&pg.seq, buf + pos + 18, 4);
        pg.flags = buf[pos + 5];
        const bool partial = seg[nsegs - 1] == 255;
        if (pages.empty()) {
            if ((pg.flags & 3) != 2) return false;
            if (seg[0] == 255 || seg[0] < 19 || memcmp(buf + body, "OpusHead", 8) != 0) return false;
        } else {
            const Page &q = pages.back();
            if ((pg.flags & 2) || (q.flags & 4) || pg.serial != q.serial || pg.seq != q.seq + 1 ||
                ((pg.flags & 1) != 0) != prev_partial)
                return false;
        }
        const uint32_t spans = done + (partial ? 1 : 0);
        at.push_back({body, body_len, spans - (prev_partial ? 1u : 0u), done, nsegs});
        pages.push_back(pg);
        prev_partial = partial;
        pos = body + body_len;
    }
    if (pages.empty() || prev_partial) return false;
    const size_t np = pages.size();
    // page ranges of about equal bytes; packet index of each range's first packet
    std::vector<size_t> cut(1, 0);
    for (size_t p = 0, acc = 0; p < np; p++) {
        acc += at[p].body_len;
        if (acc * (size_t)nthreads >= n * cut.size() && cut.size() < (size_t)nthreads && p + 1 < np)
            cut.push_back(p + 1);
    }
    cut.push_back(np);
    const int nr = (int)cut.size() - 1;
    std::vector<size_t> first_pkt(nr + 1, 0);
    for (int r = 0; r < nr; r++) {
        size_t c = 0;
        for (size_t p = cut[r]; p < cut[r + 1]; p++) c += at[p].starts;
        first_pkt[r + 1] = first_pkt[r] + c;
    }
    if (first_pkt[nr] < 2) return false;
    std::vector<uint8_t> crc_ok(opts.check_crc ? np : 0, 1);
    links.clear();
    links.emplace_back();
    LinkBuild &L = links.back();
    L.serial = pages[0].serial;
    L.eos = (pages.back().flags & 4) != 0;
    L.have_seq = true;
    L.last_seq = pages.back().seq;
    L.packets.resize(first_pkt[nr]);
    L.pages.resize(np);
    std::vector<std::vector<std::pair<uint64_t, uint32_t>>> more(nr);
    parallel_for(nr, [&](int r) {
        if (opts.check_crc) {
            for (size_t p = cut[r]; p < cut[r + 1]; p++) {
                const uint64_t hdr = at[p].body - 27 - (uint64_t)at[p].nsegs;
                uint32_t want;
                memcpy(&want, buf + hdr + 22, 4);
                static const uint8_t kZero[4] = {0, 0, 0, 0};
                uint32_t c = ogg_crc(buf + hdr, 22);
                c = ogg_crc(kZero, 4, c);
                c = ogg_crc(buf + hdr + 26, 1 + (size_t)at[p].nsegs + at[p].body_len, c);
                crc_ok[p] = c == want && buf[hdr + 4] == 0;
            }
        }
        for (size_t p = cut[r]; p < cut[r + 1]; p++) L.pages[p] = (uint32_t)p;
        Packet *out = L.packets.data() + first_pkt[r];
        std::vector<std::pair<uint64_t, uint32_t>> &mv = more[r];
        Packet cur;
        bool open = false;   // a packet of this range is being assembled
        bool skip = (pages[cut[r]].flags & 1) != 0;  // spans of a packet owned by the previous range
        for (size_t p = cut[r]; p < np; p++) {
            if (p >= cut[r + 1] && !open) break;
            const uint8_t *seg = buf + at[p].body - (uint64_t)at[p].nsegs;
            uint64_t run_start = at[p].body;
            uint32_t run_len = 0;
            for (int i = 0; i < at[p].nsegs; i++) {
                run_len += seg[i];
                const bool end = seg[i] < 255;
                if (!end && i + 1 < at[p].nsegs) continue;
                // a span: [run_start, run_start + run_len), ending a packet when `end`
                if (skip) {
                    if (end) skip = false;
                } else if (!open) {
                    if (p >= cut[r + 1]) break;
                    cur = Packet();
                    cur.off0 = run_start;
                    cur.len0 = run_len;
                    cur.nparts = 1;
                    cur.len = run_len;
                    open = true;
                } else {
                    if (cur.nparts == 1) cur.more0 = (uint32_t)mv.size();
                    mv.push_back({run_start, run_len});
                    cur.nparts++;
                    cur.len += run_len;
                }
                if (end && open) {
                    cur.page = (uint32_t)p;
                    *out++ = cur;
                    open = false;
                }
                run_start += run_len;
                run_len = 0;
            }
        }
    });
    for (uint8_t ok : crc_ok)
        if (!ok) return false;
    // the continuation pieces of all ranges in one list (more0 re-based)
    size_t nmore = 0;
    for (int r = 0; r < nr; r++) nmore += more[r].size();
    L.more.reserve(nmore);
    for (int r = 0; r < nr; r++) {
        const uint32_t base = (uint32_t)L.more.size();
        if (base)
            for (size_t k = first_pkt[r]; k < first_pkt[r + 1]; k++)
                if (L.packets[k].nparts > 1) L.packets[k].more0 += base;
        L.more.insert(L.more.end(), more[r].begin(), more[r].end());
    }
    st.pages = (uint32_t)np;
    done_spans = 0;
    for (const PageAt &a : at) done_spans += a.done;
    return true;
}
}  // namespace

bool demux(const uint8_t *buf, size_t n, Demuxed *d, std::string *err, const DemuxOpts &opts)
{
    *d = Demuxed();
    if (!buf || n == 0) { *err = "empty stream"; return false; }
    DemuxStats &st = d->stats;
    const int nthreads = n >= opts.parallel_min_bytes ? host_threads(opts) : 1;
    std::vector<Page> pages;
    std::vector<LinkBuild> links;
    size_t done_spans = 0;  // spans ending a packet on any accepted page (error messages only)
    if (nthreads <= 1 || !scan_parallel(buf, n, opts, nthreads, pages, links, st, done_spans)) {
        st = DemuxStats();
        pages.clear();
        links.clear();
        done_spans = 0;
        if (!scan_sequential(buf, n, opts, pages, links, st, done_spans, err)) return false;
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
    // audio packets run from index 2 of each link; a link's packets are cut
    // into nr ranges for the parallel passes below
    auto nranges = [&](size_t na) { return (int)std::min<size_t>((size_t)nthreads, std::max<size_t>(1, na / 4096)); };
    // range_bytes[li][r + 1]: packet bytes of link li's ranges up to r (inclusive)
    uint64_t audio_bytes = 0;
    size_t n_audio = 0;
    std::vector<std::vector<uint64_t>> range_bytes(links.size());
    for (size_t li = 0; li < links.size(); li++) {
        const LinkBuild &L = links[li];
        const size_t na = L.packets.size() > 2 ? L.packets.size() - 2 : 0;
        const int nr = nranges(na);
        std::vector<uint64_t> &rb = range_bytes[li];
        rb.assign(nr + 1, 0);
        parallel_for(nr, [&](int r) {
            uint64_t b = 0;
            for (size_t i = na * (size_t)r / nr; i < na * (size_t)(r + 1) / nr; i++) b += L.packets[i + 2].len;
            rb[r + 1] = b;
        });
        for (int r = 0; r < nr; r++) rb[r + 1] += rb[r];
        audio_bytes += rb[nr];
        n_audio += na;
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
        // packets and their TOC timeline, in parallel packet ranges: each range
        // copies its packets' bytes to their place (byte offsets from the
        // ranges' byte totals above), forms its timeline from 0 and notes per
        // page the timeline at the start of the first and the end of the last
        // packet completing on it; the ranges' timelines are then shifted by
        // the durations before them.  Those page spans drive the granule
        // cross-check below.
        const size_t na = LB.packets.size() - 2;
        struct PageSpan { uint32_t page; int64_t begin, end; };
        std::vector<PageSpan> page_end;  // in page order
        int last_dur = 0;
        int64_t cum = 0;
        {
            const int nr = nranges(na);
            auto lo = [&](int r) { return na * (size_t)r / (size_t)nr; };
            const std::vector<uint64_t> &rb = range_bytes[li];
            std::vector<int64_t> base(nr + 1, 0);
            std::vector<uint32_t> stereo(nr, 0);
            std::vector<std::vector<PageSpan>> part(nr);
            uint8_t *const data = d->data.data();
            uint32_t *const offs = d->off.data() + k;
            uint32_t *const lens = d->len.data() + k;
            int64_t *const so = d->sample_off.data() + k;
            const Packet *const pk = LB.packets.data() + 2;
            const std::pair<uint64_t, uint32_t> *const more = LB.more.data();
            parallel_for(nr, [&](int r) {
                uint64_t ww = w + rb[r];
                uint32_t sp = 0;
                int64_t c = 0;
                std::vector<PageSpan> &pe = part[r];
                for (size_t i = lo(r); i < lo(r + 1); i++) {
                    const Packet &p = pk[i];
                    uint8_t *dst = data + ww;
                    offs[i] = (uint32_t)ww;
                    lens[i] = p.len;
                    memcpy(dst, buf + p.off0, p.len0);
                    ww += p.len0;
                    for (uint32_t j = 0; j + 1 < p.nparts; j++) {
                        memcpy(data + ww, buf + more[p.more0 + j].first, more[p.more0 + j].second);
                        ww += more[p.more0 + j].second;
                    }
                    const int dur = pkt_samples48(dst, p.len);
                    so[i] = c;
                    if (pe.empty() || pe.back().page != p.page) pe.push_back({p.page, c, c + dur});
                    else pe.back().end = c + dur;
                    c += dur;
                    if (p.len > 0 && (dst[0] & 0x4)) sp++;
                }
                base[r + 1] = c;
                stereo[r] = sp;
            });
            for (int r = 0; r < nr; r++) {
                base[r + 1] += base[r];
                d->stereo_packets += stereo[r];
            }
            if (nr > 1)
                parallel_for(nr, [&](int r) {
                    if (r == 0) return;
                    for (size_t i = lo(r); i < lo(r + 1); i++) so[i] += base[r];
                });
            for (int r = 0; r < nr; r++)
                for (PageSpan x : part[r]) {
                    x.begin += base[r];
                    x.end += base[r];
                    if (!page_end.empty() && page_end.back().page == x.page) page_end.back().end = x.end;
                    else page_end.push_back(x);
                }
            cum = base[nr];
            if (na) last_dur = (int)(cum - so[na - 1]);
            w += rb[nr];
            k += na;
        }
        const int64_t page_begin_final = page_end.empty() ? 0 : page_end.back().begin;
        lk.npkts = (uint32_t)k - lk.first;
        lk.toc_total = cum;
        if (!lk.npkts) last_dur = 0;

        // ---- granule cross-check over this link's pages
        const int final_page = page_end.empty() ? -1 : (int)page_end.back().page;
        int64_t last_granule = 0;
        for (uint32_t pgi : LB.pages)
            if (pages[pgi].granule != -1 && pages[pgi].granule > last_granule) last_granule = pages[pgi].granule;
        lk.last_granule = last_granule;
        uint32_t checked = 0, match = 0;
        int64_t base = 0;  // lenient: granule minus TOC time, re-based at each mismatch
        int64_t end_trim = 0;
        for (const PageSpan &pe : page_end) {
            int64_t g = pages[pe.page].granule;
            if (g == -1) continue;
            checked++;
            const bool is_final = (int)pe.page == final_page;
            if (g == pe.end) match++;
            else if (is_final && pe.end - g >= 0 && pe.end - g < last_dur) match++;
            if (opts.strict) continue;
            int64_t want = pe.end + base;
            if (g == want) continue;
            if (is_final && g < want && want - g <= pe.end - page_begin_final) {
                end_trim = want - g;
                continue;
            }
            st.granule_warnings++;
            if (!is_final) base = g - pe.end;
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
