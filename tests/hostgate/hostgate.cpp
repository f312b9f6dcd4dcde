// hostgate.cpp -- the device decoder source compiled for the host CPU.
//
// Test infrastructure (the reference's "same-source host gate",
// harness_host.cpp): the __host__ __device__ decoder in cuopus_amd/csrc/dec
// runs here on the CPU so tests can localise a mismatch against the oracle to
// the decoder logic rather than to GPU execution.  Never used by the product
// path, which only runs the decoder on the GPU.
#include <stdlib.h>
#include <string.h>
#include "../../cuopus_amd/csrc/dec/opus_dec.h"

using namespace CUO_NS;

// Frame sink: appends each decoded frame (decoded into sc->frame_out) to the
// caller's PCM buffer.
struct HostOut {
    int16_t *buf;
    int16_t *dst;
    long pos;  // sample frames
    int cc;    // output channels, interleaved
    void emit(int n)
    {
        memcpy(dst + pos * cc, buf, sizeof(int16_t) * (size_t)n * cc);
        pos += n;
    }
};

// One chain from a fresh decoder of `channels` output channels (1 or 2; 2
// needs the stereo build) with OPUS_SET_GAIN(gain) (Q8 dB); pcm interleaved,
// cap in sample frames.
extern "C" long hg_decode_chain_cfg(const uint8_t *data, const uint32_t *off, const uint32_t *len,
                                    const uint8_t *lost, int npk, int16_t *pcm, long cap,
                                    int32_t *nsamp, uint32_t *rng, int channels, int gain)
{
    OpusDecState *st = (OpusDecState *)malloc(sizeof(OpusDecState));
    DecScratch *sc = (DecScratch *)malloc(sizeof(DecScratch));
    DecHeap *heap = (DecHeap *)malloc(sizeof(DecHeap));
    memset(sc, 0xA5, sizeof(DecScratch));  // device slots start uninitialised
    memset(st, 0x5A, sizeof(OpusDecState));
    memset(heap, 0x3C, sizeof(DecHeap));
    dec_bind(st, sc, heap);
    opus_dec_init(st, channels);
    st->decode_gain = gain;
    HostOut out{sc->frame_out, pcm, 0, st->channels};
    long total = 0;
    long ret = 0;
    for (int k = 0; k < npk; k++) {
        const uint8_t *p = data + off[k];
        int r;
        if (cap - total < 5760) { ret = OPUS_BUFFER_TOO_SMALL; break; }
        if (lost && lost[k]) {
            int dur = packet_nb_samples(p, (int)len[k]);
            if (dur <= 0) { ret = dur ? dur : OPUS_INVALID_PACKET; break; }
            const uint8_t *q = k + 1 < npk ? data + off[k + 1] : nullptr;
            if (q && !lost[k + 1] && packet_has_lbrr(q, (int)len[k + 1]) > 0)
                r = opus_decode_packet(st, q, (int)len[k + 1], out, dur, 1, sc);
            else
                r = opus_decode_packet(st, nullptr, 0, out, dur, 0, sc);
        } else {
            // as the kernel's chain loop: plain CELT packets first
            r = celt_plain_packet(st, p, (int)len[k], out, sc);
            if (r == kNotPlain) r = opus_decode_packet(st, p, (int)len[k], out, 5760, 0, sc);
        }
        if (r < 0) { ret = r; break; }
        if (nsamp) nsamp[k] = r;
        if (rng) rng[k] = st->rangeFinal;
        total += r;
        if (out.pos != total) { ret = OPUS_INTERNAL_ERROR; break; }
    }
    free(st);
    free(sc);
    free(heap);
    return ret < 0 ? ret : total;
}

extern "C" long hg_decode_chain_gain(const uint8_t *data, const uint32_t *off, const uint32_t *len,
                                     const uint8_t *lost, int npk, int16_t *pcm, long cap,
                                     int32_t *nsamp, uint32_t *rng, int gain)
{
    return hg_decode_chain_cfg(data, off, len, lost, npk, pcm, cap, nsamp, rng, 1, gain);
}

extern "C" long hg_decode_chain(const uint8_t *data, const uint32_t *off, const uint32_t *len,
                                const uint8_t *lost, int npk, int16_t *pcm, long cap,
                                int32_t *nsamp, uint32_t *rng)
{
    return hg_decode_chain_gain(data, off, len, lost, npk, pcm, cap, nsamp, rng, 0);
}

extern "C" void hg_ec_schedule(const uint8_t *buf, uint32_t n, uint32_t *out)
{
    static const uint8_t icdf_tab[8] = {120, 112, 96, 64, 32, 16, 8, 0};
    RangeDec d;
    int k = 0;
    ec_init(&d, buf, n);
    for (int i = 0; i < 64; i++) out[k++] = (uint32_t)ec_bit_logp(&d, 1);
    for (int i = 0; i < 32; i++) out[k++] = (uint32_t)ec_icdf(&d, icdf_tab, 7);
    for (int i = 0; i < 32; i++) out[k++] = ec_uint(&d, 2000);
    for (int i = 0; i < 16; i++) out[k++] = ec_bits(&d, 11);
    out[k++] = (uint32_t)ec_tell(&d);
    out[k++] = d.rng;
    out[k++] = (uint32_t)d.error;
}

extern "C" int hg_max_channels() { return kMaxCh; }

extern "C" int hg_sizes(int *out)
{
    out[0] = (int)sizeof(OpusDecState);
    out[1] = (int)sizeof(DecScratch);
    out[2] = (int)sizeof(CeltState);
    out[3] = (int)sizeof(SilkDecoder);
    out[4] = (int)sizeof(DecHeap);
    return 5;
}

// Several chains decoded one after another in the same state and scratch
// memory (no re-initialisation of the scratch between chains), the way a
// persistent wavefront reuses its LDS block.  pcm receives each chain's full
// output back to back; ret[c] = samples or the negative error.
extern "C" void hg_decode_chains_reuse(const uint8_t *data, const uint32_t *off, const uint32_t *len,
                                       const uint8_t *lost, const uint32_t *first, const uint32_t *npkts,
                                       int nch, int16_t *pcm, long cap, long *ret)
{
    OpusDecState *st = (OpusDecState *)malloc(sizeof(OpusDecState));
    DecScratch *sc = (DecScratch *)malloc(sizeof(DecScratch));
    DecHeap *heap = (DecHeap *)malloc(sizeof(DecHeap));
    memset(sc, 0xA5, sizeof(DecScratch));
    memset(st, 0x5A, sizeof(OpusDecState));
    memset(heap, 0x3C, sizeof(DecHeap));
    long pos = 0;
    for (int c = 0; c < nch; c++) {
        dec_bind(st, sc, heap);
        opus_dec_init(st, 1);
        HostOut out{sc->frame_out, pcm + pos, 0, 1};
        long r = 0;
        for (uint32_t k = 0; k < npkts[c]; k++) {
            uint32_t p = first[c] + k;
            const uint8_t *pk = data + off[p];
            int n;
            if (cap - pos - out.pos < 5760) { r = OPUS_BUFFER_TOO_SMALL; break; }
            if (lost && lost[p]) {
                int dur = packet_nb_samples(pk, (int)len[p]);
                if (dur <= 0) { r = dur ? dur : OPUS_INVALID_PACKET; break; }
                if (k + 1 < npkts[c] && !lost[p + 1] && packet_has_lbrr(data + off[p + 1], (int)len[p + 1]) > 0)
                    n = opus_decode_packet(st, data + off[p + 1], (int)len[p + 1], out, dur, 1, sc);
                else
                    n = opus_decode_packet(st, nullptr, 0, out, dur, 0, sc);
            } else {
                n = celt_plain_packet(st, pk, (int)len[p], out, sc);
                if (n == kNotPlain) n = opus_decode_packet(st, pk, (int)len[p], out, 5760, 0, sc);
            }
            if (n < 0) { r = n; break; }
        }
        ret[c] = r < 0 ? r : out.pos;
        if (r >= 0) pos += out.pos;
    }
    free(st);
    free(sc);
    free(heap);
}
