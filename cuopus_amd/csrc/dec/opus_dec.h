// opus_dec.h -- Opus packet / frame layer over the SILK and CELT decoders.
//
// Restates libopus 1.5.2 (vendor/src/opus_decoder.c) for a 48 kHz
// fixed-point decoder of one output channel, or of one or two in the stereo
// build (kMaxCh == 2; PCM interleaved): opus_decoder_init (:129-174),
// smooth_fade (:205-220),
// opus_decode_frame (:237-669: mode dispatch, celt_accum on the SILK PCM,
// SILK/hybrid/CELT transitions with the 5 ms PLC crossfade, CELT redundancy in
// both directions, rangeFinal = rng ^ redundant_rng), opus_decode_native
// (:670-811: PLC chunking, the in-band FEC branch, per-frame loop) and the
// packet parser (vendor/src/opus.c:153-354).
//
// libopus recurses into opus_decode_frame for concealment and transitions;
// here those calls go to plc_frame(), a separate instantiation of the frame
// decoder that has no data and therefore can never recurse again, so the
// device call graph is acyclic and its stack is statically sized.
#pragma once
#include "celt.h"
#include "silk.h"
#include "celt_plc.h"

namespace CUO_NS {

enum { MODE_SILK_ONLY = 1000, MODE_HYBRID = 1001, MODE_CELT_ONLY = 1002 };
enum { BW_NB = 1101, BW_MB = 1102, BW_WB = 1103, BW_SWB = 1104, BW_FB = 1105 };
enum { OPUS_OK = 0, OPUS_BAD_ARG = -1, OPUS_BUFFER_TOO_SMALL = -2, OPUS_INTERNAL_ERROR = -3,
       OPUS_INVALID_PACKET = -4, OPUS_UNIMPLEMENTED = -5 };

struct OpusDecState {
    SilkDecoder silk;
    CeltState celt;
    SilkControl silk_ctl;
    int32_t channels;         // output channels (1, or 2 in the stereo build)
    int32_t stream_channels;
    int32_t bandwidth;
    int32_t mode;
    int32_t prev_mode;
    int32_t frame_size;
    int32_t prev_redundancy;
    int32_t last_packet_duration;
    uint32_t rangeFinal;
    int32_t decode_gain;  // OPUS_SET_GAIN value, Q8 dB (0: none)
};

// Per-call temporaries.  SILK and CELT decode one after the other inside a
// frame, so their scratch shares storage.
constexpr int kMaxFrameOut = 2880;

struct DecScratch {
    union {
        CeltScratch celt;
        SilkScratch silk;
    };
    // SILK output, transition and redundancy frames (caller-owned, HBM on the
    // device: only lane-parallel loops and the frame decoders' output touch
    // them)
    int16_t *pcm_silk;
    int16_t *pcm_transition;
    int16_t *redundant_audio;
    int16_t *silk_rs_tmp;  // one channel's 48 kHz SILK output before interleaving (two output channels)
    RangeDec dec;  // the frame's range decoder (shared by SILK and CELT)
    int16_t frame_sizes[48];  // packet_parse output
    // One frame of output, caller-owned (HBM on the device): a frame decode
    // writes at most st->frame_size <= 2880 samples (60 ms) per channel.
    int16_t *frame_out;
};

// Memory of a decoder that lives outside the state/scratch block (sizes per
// output channel times kMaxCh).
struct DecHeap {
    int32_t celt_mem[kMaxCh * kMemCap];
    int16_t frame_out[kMaxCh * kMaxFrameOut];
    CeltPlcScratch plc;
    int16_t pcm_silk[kMaxCh * 480];
    int16_t pcm_transition[kMaxCh * 240];
    int16_t redundant_audio[kMaxCh * 240];
    int32_t silk_cng_exc[kMaxCh * kSilkMaxFrame];
    int32_t celt_freq[kMaxCh * 960];
    int32_t celt_etmp[kOverlap];
    int32_t silk_cng_sig[kSilkMaxFrame + kSilkMaxOrder];
    int16_t silk_rs_tmp[kMaxCh == 1 ? 2 : 960];
    int32_t silk_exc[kMaxCh == 1 ? 1 : kMaxCh * kSilkMaxFrame];  // excitation buffers (stereo build)
};

CUO_FN void dec_bind(OpusDecState *st, DecScratch *sc, DecHeap *h)
{
    st->celt.mem = st->celt.mem_base = h->celt_mem;
    st->celt.plc = &h->plc;
    st->celt.freq = h->celt_freq;
    st->celt.etmp = h->celt_etmp;
    st->silk.cng_sig = h->silk_cng_sig;
    st->silk.cng_exc = h->silk_cng_exc;
    if (kMaxCh == 2)
        for (int n = 0; n < kMaxCh; n++) st->silk.ch[n].exc_Q14.bind(h->silk_exc + kSilkMaxFrame * n);
    sc->pcm_silk = h->pcm_silk;
    sc->pcm_transition = h->pcm_transition;
    sc->redundant_audio = h->redundant_audio;
    sc->silk_rs_tmp = h->silk_rs_tmp;
    sc->frame_out = h->frame_out;
}

// Output channels of the decoder; constant 1 in the mono build.
CUO_FN int dec_cc(const OpusDecState *st) { return kMaxCh == 1 ? 1 : uniform(st->channels); }

CUO_FN void opus_dec_init(OpusDecState *st, int channels)
{
    st->channels = kMaxCh == 1 ? 1 : channels;
    silk_init_decoder(&st->silk);
    celt_init(&st->celt, st->channels);
    st->silk_ctl.internalSampleRate = 0;
    st->silk_ctl.payloadSize_ms = 0;
    st->silk_ctl.nChannelsInternal = 0;
    st->silk_ctl.nChannelsAPI = st->channels;
    st->stream_channels = st->channels;
    st->bandwidth = 0;
    st->mode = 0;
    st->prev_mode = 0;
    st->frame_size = 48000 / 400;
    st->prev_redundancy = 0;
    st->last_packet_duration = 0;
    st->rangeFinal = 0;
    st->decode_gain = 0;
}

// ------------------------------------------------------------------ TOC helpers
CUO_FN int toc_mode(uint8_t toc)
{
    if (toc & 0x80) return MODE_CELT_ONLY;
    if ((toc & 0x60) == 0x60) return MODE_HYBRID;
    return MODE_SILK_ONLY;
}
CUO_FN int toc_bandwidth(uint8_t toc)
{
    if (toc & 0x80) {
        int bw = BW_MB + ((toc >> 5) & 0x3);
        return bw == BW_MB ? BW_NB : bw;
    }
    if ((toc & 0x60) == 0x60) return (toc & 0x10) ? BW_FB : BW_SWB;
    return BW_NB + ((toc >> 5) & 0x3);
}
CUO_FN int toc_samples_per_frame(uint8_t toc)
{
    int a;
    if (toc & 0x80) { a = (toc >> 3) & 0x3; return (48000 << a) / 400; }
    if ((toc & 0x60) == 0x60) return (toc & 0x08) ? 48000 / 50 : 48000 / 100;
    a = (toc >> 3) & 0x3;
    return a == 3 ? 48000 * 60 / 1000 : (48000 << a) / 100;
}
// opus_packet_get_nb_samples (opus_decoder.c:1121-1137); <= 0 on error
CUO_FN int packet_nb_samples(const uint8_t *p, int len)
{
    if (len < 1) return OPUS_BAD_ARG;
    int count = p[0] & 0x3;
    if (count == 0) count = 1;
    else if (count != 3) count = 2;
    else if (len < 2) return OPUS_INVALID_PACKET;
    else count = p[1] & 0x3F;
    int samples = count * toc_samples_per_frame(p[0]);
    if (samples * 25 > 48000 * 3) return OPUS_INVALID_PACKET;
    return samples;
}

CUO_FN int parse_size(const uint8_t *data, int len, int16_t *size)
{
    if (len < 1) { *size = -1; return -1; }
    if (data[0] < 252) { *size = data[0]; return 1; }
    if (len < 2) { *size = -1; return -1; }
    *size = (int16_t)(4 * data[1] + data[0]);
    return 2;
}

// opus_packet_parse_impl, not self-delimited (opus.c:194-354).  Returns the
// frame count and fills size[] and *payload_offset, or a negative error.
CUO_FN int packet_parse(const uint8_t *data, int len, int16_t *size, int *payload_offset)
{
    if (len < 0) return OPUS_BAD_ARG;
    if (len == 0) return OPUS_INVALID_PACKET;
    const uint8_t *data0 = data;
    int framesize = toc_samples_per_frame(data[0]);
    uint8_t toc = *data++;
    len--;
    int32_t last_size = len;
    int count, bytes;
    switch (toc & 0x3) {
    case 0:
        count = 1;
        break;
    case 1:
        count = 2;
        if (len & 0x1) return OPUS_INVALID_PACKET;
        last_size = len / 2;
        size[0] = (int16_t)last_size;
        break;
    case 2:
        count = 2;
        bytes = parse_size(data, len, size);
        len -= bytes;
        if (size[0] < 0 || size[0] > len) return OPUS_INVALID_PACKET;
        data += bytes;
        last_size = len - size[0];
        break;
    default: {
        if (len < 1) return OPUS_INVALID_PACKET;
        uint8_t ch = *data++;
        count = ch & 0x3F;
        if (count <= 0 || framesize * count > 5760) return OPUS_INVALID_PACKET;
        len--;
        if (ch & 0x40) {
            int p;
            do {
                if (len <= 0) return OPUS_INVALID_PACKET;
                p = *data++;
                len--;
                int tmp = p == 255 ? 254 : p;
                len -= tmp;
            } while (p == 255);
        }
        if (len < 0) return OPUS_INVALID_PACKET;
        int cbr = !(ch & 0x80);
        if (!cbr) {
            last_size = len;
            for (int i = 0; i < count - 1; i++) {
                bytes = parse_size(data, len, size + i);
                len -= bytes;
                if (size[i] < 0 || size[i] > len) return OPUS_INVALID_PACKET;
                data += bytes;
                last_size -= bytes + size[i];
            }
            if (last_size < 0) return OPUS_INVALID_PACKET;
        } else {
            last_size = len / count;
            if (last_size * count != len) return OPUS_INVALID_PACKET;
            for (int i = 0; i < count - 1; i++) size[i] = (int16_t)last_size;
        }
        break;
    }
    }
    if (last_size > 1275) return OPUS_INVALID_PACKET;
    size[count - 1] = (int16_t)last_size;
    *payload_offset = (int)(data - data0);
    return count;
}

// opus_packet_has_lbrr (opus_decoder.c:1139-1162)
CUO_FN int packet_has_lbrr(const uint8_t *p, int len)
{
    if (toc_mode(p[0]) == MODE_CELT_ONLY) return 0;
    int fs = toc_samples_per_frame(p[0]);
    int nb_frames = fs > 960 ? fs / 960 : 1;
    int16_t size[48];
    int off;
    int ret = packet_parse(p, len, size, &off);
    if (ret <= 0) return ret;
    // libopus reads the first byte of the first frame even when that frame is
    // empty (then the byte lies past the packet): such a byte reads as 0 here,
    // as in the oracle's zero-padded packet copy, so the result never depends
    // on what follows the packet in memory
    const int b0 = off < len ? p[off] : 0;
    int lbrr = (b0 >> (7 - nb_frames)) & 0x1;
    if (p[0] & 0x4) lbrr = lbrr || ((b0 >> (6 - 2 * nb_frames)) & 0x1);
    return lbrr;
}

// log2(10) / (20 * 256) in Q25: Q8 dB -> Q10 log2 (QCONST16(6.48814081e-4f, 25))
constexpr int32_t kGainLog2Q25 = 21771;

// smooth_fade (opus_decoder.c:205-220) over `overlap` samples of `channels`
// interleaved channels
CUO_FN void smooth_fade(const int16_t *in1, const int16_t *in2, int16_t *out, int overlap, int channels)
{
    channels = kMaxCh == 1 ? 1 : channels;
    CUO_PFOR(e, 0, overlap * channels) {
        const int i = channels == 1 ? e : e >> 1;
        int16_t w = (int16_t)mul16_16_q15(tab::kWindow[i], tab::kWindow[i]);
        out[e] = (int16_t)((mul16_16(w, in2[e]) + mul16_16(kQ15One - w, in1[e])) >> 15);
    }
}

// ------------------------------------------------------------------ frame decode
// Three instantiations: kFrameData decodes a frame (libopus' opus_decode_frame
// with data), kFrameConceal conceals and may chunk a long gap into 20 ms
// pieces, kFrameConcealLeaf conceals at most 20 ms.  Data frames reach
// concealment only through the leaf (transition PLC is <= 5 ms), so the call
// graph has no cycle.
enum { kFrameData = 0, kFrameConceal = 1, kFrameConcealLeaf = 2 };
template <int Kind>
CUO_FN int frame_decode(OpusDecState *st, const uint8_t *data, int len, int16_t *pcm, int frame_size,
                        int decode_fec, DecScratch *sc);

CUO_FN int plc_frame(OpusDecState *st, int16_t *pcm, int frame_size, DecScratch *sc)
{
    return frame_decode<kFrameConcealLeaf>(st, nullptr, 0, pcm, frame_size, 0, sc);
}

template <int Kind>
CUO_FN int frame_decode(OpusDecState *st, const uint8_t *data, int len, int16_t *pcm, int frame_size,
                        int decode_fec, DecScratch *sc)
{
    constexpr bool HasData = Kind == kFrameData;
    const int F20 = 960, F10 = 480, F5 = 240, F2_5 = 120;
    st = lds_uniform(st);
    sc = lds_uniform(sc);
    const int CC = dec_cc(st);  // output channels, interleaved in pcm
    RangeDec &dec = sc->dec;
    if (frame_size < F2_5) return OPUS_BUFFER_TOO_SMALL;
    frame_size = imin(frame_size, 48000 / 25 * 3);
    if constexpr (HasData) {
        if (len <= 1)  // a payload of <= 1 byte is concealed (opus_decoder.c:285-290)
            return frame_decode<kFrameConceal>(st, nullptr, 0, pcm, frame_size, 0, sc);
    }
    if (!HasData) {
        data = nullptr;
        frame_size = imin(frame_size, st->frame_size);
    }
    int audiosize, mode, bandwidth;
    if (data != nullptr) {
        audiosize = st->frame_size;
        mode = st->mode;
        bandwidth = st->bandwidth;
        ec_init(&dec, data, (uint32_t)len);
    } else {
        audiosize = frame_size;
        mode = st->prev_redundancy ? MODE_CELT_ONLY : st->prev_mode;
        bandwidth = 0;
        if (mode == 0) {
            CUO_PFOR(i, 0, audiosize * CC) pcm[i] = 0;
            return audiosize;
        }
        if constexpr (Kind == kFrameConceal) {
            if (audiosize > F20) {
                // chunk long concealment into <= 20 ms calls (opus_decoder.c:321-333)
                do {
                    int ret = frame_decode<kFrameConcealLeaf>(st, nullptr, 0, pcm, imin(audiosize, F20), 0, sc);
                    if (ret < 0) return ret;
                    pcm += ret * CC;
                    audiosize -= ret;
                } while (audiosize > 0);
                return frame_size;
            }
        }
        if (audiosize < F20) {
            if (audiosize > F10) audiosize = F10;
            else if (mode != MODE_SILK_ONLY && audiosize > F5 && audiosize < F10) audiosize = F5;
        }
    }
    const int celt_accum = (mode != MODE_CELT_ONLY) && (frame_size >= F10);
    int transition = 0;
    int16_t *pcm_transition = nullptr;
    if (HasData && data != nullptr && st->prev_mode > 0 &&
        ((mode == MODE_CELT_ONLY && st->prev_mode != MODE_CELT_ONLY && !st->prev_redundancy) ||
         (mode != MODE_CELT_ONLY && st->prev_mode == MODE_CELT_ONLY)))
        transition = 1;
    if constexpr (HasData) {
        if (transition && mode == MODE_CELT_ONLY) {
            pcm_transition = sc->pcm_transition;
            plc_frame(st, pcm_transition, imin(F5, audiosize), sc);
        }
    }
    if (audiosize > frame_size) return OPUS_BAD_ARG;
    frame_size = audiosize;

    int16_t *pcm_silk = sc->pcm_silk;
    if (mode != MODE_CELT_ONLY) {
        int16_t *pcm_ptr = celt_accum ? pcm : pcm_silk;
        if (st->prev_mode == MODE_CELT_ONLY) silk_reset_decoder(&st->silk);
        st->silk_ctl.payloadSize_ms = imax(10, 1000 * audiosize / 48000);
        if (data != nullptr) {
            st->silk_ctl.nChannelsInternal = st->stream_channels;
            if (mode == MODE_SILK_ONLY) {
                if (bandwidth == BW_NB) st->silk_ctl.internalSampleRate = 8000;
                else if (bandwidth == BW_MB) st->silk_ctl.internalSampleRate = 12000;
                else st->silk_ctl.internalSampleRate = 16000;
            } else {
                st->silk_ctl.internalSampleRate = 16000;
            }
        }
        int lost_flag = data == nullptr ? 1 : 2 * !!decode_fec;
        int decoded = 0;
        do {
            int32_t silk_frame_size;
            int first_frame = decoded == 0;
            int silk_ret = silk_decode(&st->silk, &st->silk_ctl, lost_flag, first_frame, &dec, pcm_ptr,
                                       &silk_frame_size, &sc->silk, sc->silk_rs_tmp);
            if (silk_ret) {
                if (lost_flag) {
                    silk_frame_size = frame_size;
                    CUO_PFOR(i, 0, frame_size * CC) pcm_ptr[i] = 0;
                } else {
                    return OPUS_INTERNAL_ERROR;
                }
            }
            pcm_ptr += silk_frame_size * CC;
            decoded += silk_frame_size;
        } while (decoded < frame_size);
    }

    int start_band = 0, redundancy = 0, redundancy_bytes = 0, celt_to_silk = 0;
    if (HasData && !decode_fec && mode != MODE_CELT_ONLY && data != nullptr &&
        ec_tell(&dec) + 17 + 20 * (mode == MODE_HYBRID) <= 8 * len) {
        redundancy = mode == MODE_HYBRID ? ec_bit_logp(&dec, 12) : 1;
        if (redundancy) {
            celt_to_silk = ec_bit_logp(&dec, 1);
            redundancy_bytes = mode == MODE_HYBRID ? (int)ec_uint(&dec, 256) + 2 : len - ((ec_tell(&dec) + 7) >> 3);
            len -= redundancy_bytes;
            if (len * 8 < ec_tell(&dec)) {
                len = 0;
                redundancy_bytes = 0;
                redundancy = 0;
            }
            dec.storage -= (uint32_t)redundancy_bytes;
        }
    }
    if (mode != MODE_CELT_ONLY) start_band = 17;
    if (redundancy) transition = 0;
    if constexpr (HasData) {
        if (transition && mode != MODE_CELT_ONLY) {
            pcm_transition = sc->pcm_transition;
            plc_frame(st, pcm_transition, imin(F5, audiosize), sc);
        }
    }
    if (bandwidth) {
        int endband = 21;
        switch (bandwidth) {
        case BW_NB: endband = 13; break;
        case BW_MB: case BW_WB: endband = 17; break;
        case BW_SWB: endband = 19; break;
        default: endband = 21; break;
        }
        st->celt.end = endband;
    }
    st->celt.stream_channels = st->stream_channels;  // CELT_SET_CHANNELS
    int16_t *redundant_audio = sc->redundant_audio;
    uint32_t redundant_rng = 0;
    if (HasData && redundancy && celt_to_silk) {
        st->celt.start = 0;
        celt_decode(&st->celt, data + len, redundancy_bytes, redundant_audio, F5, nullptr, 0, &sc->celt);
        redundant_rng = st->celt.rng;
    }
    st->celt.start = start_band;
    int celt_ret = 0;
    if (mode != MODE_SILK_ONLY) {
        int celt_frame_size = imin(F20, frame_size);
        if (mode != st->prev_mode && st->prev_mode > 0 && !st->prev_redundancy) celt_reset(&st->celt);
        celt_ret = celt_decode(&st->celt, decode_fec ? nullptr : data, len, pcm, celt_frame_size,
                               data != nullptr ? &dec : nullptr, celt_accum, &sc->celt);
    } else {
        alignas(4) static constexpr uint8_t silence[4] = {0xFF, 0xFF, 0, 0};  // a 2-byte frame
        if (!celt_accum) CUO_PFOR(i, 0, frame_size * CC) pcm[i] = 0;
        if (st->prev_mode == MODE_HYBRID && !(redundancy && celt_to_silk && st->prev_redundancy)) {
            st->celt.start = 0;
            celt_decode(&st->celt, silence, 2, pcm, F2_5, nullptr, celt_accum, &sc->celt);
        }
    }
    if (mode != MODE_CELT_ONLY && !celt_accum) CUO_PFOR(i, 0, frame_size * CC) pcm[i] = sat16(pcm[i] + pcm_silk[i]);
    if (HasData && redundancy && !celt_to_silk) {
        celt_reset(&st->celt);
        st->celt.start = 0;
        celt_decode(&st->celt, data + len, redundancy_bytes, redundant_audio, F5, nullptr, 0, &sc->celt);
        redundant_rng = st->celt.rng;
        smooth_fade(pcm + CC * (frame_size - F2_5), redundant_audio + CC * F2_5, pcm + CC * (frame_size - F2_5),
                    F2_5, CC);
    }
    if (HasData && redundancy && celt_to_silk && (st->prev_mode != MODE_SILK_ONLY || st->prev_redundancy)) {
        CUO_PFOR(i, 0, F2_5 * CC) pcm[i] = redundant_audio[i];
        smooth_fade(redundant_audio + CC * F2_5, pcm + CC * F2_5, pcm + CC * F2_5, F2_5, CC);
    }
    if (HasData && transition) {
        if (audiosize >= F5) {
            CUO_PFOR(i, 0, F2_5 * CC) pcm[i] = pcm_transition[i];
            smooth_fade(pcm_transition + CC * F2_5, pcm + CC * F2_5, pcm + CC * F2_5, F2_5, CC);
        } else {
            smooth_fade(pcm_transition, pcm, pcm, F2_5, CC);
        }
    }
    if (st->decode_gain) {
        // OPUS_SET_GAIN (opus_decoder.c:639-648): gain = 2^(decode_gain / (20 log10(2) 256)), applied
        // as a Q16 multiply with saturation
        const int32_t gain = celt_exp2((int16_t)mul16_16_p15(kGainLog2Q25, st->decode_gain));
        CUO_PFOR(i, 0, frame_size * CC) pcm[i] = (int16_t)saturate(mul16_32_p16(pcm[i], gain), 32767);
    }
    st->rangeFinal = len <= 1 ? 0 : (HasData ? dec.rng ^ redundant_rng : 0);
    st->prev_mode = mode;
    st->prev_redundancy = redundancy && !celt_to_silk;
    return celt_ret < 0 ? celt_ret : audiosize;
}

// Output of the packet decoder: every frame is decoded into out.buf
// (DecScratch::frame_out) and handed to out.emit(n) before the next frame.

// opus_decode_native's concealment branch (opus_decoder.c:713-729)
template <class Out>
CUO_FN int decode_plc(OpusDecState *st, Out &out, int frame_size, DecScratch *sc)
{
    if (frame_size % 120 != 0) return OPUS_BAD_ARG;
    int count = 0;
    do {
        int ret = frame_decode<kFrameConceal>(st, nullptr, 0, out.buf, frame_size - count, 0, sc);
        if (ret < 0) return ret;
        out.emit(ret);
        count += ret;
    } while (count < frame_size);
    st->last_packet_duration = count;
    return count;
}

// opus_decode (opus_decoder.c:670-811) for int16 output.  data == nullptr or
// len == 0 conceals frame_size samples.  Returns samples or a negative error.
template <class Out>
CUO_FN int opus_decode_packet(OpusDecState *st, const uint8_t *data, int len, Out &out, int frame_size,
                              int decode_fec, DecScratch *sc)
{
    if (frame_size <= 0) return OPUS_BAD_ARG;
    if (decode_fec < 0 || decode_fec > 1) return OPUS_BAD_ARG;
    if ((decode_fec || len == 0 || data == nullptr) && frame_size % 120 != 0) return OPUS_BAD_ARG;
    if (len == 0 || data == nullptr) return decode_plc(st, out, frame_size, sc);
    if (len < 0) return OPUS_BAD_ARG;
    int packet_mode = toc_mode(data[0]);
    int packet_bandwidth = toc_bandwidth(data[0]);
    int packet_frame_size = toc_samples_per_frame(data[0]);
    int packet_stream_channels = (data[0] & 0x4) ? 2 : 1;
    int16_t *size = lds_uniform(sc->frame_sizes);
    int offset;
    int count = packet_parse(data, len, size, &offset);
    if (count < 0) return count;
    // the mono build keeps one coded channel: stereo packets go to the stereo build
    if (kMaxCh == 1 && packet_stream_channels != 1) return OPUS_UNIMPLEMENTED;
    data += offset;
    int nframes = count;
    if (decode_fec) {
        if (frame_size < packet_frame_size || packet_mode == MODE_CELT_ONLY || st->mode == MODE_CELT_ONLY)
            return decode_plc(st, out, frame_size, sc);
        int duration_copy = st->last_packet_duration;
        if (frame_size - packet_frame_size != 0) {
            int ret = decode_plc(st, out, frame_size - packet_frame_size, sc);
            if (ret < 0) { st->last_packet_duration = duration_copy; return ret; }
        }
        nframes = 1;  // the LBRR of the first frame, over packet_frame_size samples
    } else if (count * packet_frame_size > frame_size) {
        return OPUS_BUFFER_TOO_SMALL;
    }
    st->mode = packet_mode;
    st->bandwidth = packet_bandwidth;
    st->frame_size = packet_frame_size;
    st->stream_channels = packet_stream_channels;
    int nb_samples = 0;
    for (int i = 0; i < nframes; i++) {
        const int limit = decode_fec ? packet_frame_size : frame_size - nb_samples;
        const int ret = frame_decode<kFrameData>(st, data, size[i], out.buf, limit, decode_fec, sc);
        if (ret < 0) return ret;
        out.emit(ret);
        data += size[i];
        nb_samples += ret;
    }
    st->last_packet_duration = decode_fec ? frame_size : nb_samples;
    return decode_fec ? frame_size : nb_samples;
}

// A plain CELT packet: one coded channel, CELT-only, decoded after a CELT-only
// packet or first after a reset, every frame longer than one byte.  For such
// a packet opus_decode_frame (opus_decoder.c:222-653) reduces to the CELT
// decoder alone: no mode transition, no redundancy, no SILK, no CELT reset,
// no accumulation.  Decodes it as opus_decode_packet would (same state
// updates, same outputs) with the CELT frame decoder inlined, so a caller that
// loops over such packets pays no call per frame.  Returns the packet's
// samples, a negative decoder error, or kNotPlain (nothing touched) when the
// packet is not of this kind.
constexpr int kNotPlain = -0x7fff;
template <class Out>
CUO_FI int celt_plain_packet(OpusDecState *st, const uint8_t *data, int len, Out &out, DecScratch *sc)
{
    if (kMaxCh != 1 || len <= 1) return kNotPlain;
    const uint8_t toc = data[0];
    if (toc_mode(toc) != MODE_CELT_ONLY || (toc & 0x4)) return kNotPlain;
    if (st->prev_mode != 0 && st->prev_mode != MODE_CELT_ONLY) return kNotPlain;
    int16_t *size = lds_uniform(sc->frame_sizes);
    int offset;
    const int count = packet_parse(data, len, size, &offset);
    if (count <= 0) return kNotPlain;  // the general path reports the error
    for (int i = 0; i < count; i++)
        if (size[i] <= 1) return kNotPlain;
    const int fs = toc_samples_per_frame(toc);
    const int bw = toc_bandwidth(toc);
    st->mode = MODE_CELT_ONLY;
    st->bandwidth = bw;
    st->frame_size = fs;
    st->stream_channels = 1;
    const int endband = bw == BW_NB ? 13 : (bw == BW_MB || bw == BW_WB ? 17 : (bw == BW_SWB ? 19 : 21));
    data += offset;
    int nb_samples = 0;
    for (int i = 0; i < count; i++) {
        RangeDec &dec = sc->dec;
        const int n = size[i];
        ec_init(&dec, data, (uint32_t)n);
        st->celt.end = endband;
        st->celt.stream_channels = 1;
        st->celt.start = 0;
        const int r = celt_decode_frame<false>(&st->celt, data, n, out.buf, fs, &dec, 0, &sc->celt);
        if (st->decode_gain) {
            const int32_t gain = celt_exp2((int16_t)mul16_16_p15(kGainLog2Q25, st->decode_gain));
            CUO_PFOR(j, 0, fs) out.buf[j] = (int16_t)saturate(mul16_32_p16(out.buf[j], gain), 32767);
        }
        st->rangeFinal = dec.rng;
        st->prev_mode = MODE_CELT_ONLY;
        st->prev_redundancy = 0;
        if (r < 0) return r;
        out.emit(fs);
        data += n;
        nb_samples += fs;
    }
    st->last_packet_duration = nb_samples;
    return nb_samples;
}

}  // namespace CUO_NS
