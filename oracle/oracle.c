/* oracle.c -- CPU oracle for the batched Opus decode path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, the smoke check and bench.py's
 * cpu_baseline leg may load this library, and only as the checker (or as the
 * reported CPU baseline), never as the thing measured or shipped.
 *
 * The decode arithmetic is the reference's own oracle: vendored libopus 1.5.2
 * compiled FIXED_POINT from its sources (see Makefile).  This file restates the
 * glue the reference wraps around it:
 *   - the chain protocol: a fresh OpusDecoder per chain, every packet through
 *     opus_decode(..., 5760, fec) with per-packet sample count and final range
 *     (cuopus_kernels.cuh:386-414; harness2.cu:51-91, 200-253);
 *   - the loss rule: lost packet p is decoded as FEC from p+1 when p+1 is in
 *     the chain, not lost and carries LBRR (opus_packet_has_lbrr,
 *     vendor/src/opus_decoder.c:1139-1162); otherwise PLC with p's TOC
 *     duration (vendor/src/opus_decoder.c:683-698, 745-779);
 *   - the 49-tap Kaiser 48k->16k resampler and its taps
 *     (cuopus_kernels.cuh:441-456, 468-494; gpu_resample.py:20-49), summed in
 *     float32 in tap order;
 *   - the phase-1 range decoder schedule (schedule.inc:19-37) through the real
 *     libopus entdec.
 * Parity of the restated glue is pinned in tests/ (closed forms for the taps,
 * float64 evaluation for the resampler, golden vectors for the schedule).
 */
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "opus.h"
#include "entdec.h"

#define ORC_MAX_PKT_SAMPLES 5760
#define RS_TAPS 49
#define RS_PAD 24
#define RS_BETA 8.6

int orc_version(void) { return 1; }

/* opus_packet_has_lbrr on a copy of the packet followed by zero bytes:
 * libopus reads the first byte of the first frame even when that frame is
 * empty, i.e. possibly past the packet; in the copy that byte is 0, so the
 * FEC decision never depends on what follows the packet in memory. */
static int orc_has_lbrr(const unsigned char *q, opus_int32 len)
{
    unsigned char small[1536];
    unsigned char *buf = (size_t)len + 8 <= sizeof(small) ? small : (unsigned char *)malloc((size_t)len + 8);
    if (!buf) return OPUS_ALLOC_FAIL;
    memcpy(buf, q, (size_t)len);
    memset(buf + len, 0, 8);
    int r = opus_packet_has_lbrr(buf, len);
    if (buf != small) free(buf);
    return r;
}

/* Decode one chain with a fresh decoder of `channels` output channels and
 * OPUS_SET_GAIN(gain) (Q8 dB; 0: none).  Packet k is data[off[k] ..
 * off[k]+len[k]).  lost may be NULL.  pcm receives every decoded int16
 * sample, interleaved when channels == 2 (capacity cap samples per channel);
 * nsamp[k] and rng[k] receive opus_decode's return and OPUS_GET_FINAL_RANGE
 * after packet k (either may be NULL).  Returns total samples per channel, or
 * a negative opus error. */
long orc_decode_chain_cfg(const unsigned char *data, const uint32_t *off, const uint32_t *len, const uint8_t *lost,
                          int npk, int16_t *pcm, long cap, int32_t *nsamp, uint32_t *rng, int channels, int gain)
{
    int err = 0;
    OpusDecoder *dec = opus_decoder_create(48000, channels, &err);
    if (!dec) return err ? err : -1;
    if (gain) opus_decoder_ctl(dec, OPUS_SET_GAIN(gain));
    long total = 0;
    for (int k = 0; k < npk; k++) {
        const unsigned char *p = data + off[k];
        int16_t *o = pcm + total * channels;
        int ret;
        if (cap - total < ORC_MAX_PKT_SAMPLES) { ret = OPUS_BUFFER_TOO_SMALL; goto done; }
        if (lost && lost[k]) {
            int dur = opus_packet_get_nb_samples(p, (opus_int32)len[k], 48000);
            if (dur <= 0) { ret = dur ? dur : OPUS_INVALID_PACKET; goto done; }
            const unsigned char *q = (k + 1 < npk) ? data + off[k + 1] : NULL;
            if (q && !lost[k + 1] && orc_has_lbrr(q, (opus_int32)len[k + 1]) > 0)
                ret = opus_decode(dec, q, (opus_int32)len[k + 1], o, dur, 1);
            else
                ret = opus_decode(dec, NULL, 0, o, dur, 0);
        } else {
            ret = opus_decode(dec, p, (opus_int32)len[k], o, ORC_MAX_PKT_SAMPLES, 0);
        }
        if (ret < 0) goto done;
        if (nsamp) nsamp[k] = ret;
        if (rng) {
            opus_uint32 r = 0;
            opus_decoder_ctl(dec, OPUS_GET_FINAL_RANGE(&r));
            rng[k] = r;
        }
        total += ret;
        continue;
done:
        opus_decoder_destroy(dec);
        return ret;
    }
    opus_decoder_destroy(dec);
    return total;
}

/* The chain protocol's default decoder: mono, no gain. */
long orc_decode_chain(const unsigned char *data, const uint32_t *off, const uint32_t *len, const uint8_t *lost,
                      int npk, int16_t *pcm, long cap, int32_t *nsamp, uint32_t *rng)
{
    return orc_decode_chain_cfg(data, off, len, lost, npk, pcm, cap, nsamp, rng, 1, 0);
}

/* ---- many chains, optionally on a thread pool (CPU baseline) ---- */
typedef struct {
    const unsigned char *data;
    const uint32_t *off, *len;
    const uint8_t *lost;
    const uint32_t *first, *npkts;
    const uint64_t *pcm_off;
    int16_t *pcm;
    const uint64_t *pcm_cap;
    int32_t *nsamp;
    uint32_t *rng;
    int64_t *chain_ret;
    int n_chains;
    int channels;            /* decoder output channels */
    const int32_t *gain;     /* per chain OPUS_SET_GAIN value, or NULL */
    volatile int next;
    pthread_mutex_t mu;
} orc_job;

static void *orc_worker(void *arg)
{
    orc_job *j = (orc_job *)arg;
    for (;;) {
        pthread_mutex_lock(&j->mu);
        int c0 = j->next;
        j->next += 4;
        pthread_mutex_unlock(&j->mu);
        if (c0 >= j->n_chains) break;
        for (int c = c0; c < c0 + 4 && c < j->n_chains; c++) {
            uint32_t f = j->first[c], np = j->npkts[c];
            int32_t *ns = (int32_t *)calloc(np ? np : 1, sizeof(int32_t));
            uint32_t *rg = (uint32_t *)calloc(np ? np : 1, sizeof(uint32_t));
            j->chain_ret[c] = orc_decode_chain_cfg(
                j->data, j->off + f, j->len + f, j->lost ? j->lost + f : NULL,
                (int)np, j->pcm + j->pcm_off[c] * (uint64_t)j->channels, (long)j->pcm_cap[c], ns, rg,
                j->channels, j->gain ? j->gain[c] : 0);
            /* A packet inside several chains (preroll overlap) is recorded by
             * the last chain that holds it: chain c owns packets below the
             * first packet of chain c + 1.  Deterministic on any thread count. */
            uint32_t own_end = c + 1 < j->n_chains ? j->first[c + 1] : 0xFFFFFFFFu;
            for (uint32_t k = 0; k < np && f + k < own_end; k++) {
                if (j->nsamp) j->nsamp[f + k] = ns[k];
                if (j->rng) j->rng[f + k] = rg[k];
            }
            free(ns);
            free(rg);
        }
    }
    return NULL;
}

/* Decode n_chains chains with nthreads workers (dynamic chunks of 4), each
 * with a fresh decoder of `channels` output channels and the chain's gain
 * (gain may be NULL).  pcm_off / pcm_cap are in samples per channel.
 * Returns 0 if every chain decoded, else the first negative chain return. */
int orc_decode_chains_cfg(const unsigned char *data, const uint32_t *off, const uint32_t *len, const uint8_t *lost,
                          const uint32_t *first, const uint32_t *npkts, int n_chains, const uint64_t *pcm_off,
                          const uint64_t *pcm_cap, int16_t *pcm, int32_t *nsamp, uint32_t *rng, int64_t *chain_ret,
                          int nthreads, int channels, const int32_t *gain)
{
    orc_job j;
    memset(&j, 0, sizeof(j));
    j.data = data; j.off = off; j.len = len; j.lost = lost;
    j.first = first; j.npkts = npkts; j.pcm_off = pcm_off; j.pcm_cap = pcm_cap;
    j.pcm = pcm; j.nsamp = nsamp; j.rng = rng; j.chain_ret = chain_ret;
    j.n_chains = n_chains; j.next = 0;
    j.channels = channels == 2 ? 2 : 1;
    j.gain = gain;
    pthread_mutex_init(&j.mu, NULL);
    if (nthreads < 1) nthreads = 1;
    if (nthreads == 1) {
        orc_worker(&j);
    } else {
        pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * (size_t)nthreads);
        for (int t = 0; t < nthreads; t++) pthread_create(&th[t], NULL, orc_worker, &j);
        for (int t = 0; t < nthreads; t++) pthread_join(th[t], NULL);
        free(th);
    }
    pthread_mutex_destroy(&j.mu);
    for (int c = 0; c < n_chains; c++)
        if (chain_ret[c] < 0) return (int)chain_ret[c];
    return 0;
}

int orc_decode_chains(const unsigned char *data, const uint32_t *off, const uint32_t *len, const uint8_t *lost,
                      const uint32_t *first, const uint32_t *npkts, int n_chains, const uint64_t *pcm_off,
                      const uint64_t *pcm_cap, int16_t *pcm, int32_t *nsamp, uint32_t *rng, int64_t *chain_ret,
                      int nthreads)
{
    return orc_decode_chains_cfg(data, off, len, lost, first, npkts, n_chains, pcm_off, pcm_cap, pcm, nsamp, rng,
                                 chain_ret, nthreads, 1, NULL);
}

/* ---- resampler ---- */

/* Modified Bessel I0 by its power series, in double. */
static double orc_i0(double x)
{
    double s = 1.0, t = 1.0, q = x * x / 4.0;
    for (int k = 1; k < 64; k++) {
        t *= q / ((double)k * k);
        s += t;
        if (t < 1e-18 * s) break;
    }
    return s;
}

/* 49 taps: sinc at cutoff 1/3 (48 kHz Nyquist units) x Kaiser(beta 8.6),
 * normalised to unity DC in double, rounded to float (cuopus_kernels.cuh:481-494). */
void orc_taps(float *out)
{
    double h[RS_TAPS], sum = 0.0;
    const double fc = 1.0 / 3.0;
    for (int i = 0; i < RS_TAPS; i++) {
        double n = i - (RS_TAPS - 1) / 2.0;
        double s = (n == 0.0) ? fc : sin(M_PI * fc * n) / (M_PI * n);
        double r = 2.0 * i / (RS_TAPS - 1) - 1.0;
        double w = orc_i0(RS_BETA * sqrt(1.0 - r * r)) / orc_i0(RS_BETA);
        h[i] = s * w;
        sum += h[i];
    }
    for (int i = 0; i < RS_TAPS; i++) out[i] = (float)(h[i] / sum);
}

/* out[i] = sum_t taps[t] * x[3i - 24 + t] / 32768, zero outside [0, n48),
 * float32 accumulation in tap order (cuopus_kernels.cuh:441-456). */
/* Outputs [o0, o1) of orc_resample (same arithmetic), so threads can split
 * one stream. */
void orc_resample_range(const int16_t *x, int64_t n48, float *out, int64_t o0, int64_t o1)
{
    float taps[RS_TAPS];
    orc_taps(taps);
    for (int64_t i = o0; i < o1; i++) {
        int64_t c = 3 * i;
        float acc = 0.f;
        for (int t = 0; t < RS_TAPS; t++) {
            int64_t li = c - RS_PAD + t;
            float v = (li >= 0 && li < n48) ? (float)x[li] * (1.0f / 32768.0f) : 0.f;
            acc += taps[t] * v;
        }
        out[i] = acc;
    }
}

void orc_resample(const int16_t *x, int64_t n48, float *out, int64_t n_out)
{
    float taps[RS_TAPS];
    orc_taps(taps);
    for (int64_t i = 0; i < n_out; i++) {
        int64_t c = 3 * i;
        float acc = 0.f;
        for (int t = 0; t < RS_TAPS; t++) {
            int64_t li = c - RS_PAD + t;
            float v = (li >= 0 && li < n48) ? (float)x[li] * (1.0f / 32768.0f) : 0.f;
            acc += taps[t] * v;
        }
        out[i] = acc;
    }
}

/* ---- range decoder schedule (schedule.inc:19-37) through libopus entdec ---- */
void orc_ec_schedule(const unsigned char *buf, uint32_t n, uint32_t *out)
{
    static const unsigned char icdf_tab[8] = {120, 112, 96, 64, 32, 16, 8, 0};
    ec_dec dec;
    int k = 0;
    ec_dec_init(&dec, (unsigned char *)buf, n);
    for (int i = 0; i < 64; i++) out[k++] = (uint32_t)ec_dec_bit_logp(&dec, 1);
    for (int i = 0; i < 32; i++) out[k++] = (uint32_t)ec_dec_icdf(&dec, icdf_tab, 7);
    for (int i = 0; i < 32; i++) out[k++] = (uint32_t)ec_dec_uint(&dec, 2000);
    for (int i = 0; i < 16; i++) out[k++] = (uint32_t)ec_dec_bits(&dec, 11);
    out[k++] = (uint32_t)ec_tell(&dec);
    out[k++] = dec.rng;
    out[k++] = (uint32_t)dec.error;
}

/* Entry points for KATs of single functions (tests only). */
int orc_packet_has_lbrr(const unsigned char *p, int len) { return opus_packet_has_lbrr(p, len); }
int orc_packet_nb_samples(const unsigned char *p, int len) { return opus_packet_get_nb_samples(p, len, 48000); }
