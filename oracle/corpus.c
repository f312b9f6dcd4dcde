/* corpus.c -- seeded synthetic Ogg/Opus corpus generator (test data only).
 *
 * Synthesises mono 48 kHz signals in float64, dithers them to int16, encodes
 * them with the vendored libopus 1.5.2 encoder and writes an Ogg/Opus stream
 * with this file's own muxer (RFC 7845: OpusHead v1, OpusTags, CRC-32 with
 * polynomial 0x04C11DB7, ~1 s pages, end-trimmed final granule).
 *
 * Signals:
 *   speech  Rosenberg glottal-flow-derivative pulse train, f0 ~ U(100,250) Hz
 *           redrawn every 200 ms (phase continuous; 25% of segments unvoiced),
 *           through a seeded stable AR(16) (reflection coefficients |k|<=0.9,
 *           redrawn per segment), plus white noise at -30 dB re signal RMS,
 *           scaled to -20 dBFS RMS.
 *   music   pink noise + log sine sweep 50 Hz..20 kHz (10 s period) + 8 tones
 *           with U(-30,-10) dBFS amplitudes redrawn every minute.
 *   mixed   each second is speech or music (coin flip) and the encoder's
 *           signal hint follows it, so the encoder switches modes.
 *
 * usage: mkcorpus --out F --seed S --dur SEC [--signal speech|music]
 *        [--signal speech|music|mixed] [--app voip|audio|lowdelay] [--bw nb|mb|wb|swb|fb]
 *        [--force auto|silk|hybrid|celt] [--vbr 0|1] [--bitrate B]
 *        [--brange LO HI] [--frame N] [--frameset a,b,c] [--fec 0|1] [--dtx 0|1] [--gaps P]
 *        [--loss P] [--complexity C] [--pcm F] [--gain Q8] [--channels 1|2] [--monoflip P]
 * --brange redraws the bitrate every second from U(LO,HI); --frameset draws
 * the frame size (48 kHz samples) per second.  --gain writes the OpusHead
 * output gain (Q7.8 dB).  --channels 2 encodes a stereo stream: left is the
 * signal, right is 0.7 x the signal 20 samples later plus 0.5 x a second,
 * independent signal of the same kind; --monoflip P forces mono packets for
 * each second with probability P (stereo <-> mono switches in one stream).
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "opus.h"

#define FS 48000
#define SET_FORCE_MODE 11002 /* OPUS_SET_FORCE_MODE_REQUEST (vendor/src/opus_private.h:141) */
#define MODE_SILK 1000
#define MODE_HYB 1001
#define MODE_CELT 1002

/* ---- seeded RNG: splitmix64 ---- */
static uint64_t rs;
static uint64_t rnext(void)
{
    uint64_t z = (rs += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
static double runif(void) { return (rnext() >> 11) * (1.0 / 9007199254740992.0); }
static double rgauss(void)
{
    double u = runif(), v = runif();
    if (u < 1e-300) u = 1e-300;
    return sqrt(-2.0 * log(u)) * cos(2.0 * M_PI * v);
}

/* ---- signals ---- */
static double rosenberg_deriv(double ph)
{
    const double tp = 0.40, tn = 0.16;
    if (ph < tp) return 0.5 * M_PI / tp * sin(M_PI * ph / tp);
    if (ph < tp + tn) return -M_PI / (2.0 * tn) * sin(M_PI * (ph - tp) / (2.0 * tn));
    return 0.0;
}

static void ar_from_reflection(double *a, int order)
{
    /* step-up recursion: A(z) = 1 + sum a[i] z^-i from reflection coefficients */
    double k[32], tmp[32];
    for (int i = 0; i < order; i++) k[i] = (runif() * 2.0 - 1.0) * 0.9;
    for (int i = 0; i < order; i++) a[i] = 0.0;
    for (int m = 0; m < order; m++) {
        for (int i = 0; i < m; i++) tmp[i] = a[i] + k[m] * a[m - 1 - i];
        for (int i = 0; i < m; i++) a[i] = tmp[i];
        a[m] = k[m];
    }
}

static void gen_speech(double *x, long n)
{
    double a[16], st[16];
    memset(st, 0, sizeof(st));
    double ph = 0.0, f0 = 150.0;
    int voiced = 1;
    long seg = FS / 5;
    for (long i = 0; i < n; i++) {
        if (i % seg == 0) {
            f0 = 100.0 + 150.0 * runif();
            voiced = runif() >= 0.25;
            ar_from_reflection(a, 16);
        }
        double e;
        if (voiced) {
            ph += f0 / FS;
            if (ph >= 1.0) ph -= 1.0;
            e = rosenberg_deriv(ph) + 0.02 * rgauss();
        } else {
            e = 0.3 * rgauss();
        }
        /* syllable-rate envelope, 4 Hz */
        e *= 0.55 + 0.45 * sin(2.0 * M_PI * 4.0 * i / FS);
        double y = e;
        for (int j = 0; j < 16; j++) y -= a[j] * st[j];
        if (!isfinite(y) || fabs(y) > 1e6) { y = 0; memset(st, 0, sizeof(st)); }
        for (int j = 15; j > 0; j--) st[j] = st[j - 1];
        st[0] = y;
        x[i] = y;
    }
    double e2 = 0;
    for (long i = 0; i < n; i++) e2 += x[i] * x[i];
    double rms = sqrt(e2 / (n ? n : 1));
    double nrm = rms * pow(10.0, -30.0 / 20.0);
    for (long i = 0; i < n; i++) x[i] += nrm * rgauss();
    e2 = 0;
    for (long i = 0; i < n; i++) e2 += x[i] * x[i];
    rms = sqrt(e2 / (n ? n : 1));
    double g = (rms > 0) ? pow(10.0, -20.0 / 20.0) / rms : 0.0;
    for (long i = 0; i < n; i++) x[i] *= g;
}

static void gen_music(double *x, long n)
{
    double b0 = 0, b1 = 0, b2 = 0, b3 = 0, b4 = 0, b5 = 0, b6 = 0;
    double sweep_ph = 0, tone_ph[8] = {0}, tone_f[8], tone_a[8];
    const double f_lo = 50.0, f_hi = 20000.0, period = 10.0;
    for (long i = 0; i < n; i++) {
        if (i % (60L * FS) == 0) {
            for (int t = 0; t < 8; t++) {
                tone_f[t] = 80.0 * pow(2.0, runif() * 7.5);
                tone_a[t] = pow(10.0, (-30.0 + 20.0 * runif()) / 20.0);
            }
        }
        /* pink noise (Kellet's economy filter over white noise) */
        double w = rgauss() * 0.05;
        b0 = 0.99886 * b0 + w * 0.0555179;
        b1 = 0.99332 * b1 + w * 0.0750759;
        b2 = 0.96900 * b2 + w * 0.1538520;
        b3 = 0.86650 * b3 + w * 0.3104856;
        b4 = 0.55000 * b4 + w * 0.5329522;
        b5 = -0.7616 * b5 - w * 0.0168980;
        double pink = b0 + b1 + b2 + b3 + b4 + b5 + b6 + w * 0.5362;
        b6 = w * 0.115926;
        double tt = fmod((double)i / FS, period) / period;
        double f = f_lo * pow(f_hi / f_lo, tt);
        sweep_ph += f / FS;
        sweep_ph -= floor(sweep_ph);
        double s = 0.5 * pink + 0.1 * sin(2.0 * M_PI * sweep_ph);
        for (int t = 0; t < 8; t++) {
            tone_ph[t] += tone_f[t] / FS;
            tone_ph[t] -= floor(tone_ph[t]);
            s += tone_a[t] * sin(2.0 * M_PI * tone_ph[t]);
        }
        x[i] = s;
    }
    double pk = 0;
    for (long i = 0; i < n; i++) if (fabs(x[i]) > pk) pk = fabs(x[i]);
    double g = pk > 0 ? 0.5 / pk : 0.0;
    for (long i = 0; i < n; i++) x[i] *= g;
}

/* ---- Ogg muxer ---- */
static uint32_t crc_tab[256];
static void crc_init(void)
{
    for (uint32_t i = 0; i < 256; i++) {
        uint32_t r = i << 24;
        for (int k = 0; k < 8; k++) r = (r & 0x80000000u) ? (r << 1) ^ 0x04C11DB7u : (r << 1);
        crc_tab[i] = r;
    }
}
typedef struct {
    FILE *f;
    uint32_t serial, seq;
    uint8_t body[1 << 20];
    size_t body_n;
    uint8_t lace[255];
    int nlace;
    int64_t granule;
} ogg_w;

static void ogg_flush(ogg_w *w, int flags)
{
    uint8_t hdr[27 + 255];
    memcpy(hdr, "OggS", 4);
    hdr[4] = 0;
    hdr[5] = (uint8_t)flags;
    for (int i = 0; i < 8; i++) hdr[6 + i] = (uint8_t)((uint64_t)w->granule >> (8 * i));
    for (int i = 0; i < 4; i++) hdr[14 + i] = (uint8_t)(w->serial >> (8 * i));
    for (int i = 0; i < 4; i++) hdr[18 + i] = (uint8_t)(w->seq >> (8 * i));
    memset(hdr + 22, 0, 4);
    hdr[26] = (uint8_t)w->nlace;
    memcpy(hdr + 27, w->lace, (size_t)w->nlace);
    size_t hn = 27 + (size_t)w->nlace;
    uint32_t c = 0;
    for (size_t i = 0; i < hn; i++) c = (c << 8) ^ crc_tab[((c >> 24) ^ hdr[i]) & 0xFF];
    for (size_t i = 0; i < w->body_n; i++) c = (c << 8) ^ crc_tab[((c >> 24) ^ w->body[i]) & 0xFF];
    for (int i = 0; i < 4; i++) hdr[22 + i] = (uint8_t)(c >> (8 * i));
    fwrite(hdr, 1, hn, w->f);
    fwrite(w->body, 1, w->body_n, w->f);
    w->seq++;
    w->body_n = 0;
    w->nlace = 0;
}

/* Append one packet; packets never span pages here (they are < 64 KB). */
static void ogg_packet(ogg_w *w, const uint8_t *p, size_t n, int64_t granule_after)
{
    int need = (int)(n / 255) + 1;
    if (w->nlace + need > 255) ogg_flush(w, 0);
    for (size_t k = n; ; k -= 255) {
        if (k >= 255) { w->lace[w->nlace++] = 255; }
        else { w->lace[w->nlace++] = (uint8_t)k; break; }
    }
    memcpy(w->body + w->body_n, p, n);
    w->body_n += n;
    w->granule = granule_after;
}

static int parse_bw(const char *s)
{
    if (!strcmp(s, "nb")) return OPUS_BANDWIDTH_NARROWBAND;
    if (!strcmp(s, "mb")) return OPUS_BANDWIDTH_MEDIUMBAND;
    if (!strcmp(s, "wb")) return OPUS_BANDWIDTH_WIDEBAND;
    if (!strcmp(s, "swb")) return OPUS_BANDWIDTH_SUPERWIDEBAND;
    return OPUS_BANDWIDTH_FULLBAND;
}

int main(int argc, char **argv)
{
    const char *out = NULL, *pcm_out = NULL, *signal = "speech", *app = "audio";
    const char *bw = "fb", *force = "auto", *frameset = NULL;
    uint64_t seed = 1;
    double dur = 10.0;
    int vbr = 1, bitrate = 32000, blo = 0, bhi = 0, frame = 960, fec = 0, loss = 0, cx = 10, dtx = 0;
    double gaps = 0.0, monoflip = 0.0;
    int gain_q8 = 0, channels = 1;
    for (int i = 1; i < argc; i++) {
        const char *a = argv[i];
#define NEXT (i + 1 < argc ? argv[++i] : "")
        if (!strcmp(a, "--out")) out = NEXT;
        else if (!strcmp(a, "--pcm")) pcm_out = NEXT;
        else if (!strcmp(a, "--seed")) seed = strtoull(NEXT, NULL, 0);
        else if (!strcmp(a, "--dur")) dur = atof(NEXT);
        else if (!strcmp(a, "--signal")) signal = NEXT;
        else if (!strcmp(a, "--app")) app = NEXT;
        else if (!strcmp(a, "--bw")) bw = NEXT;
        else if (!strcmp(a, "--force")) force = NEXT;
        else if (!strcmp(a, "--vbr")) vbr = atoi(NEXT);
        else if (!strcmp(a, "--bitrate")) bitrate = atoi(NEXT);
        else if (!strcmp(a, "--brange")) { blo = atoi(NEXT); bhi = atoi(NEXT); }
        else if (!strcmp(a, "--frame")) frame = atoi(NEXT);
        else if (!strcmp(a, "--frameset")) frameset = NEXT;
        else if (!strcmp(a, "--fec")) fec = atoi(NEXT);
        else if (!strcmp(a, "--loss")) loss = atoi(NEXT);
        else if (!strcmp(a, "--complexity")) cx = atoi(NEXT);
        else if (!strcmp(a, "--dtx")) dtx = atoi(NEXT);
        else if (!strcmp(a, "--gaps")) gaps = atof(NEXT);
        else if (!strcmp(a, "--gain")) gain_q8 = atoi(NEXT);
        else if (!strcmp(a, "--channels")) channels = atoi(NEXT) == 2 ? 2 : 1;
        else if (!strcmp(a, "--monoflip")) monoflip = atof(NEXT);
        else { fprintf(stderr, "unknown arg %s\n", a); return 2; }
#undef NEXT
    }
    if (!out) { fprintf(stderr, "--out required\n"); return 2; }
    int fs_list[8], nfs = 0;
    if (frameset) {
        const char *s = frameset;
        while (*s && nfs < 8) {
            fs_list[nfs++] = atoi(s);
            while (*s && *s != ',') s++;
            if (*s == ',') s++;
        }
    }
    rs = seed;
    crc_init();

    long n = (long)(dur * FS + 0.5);
    double *x = (double *)calloc((size_t)n + 1, sizeof(double));
    long nsec = n / FS + 2;
    unsigned char *is_music = (unsigned char *)calloc((size_t)nsec, 1);
    int mixed = !strcmp(signal, "mixed");
    if (!strcmp(signal, "music")) gen_music(x, n);
    else if (mixed) {
        double *y = (double *)calloc((size_t)n + 1, sizeof(double));
        gen_speech(x, n);
        gen_music(y, n);
        for (long s = 0; s < nsec; s++) is_music[s] = runif() < 0.5;
        for (long i = 0; i < n; i++) if (is_music[i / FS]) x[i] = y[i];
        free(y);
    } else gen_speech(x, n);
    double *xr = NULL;
    if (channels == 2) {
        double *y = (double *)calloc((size_t)n + 1, sizeof(double));
        if (!strcmp(signal, "music") || mixed) gen_music(y, n);
        else gen_speech(y, n);
        xr = (double *)calloc((size_t)n + 1, sizeof(double));
        for (long i = 0; i < n; i++) xr[i] = 0.7 * (i >= 20 ? x[i - 20] : 0.0) + 0.5 * y[i];
        free(y);
    }
    /* --gaps P: each 500 ms segment is silent with probability P (only the
       int16 dither remains), so the encoder's DTX has inactive stretches */
    if (gaps > 0.0) {
        const long seg = FS / 2;
        for (long s0 = 0; s0 < n; s0 += seg)
            if (runif() < gaps)
                for (long i = s0; i < n && i < s0 + seg; i++) {
                    x[i] = 0.0;
                    if (xr) xr[i] = 0.0;
                }
    }

    int application = !strcmp(app, "voip") ? OPUS_APPLICATION_VOIP
                    : !strcmp(app, "lowdelay") ? OPUS_APPLICATION_RESTRICTED_LOWDELAY
                    : OPUS_APPLICATION_AUDIO;
    int err = 0;
    OpusEncoder *enc = opus_encoder_create(FS, channels, application, &err);
    if (!enc) { fprintf(stderr, "encoder: %s\n", opus_strerror(err)); return 1; }
    opus_encoder_ctl(enc, OPUS_SET_COMPLEXITY(cx));
    opus_encoder_ctl(enc, OPUS_SET_VBR(vbr));
    opus_encoder_ctl(enc, OPUS_SET_MAX_BANDWIDTH(parse_bw(bw)));
    opus_encoder_ctl(enc, OPUS_SET_BITRATE(bitrate));
    opus_encoder_ctl(enc, OPUS_SET_INBAND_FEC(fec));
    opus_encoder_ctl(enc, OPUS_SET_PACKET_LOSS_PERC(loss));
    opus_encoder_ctl(enc, OPUS_SET_DTX(dtx));
    if (!strcmp(force, "silk")) opus_encoder_ctl(enc, SET_FORCE_MODE, MODE_SILK);
    else if (!strcmp(force, "hybrid")) opus_encoder_ctl(enc, SET_FORCE_MODE, MODE_HYB);
    else if (!strcmp(force, "celt")) opus_encoder_ctl(enc, SET_FORCE_MODE, MODE_CELT);
    if (!strcmp(signal, "music")) opus_encoder_ctl(enc, OPUS_SET_SIGNAL(OPUS_SIGNAL_MUSIC));
    else opus_encoder_ctl(enc, OPUS_SET_SIGNAL(OPUS_SIGNAL_VOICE));
    opus_int32 lookahead = 0;
    opus_encoder_ctl(enc, OPUS_GET_LOOKAHEAD(&lookahead));

    /* int16 with TPDF dither; padded with silence to flush the lookahead */
    long total_in = n + lookahead + 2 * 2880;
    int16_t *pcm = (int16_t *)calloc((size_t)total_in * channels, sizeof(int16_t));
    for (long i = 0; i < n; i++) {
        for (int c = 0; c < channels; c++) {
            double v = (c ? xr[i] : x[i]) * 32768.0 + (runif() - runif());
            v = floor(v + 0.5);
            if (v > 32767) v = 32767;
            if (v < -32768) v = -32768;
            pcm[i * channels + c] = (int16_t)v;
        }
    }
    free(x);
    free(xr);
    if (pcm_out) {
        FILE *pf = fopen(pcm_out, "wb");
        if (pf) { fwrite(pcm, 2, (size_t)n * channels, pf); fclose(pf); }
    }

    static ogg_w w;
    w.f = fopen(out, "wb");
    if (!w.f) { perror(out); return 1; }
    w.serial = (uint32_t)(seed * 2654435761u) | 1u;
    uint8_t head[19];
    memcpy(head, "OpusHead", 8);
    head[8] = 1; head[9] = (uint8_t)channels;
    head[10] = (uint8_t)(lookahead & 0xFF); head[11] = (uint8_t)(lookahead >> 8);
    head[12] = FS & 0xFF; head[13] = (FS >> 8) & 0xFF; head[14] = (FS >> 16) & 0xFF; head[15] = 0;
    head[16] = (uint8_t)(gain_q8 & 0xFF); head[17] = (uint8_t)((gain_q8 >> 8) & 0xFF); head[18] = 0;
    ogg_packet(&w, head, 19, 0);
    ogg_flush(&w, 0x02);
    uint8_t tags[8 + 4 + 8 + 4];
    memcpy(tags, "OpusTags", 8);
    const char *vendor = "cuopus-corpus";
    uint32_t vl = (uint32_t)strlen(vendor);
    size_t tn = 0;
    uint8_t tbuf[64];
    memcpy(tbuf, tags, 8); tn = 8;
    for (int i = 0; i < 4; i++) tbuf[tn++] = (uint8_t)(vl >> (8 * i));
    memcpy(tbuf + tn, vendor, vl); tn += vl;
    for (int i = 0; i < 4; i++) tbuf[tn++] = 0;
    ogg_packet(&w, tbuf, tn, 0);
    ogg_flush(&w, 0);

    int64_t end_granule = (int64_t)n + lookahead;
    int64_t done = 0;      /* samples fed to the encoder */
    int64_t page_start = 0;
    uint8_t pkt[4000];
    long sec = -1;
    int fsz = frame;
    while (done < end_granule) {
        long s = (long)(done / FS);
        if (s != sec) {
            sec = s;
            if (blo > 0 && bhi >= blo) {
                int br = blo + (int)(runif() * (bhi - blo + 1));
                opus_encoder_ctl(enc, OPUS_SET_BITRATE(br));
            }
            if (nfs > 0) fsz = fs_list[(int)(runif() * nfs) % nfs];
            if (mixed)
                opus_encoder_ctl(enc, OPUS_SET_SIGNAL(is_music[s] ? OPUS_SIGNAL_MUSIC : OPUS_SIGNAL_VOICE));
            if (channels == 2 && monoflip > 0.0)
                opus_encoder_ctl(enc, OPUS_SET_FORCE_CHANNELS(runif() < monoflip ? 1 : OPUS_AUTO));
            /* keep frames inside the second so the draw applies per segment */
        }
        int this_fs = fsz;
        int64_t to_boundary = (int64_t)(sec + 1) * FS - done;
        if (to_boundary < this_fs) this_fs = (int)to_boundary;
        int nb = opus_encode(enc, pcm + done * channels, this_fs, pkt, (opus_int32)sizeof(pkt));
        if (nb < 0) { fprintf(stderr, "encode: %s (fs %d)\n", opus_strerror(nb), this_fs); return 1; }
        done += this_fs;
        int64_t g = done >= end_granule ? end_granule : done;
        ogg_packet(&w, pkt, (size_t)nb, g);
        if (done >= end_granule) break;
        if (done - page_start >= FS) { ogg_flush(&w, 0); page_start = done; }
    }
    ogg_flush(&w, 0x04);
    fclose(w.f);
    opus_encoder_destroy(enc);
    free(pcm);
    return 0;
}
