// celt_plc.h -- pitch-based CELT loss concealment, fixed point.
//
// Restates the periodic branch of celt_decode_lost (vendor/celt/
// celt_decoder.c:700-969) and its helpers: celt_plc_pitch_search (:499-513),
// pitch_downsample / celt_fir5 / find_best_pitch / pitch_search
// (vendor/celt/pitch.c), _celt_autocorr / _celt_lpc / celt_fir / celt_iir
// (vendor/celt/celt_lpc.c).  Integer sums are formed in the same int32
// arithmetic; where libopus unrolls a correlation into 4-wide kernels the sums
// are the same integers, so they are written here as plain loops.
#pragma once
#include "celt.h"

namespace CUO_NS {

constexpr int kPlcLagMax = 720;
constexpr int kPlcLagMin = 100;

CUO_FN int32_t maxabs16(const int16_t *x, int len)
{
    int16_t mx = 0, mn = 0;
    for (int i = 0; i < len; i++) { mx = max16(mx, x[i]); mn = min16(mn, x[i]); }
    return max32((int32_t)mx, -(int32_t)mn);
}
CUO_FN int32_t maxabs32(const int32_t *x, int len)
{
    int32_t mx = 0, mn = 0;
    for (int i = 0; i < len; i++) { mx = max32(mx, x[i]); mn = min32(mn, x[i]); }
    return max32(mx, -mn);
}

// _celt_autocorr (celt_lpc.c:237-317); xx: n int16 of scratch
CUO_FN int celt_autocorr(const int16_t *x, int32_t *ac, const int16_t *window, int overlap, int lag, int n,
                         int16_t *xx)
{
    const int16_t *xp = x;
    if (overlap != 0) {
        for (int i = 0; i < n; i++) xx[i] = x[i];
        for (int i = 0; i < overlap; i++) {
            xx[i] = (int16_t)mul16_16_q15(x[i], window[i]);
            xx[n - i - 1] = (int16_t)mul16_16_q15(x[n - i - 1], window[i]);
        }
        xp = xx;
    }
    int shift;
    {
        int32_t ac0 = 1 + (n << 7);
        if (n & 1) ac0 += mul16_16(xp[0], xp[0]) >> 9;
        for (int i = (n & 1); i < n; i += 2) {
            ac0 += mul16_16(xp[i], xp[i]) >> 9;
            ac0 += mul16_16(xp[i + 1], xp[i + 1]) >> 9;
        }
        shift = celt_ilog2(ac0) - 30 + 10;
        shift = shift / 2;
        if (shift > 0) {
            for (int i = 0; i < n; i++) xx[i] = (int16_t)pshr32(xp[i], shift);
            xp = xx;
        } else {
            shift = 0;
        }
    }
    for (int k = 0; k <= lag; k++) {
        int32_t d = 0;
        for (int i = k; i < n; i++) d = (int32_t)((uint32_t)d + (uint32_t)mul16_16(xp[i], xp[i - k]));
        ac[k] = d;
    }
    shift = 2 * shift;
    if (shift <= 0) ac[0] += shl32(1, -shift);
    if (ac[0] < 268435456) {
        int shift2 = 29 - ilog((uint32_t)ac[0]);
        for (int i = 0; i <= lag; i++) ac[i] = shl32(ac[i], shift2);
        shift -= shift2;
    } else if (ac[0] >= 536870912) {
        int shift2 = 1;
        if (ac[0] >= 1073741824) shift2++;
        for (int i = 0; i <= lag; i++) ac[i] = ac[i] >> shift2;
        shift += shift2;
    }
    return shift;
}

// _celt_lpc (celt_lpc.c:37-135), p <= 24
CUO_FN void celt_lpc(int16_t *lpc_out, const int32_t *ac, int p)
{
    int32_t lpc[kLpcOrder];
    int32_t error = ac[0];
    for (int i = 0; i < p; i++) lpc[i] = 0;
    if (ac[0] != 0) {
        for (int i = 0; i < p; i++) {
            int32_t rr = 0;
            for (int j = 0; j < i; j++) rr += mul32_32_q31(lpc[j], ac[i - j]);
            rr += ac[i + 1] >> 6;
            int32_t r = -frac_div32(shl32(rr, 6), error);
            lpc[i] = r >> 6;
            for (int j = 0; j < (i + 1) >> 1; j++) {
                int32_t t1 = lpc[j], t2 = lpc[i - 1 - j];
                lpc[j] = t1 + mul32_32_q31(r, t2);
                lpc[i - 1 - j] = t2 + mul32_32_q31(r, t1);
            }
            error = error - mul32_32_q31(mul32_32_q31(r, r), error);
            if (error <= (ac[0] >> 10)) break;
        }
    }
    int iter, idx = 0;
    for (iter = 0; iter < 10; iter++) {
        int32_t maxabs = 0;
        for (int i = 0; i < p; i++) {
            int32_t a = abs32(lpc[i]);
            if (a > maxabs) { maxabs = a; idx = i; }
        }
        maxabs = pshr32(maxabs, 13);
        if (maxabs > 32767) {
            maxabs = min32(maxabs, 163838);
            int32_t chirp = 65470 - shl32(maxabs - 32767, 14) / ((maxabs * (idx + 1)) >> 2);
            int32_t cm1 = chirp - 65536;
            for (int i = 0; i < p - 1; i++) {
                lpc[i] = mul32_32_q16(chirp, lpc[i]);
                chirp += pshr32(chirp * cm1, 16);
            }
            lpc[p - 1] = mul32_32_q16(chirp, lpc[p - 1]);
        } else {
            break;
        }
    }
    if (iter == 10) {
        for (int i = 0; i < p; i++) lpc_out[i] = 0;
        lpc_out[0] = 4096;
    } else {
        for (int i = 0; i < p; i++) lpc_out[i] = (int16_t)pshr32(lpc[i], 13);
    }
}

// celt_fir_c (celt_lpc.c:137-185): y[i] = x[i] + sum_m num[m] x[i-m-1], Q12;
// x must have ord samples of history before x[0]
CUO_FN void celt_fir(const int16_t *x, const int16_t *num, int16_t *y, int N, int ord)
{
    for (int i = 0; i < N; i++) {
        int32_t sum = shl32((int32_t)x[i], kSigShift);
        for (int j = 0; j < ord; j++) sum = mac16_16(sum, num[ord - j - 1], x[i + j - ord]);
        y[i] = sround16(sum, kSigShift);
    }
}

// celt_iir (celt_lpc.c:187-235, the 4-sample blocked form).  x may equal y.
// ybuf: N + ord int16 of scratch.  mem receives the last outputs narrowed to
// int16, as libopus stores them.
CUO_FN void celt_iir(const int32_t *x, const int16_t *den, int32_t *y, int N, int ord, int16_t *mem, int16_t *ybuf)
{
    for (int i = 0; i < ord; i++) ybuf[i] = (int16_t)-mem[ord - i - 1];
    for (int i = ord; i < N + ord; i++) ybuf[i] = 0;
    int i;
    for (i = 0; i < N - 3; i += 4) {
        int32_t s[4] = {x[i], x[i + 1], x[i + 2], x[i + 3]};
        for (int j = 0; j < ord; j++) {
            int16_t rd = den[ord - j - 1];
            s[0] = mac16_16(s[0], rd, ybuf[i + j]);
            s[1] = mac16_16(s[1], rd, ybuf[i + j + 1]);
            s[2] = mac16_16(s[2], rd, ybuf[i + j + 2]);
            s[3] = mac16_16(s[3], rd, ybuf[i + j + 3]);
        }
        ybuf[i + ord] = (int16_t)-sround16(s[0], kSigShift);
        y[i] = s[0];
        s[1] = mac16_16(s[1], ybuf[i + ord], den[0]);
        ybuf[i + ord + 1] = (int16_t)-sround16(s[1], kSigShift);
        y[i + 1] = s[1];
        s[2] = mac16_16(s[2], ybuf[i + ord + 1], den[0]);
        s[2] = mac16_16(s[2], ybuf[i + ord], den[1]);
        ybuf[i + ord + 2] = (int16_t)-sround16(s[2], kSigShift);
        y[i + 2] = s[2];
        s[3] = mac16_16(s[3], ybuf[i + ord + 2], den[0]);
        s[3] = mac16_16(s[3], ybuf[i + ord + 1], den[1]);
        s[3] = mac16_16(s[3], ybuf[i + ord], den[2]);
        ybuf[i + ord + 3] = (int16_t)-sround16(s[3], kSigShift);
        y[i + 3] = s[3];
    }
    for (; i < N; i++) {
        int32_t sum = x[i];
        for (int j = 0; j < ord; j++) sum -= mul16_16(den[ord - j - 1], ybuf[i + j]);
        ybuf[i + ord] = sround16(sum, kSigShift);
        y[i] = sum;
    }
    for (int k = 0; k < ord; k++) mem[k] = (int16_t)y[N - k - 1];
}

// pitch.c:75-114
CUO_FN void find_best_pitch(const int32_t *xcorr, const int16_t *y, int len, int max_pitch, int *best_pitch,
                            int yshift, int32_t maxcorr)
{
    int32_t Syy = 1;
    int16_t best_num[2] = {-1, -1};
    int32_t best_den[2] = {0, 0};
    int xshift = celt_ilog2(maxcorr) - 14;
    best_pitch[0] = 0;
    best_pitch[1] = 1;
    for (int j = 0; j < len; j++) Syy = Syy + (mul16_16(y[j], y[j]) >> yshift);
    for (int i = 0; i < max_pitch; i++) {
        if (xcorr[i] > 0) {
            int16_t xcorr16 = (int16_t)vshr32(xcorr[i], xshift);
            int16_t num = (int16_t)mul16_16_q15(xcorr16, xcorr16);
            if (mul16_32_q15(num, best_den[1]) > mul16_32_q15(best_num[1], Syy)) {
                if (mul16_32_q15(num, best_den[0]) > mul16_32_q15(best_num[0], Syy)) {
                    best_num[1] = best_num[0];
                    best_den[1] = best_den[0];
                    best_pitch[1] = best_pitch[0];
                    best_num[0] = num;
                    best_den[0] = Syy;
                    best_pitch[0] = i;
                } else {
                    best_num[1] = num;
                    best_den[1] = Syy;
                    best_pitch[1] = i;
                }
            }
        }
        Syy += (mul16_16(y[i + len], y[i + len]) >> yshift) - (mul16_16(y[i], y[i]) >> yshift);
        Syy = max32(1, Syy);
    }
}

// pitch_downsample (pitch.c:140-213); C channels, x1 the second (or null)
CUO_FN void pitch_downsample(const int32_t *x, const int32_t *x1, int C, int16_t *x_lp, int len, int16_t *xx)
{
    C = kMaxCh == 1 ? 1 : C;
    int32_t maxabs = maxabs32(x, len);
    if (C == 2) maxabs = max32(maxabs, maxabs32(x1, len));
    if (maxabs < 1) maxabs = 1;
    int shift = celt_ilog2(maxabs) - 10;
    if (shift < 0) shift = 0;
    if (C == 2) shift++;
    for (int i = 1; i < len >> 1; i++)
        x_lp[i] = (int16_t)((x[2 * i - 1] >> (shift + 2)) + (x[2 * i + 1] >> (shift + 2)) + (x[2 * i] >> (shift + 1)));
    x_lp[0] = (int16_t)((x[1] >> (shift + 2)) + (x[0] >> (shift + 1)));
    if (C == 2) {
        for (int i = 1; i < len >> 1; i++)
            x_lp[i] = (int16_t)(x_lp[i] + (x1[2 * i - 1] >> (shift + 2)) + (x1[2 * i + 1] >> (shift + 2))
                                + (x1[2 * i] >> (shift + 1)));
        x_lp[0] = (int16_t)(x_lp[0] + (x1[1] >> (shift + 2)) + (x1[0] >> (shift + 1)));
    }
    int32_t ac[5];
    celt_autocorr(x_lp, ac, nullptr, 0, 4, len >> 1, xx);
    ac[0] += ac[0] >> 13;
    for (int i = 1; i <= 4; i++) ac[i] -= mul16_32_q15(2 * i * i, ac[i]);
    int16_t lpc[4], lpc2[5];
    celt_lpc(lpc, ac, 4);
    int16_t tmp = kQ15One;
    for (int i = 0; i < 4; i++) {
        tmp = (int16_t)mul16_16_q15(29491, tmp);
        lpc[i] = (int16_t)mul16_16_q15(lpc[i], tmp);
    }
    const int16_t c1 = 26214;
    lpc2[0] = (int16_t)(lpc[0] + 3277);
    lpc2[1] = (int16_t)(lpc[1] + mul16_16_q15(c1, lpc[0]));
    lpc2[2] = (int16_t)(lpc[2] + mul16_16_q15(c1, lpc[1]));
    lpc2[3] = (int16_t)(lpc[3] + mul16_16_q15(c1, lpc[2]));
    lpc2[4] = (int16_t)mul16_16_q15(c1, lpc[3]);
    // celt_fir5, in place
    int32_t m0 = 0, m1 = 0, m2 = 0, m3 = 0, m4 = 0;
    for (int i = 0; i < len >> 1; i++) {
        int32_t sum = shl32((int32_t)x_lp[i], kSigShift);
        sum = mac16_16(sum, lpc2[0], m0);
        sum = mac16_16(sum, lpc2[1], m1);
        sum = mac16_16(sum, lpc2[2], m2);
        sum = mac16_16(sum, lpc2[3], m3);
        sum = mac16_16(sum, lpc2[4], m4);
        m4 = m3; m3 = m2; m2 = m1; m1 = m0;
        m0 = x_lp[i];
        x_lp[i] = round16(sum, kSigShift);
    }
}

// pitch_search (pitch.c:300-413); scratch: x4 (len/4), y4 (lag/4), xcorr (max_pitch/2)
CUO_FN int pitch_search(const int16_t *x_lp, const int16_t *y, int len, int max_pitch, int16_t *x4, int16_t *y4,
                        int32_t *xcorr)
{
    int lag = len + max_pitch;
    int best_pitch[2] = {0, 0};
    for (int j = 0; j < len >> 2; j++) x4[j] = x_lp[2 * j];
    for (int j = 0; j < lag >> 2; j++) y4[j] = y[2 * j];
    int32_t xmax = maxabs16(x4, len >> 2), ymax = maxabs16(y4, lag >> 2);
    int shift = celt_ilog2(max32(1, max32(xmax, ymax))) - 11;
    if (shift > 0) {
        for (int j = 0; j < len >> 2; j++) x4[j] = (int16_t)(x4[j] >> shift);
        for (int j = 0; j < lag >> 2; j++) y4[j] = (int16_t)(y4[j] >> shift);
        shift *= 2;
    } else {
        shift = 0;
    }
    int32_t maxcorr = 1;
    for (int i = 0; i < max_pitch >> 2; i++) {
        int32_t s = 0;
        for (int j = 0; j < len >> 2; j++) s = mac16_16(s, x4[j], y4[i + j]);
        xcorr[i] = s;
        maxcorr = max32(maxcorr, s);
    }
    find_best_pitch(xcorr, y4, len >> 2, max_pitch >> 2, best_pitch, 0, maxcorr);
    maxcorr = 1;
    for (int i = 0; i < max_pitch >> 1; i++) {
        xcorr[i] = 0;
        int d0 = i - 2 * best_pitch[0], d1 = i - 2 * best_pitch[1];
        if ((d0 < 0 ? -d0 : d0) > 2 && (d1 < 0 ? -d1 : d1) > 2) continue;
        int32_t sum = 0;
        for (int j = 0; j < len >> 1; j++) sum += mul16_16(x_lp[j], y[i + j]) >> shift;
        xcorr[i] = max32(-1, sum);
        maxcorr = max32(maxcorr, sum);
    }
    find_best_pitch(xcorr, y, len >> 1, max_pitch >> 1, best_pitch, shift + 1, maxcorr);
    int offset;
    if (best_pitch[0] > 0 && best_pitch[0] < (max_pitch >> 1) - 1) {
        int32_t a = xcorr[best_pitch[0] - 1], b = xcorr[best_pitch[0]], c = xcorr[best_pitch[0] + 1];
        if ((c - a) > mul16_32_q15(22938, b - a)) offset = 1;
        else if ((a - c) > mul16_32_q15(22938, b - c)) offset = -1;
        else offset = 0;
    } else {
        offset = 0;
    }
    return 2 * best_pitch[0] - offset;
}

// celt_decode_lost, periodic branch (celt_decoder.c:700-955): one pitch
// search over all output channels, then each channel extrapolated with its
// own LPC filter.
CUO_FN int celt_plc_periodic(CeltState *st, int N, CeltScratch *sc)
{
    const int CC = celt_cc(st);
    const int16_t *window = tab::kWindow;
    int16_t fade = kQ15One;
    int pitch_index;
    if (st->loss_duration == 0) {
        int16_t *lp = st->plc->lp_pitch;
        pitch_downsample(st->mem, st->mem + kMemCap, CC, lp, kDecBuf, st->plc->plc_a);
        pitch_index = pitch_search(lp + (kPlcLagMax >> 1), lp, kDecBuf - kPlcLagMax, kPlcLagMax - kPlcLagMin,
                                   st->plc->plc_a, st->plc->plc_b, st->plc->xcorr);
        pitch_index = kPlcLagMax - pitch_index;
        st->last_pitch_index = pitch_index;
    } else {
        pitch_index = st->last_pitch_index;
        fade = 26214;  // QCONST16(.8f, 15)
    }
    const int exc_length = imin(2 * pitch_index, kMaxPeriod);
    for (int c = 0; c < CC; c++) {
        int32_t *buf = st->mem + kMemCap * c;
        int16_t *lpc = st->lpc + kLpcOrder * c;
        int16_t *exc = st->plc->exc + kLpcOrder;
        for (int i = 0; i < kMaxPeriod + kLpcOrder; i++)
            exc[i - kLpcOrder] = sround16(buf[kDecBuf - kMaxPeriod - kLpcOrder + i], kSigShift);
        if (st->loss_duration == 0) {
            int32_t ac[kLpcOrder + 1];
            celt_autocorr(exc, ac, window, kOverlap, kLpcOrder, kMaxPeriod, st->plc->plc_a);
            ac[0] += ac[0] >> 13;
            for (int i = 1; i <= kLpcOrder; i++) ac[i] -= mul16_32_q15(2 * i * i, ac[i]);
            celt_lpc(lpc, ac, kLpcOrder);
            while (1) {
                int16_t tmp = kQ15One;
                int32_t sum = 1 << kSigShift;
                for (int i = 0; i < kLpcOrder; i++) sum += lpc[i] < 0 ? -lpc[i] : lpc[i];
                if (sum < 65535) break;
                for (int i = 0; i < kLpcOrder; i++) {
                    tmp = (int16_t)mul16_16_q15(32440, tmp);
                    lpc[i] = (int16_t)mul16_16_q15(lpc[i], tmp);
                }
            }
        }
        celt_fir(exc + kMaxPeriod - exc_length, lpc, st->plc->fir_tmp, exc_length, kLpcOrder);
        for (int i = 0; i < exc_length; i++) exc[kMaxPeriod - exc_length + i] = st->plc->fir_tmp[i];
        int16_t decay;
        {
            int32_t E1 = 1, E2 = 1;
            int shift = imax(0, 2 * celt_zlog2(maxabs16(&exc[kMaxPeriod - exc_length], exc_length)) - 20);
            int decay_length = exc_length >> 1;
            for (int i = 0; i < decay_length; i++) {
                int16_t e = exc[kMaxPeriod - decay_length + i];
                E1 += mul16_16(e, e) >> shift;
                e = exc[kMaxPeriod - 2 * decay_length + i];
                E2 += mul16_16(e, e) >> shift;
            }
            E1 = min32(E1, E2);
            decay = (int16_t)celt_sqrt(frac_div32(E1 >> 1, E2));
        }
        for (int i = 0; i < kDecBuf - N; i++) buf[i] = buf[i + N];
        const int extrapolation_offset = kMaxPeriod - pitch_index;
        const int extrapolation_len = N + kOverlap;
        int16_t attenuation = (int16_t)mul16_16_q15(fade, decay);
        int32_t S1 = 0;
        for (int i = 0, j = 0; i < extrapolation_len; i++, j++) {
            if (j >= pitch_index) {
                j -= pitch_index;
                attenuation = (int16_t)mul16_16_q15(attenuation, decay);
            }
            buf[kDecBuf - N + i] =
                shl32((int32_t)(int16_t)mul16_16_q15(attenuation, exc[extrapolation_offset + j]), kSigShift);
            int16_t tmp = sround16(buf[kDecBuf - kMaxPeriod - N + extrapolation_offset + j], kSigShift);
            S1 += mul16_16(tmp, tmp) >> 10;
        }
        {
            int16_t lpc_mem[kLpcOrder];
            for (int i = 0; i < kLpcOrder; i++) lpc_mem[i] = sround16(buf[kDecBuf - N - 1 - i], kSigShift);
            celt_iir(buf + kDecBuf - N, lpc, buf + kDecBuf - N, extrapolation_len, kLpcOrder, lpc_mem,
                     st->plc->plc_b);
            for (int i = 0; i < extrapolation_len; i++)
                buf[kDecBuf - N + i] = saturate(buf[kDecBuf - N + i], kSigSat);
        }
        {
            int32_t S2 = 0;
            for (int i = 0; i < extrapolation_len; i++) {
                int16_t tmp = sround16(buf[kDecBuf - N + i], kSigShift);
                S2 += mul16_16(tmp, tmp) >> 10;
            }
            if (!(S1 > (S2 >> 2))) {
                for (int i = 0; i < extrapolation_len; i++) buf[kDecBuf - N + i] = 0;
            } else if (S1 < S2) {
                int16_t ratio = (int16_t)celt_sqrt(frac_div32((S1 >> 1) + 1, S2 + 1));
                for (int i = 0; i < kOverlap; i++) {
                    int16_t tmp_g = (int16_t)(kQ15One - mul16_16_q15(window[i], kQ15One - ratio));
                    buf[kDecBuf - N + i] = mul16_32_q15(tmp_g, buf[kDecBuf - N + i]);
                }
                for (int i = kOverlap; i < extrapolation_len; i++)
                    buf[kDecBuf - N + i] = mul16_32_q15(ratio, buf[kDecBuf - N + i]);
            }
        }
    }
    return 0;
}

}  // namespace CUO_NS
