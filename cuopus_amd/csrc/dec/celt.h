// celt.h -- CELT fixed-point decoder (48 kHz, 960-sample mode; one coded
// channel, or two in the stereo build: kMaxCh == 2).
//
// Restates libopus 1.5.2's celt_decode_with_ec (vendor/celt/celt_decoder.c:
// 970-1369) and its closure: coarse/fine energy (quant_bands.c:428-543,
// laplace.c:88-135), tf/spread/dynalloc/trim (celt_decoder.c:460-497,
// 1206-1250), bit allocation (rate.c:245-646), PVQ band decoding
// (bands.c:700-1672 with quant_partition made iterative; vq.c; cwrs.c:463-541),
// anti-collapse (bands.c:268-370), synthesis (celt_decoder.c:382-458,
// bands.c:196-267, mdct.c:242-342, kiss_fft.c:48-567), comb post-filter
// (celt.c:98-256), deemphasis (celt_decoder.c:279-377) and loss concealment
// (celt_decoder.c:515-969).
//
// Layout: the persistent per-chain state is CeltState; per-frame temporaries
// live in CeltScratch, a plain struct the caller places in HBM (or LDS) so no
// large private arrays are spilled to scratch memory.
#pragma once
#include "entdec.h"
#include "tables_gen.h"

namespace CUO_NS {

constexpr int kDecBuf = 2048;
constexpr int kOverlap = 120;
constexpr int kLpcOrder = 24;
constexpr int kNbBands = 21;
constexpr int kMaxPeriod = 1024;
constexpr int kCombMinPeriod = 15;

enum { SPREAD_NONE = 0, SPREAD_LIGHT = 1, SPREAD_NORMAL = 2, SPREAD_AGGRESSIVE = 3 };

// Decode history of one output channel: kDecBuf + kOverlap samples.
constexpr int kMemStride = kDecBuf + kOverlap;
// Each channel's history window slides forward through a buffer of kMemCap
// samples (advance_history); the kMemSlack samples beyond one window let it
// move on for several frames before the window is copied back to the start.
constexpr int kMemSlack = 3840;
constexpr int kMemCap = kMemStride + kMemSlack;

struct CeltState {
    // decode history: the current window of kMemStride samples per output
    // channel (channel c at c * kMemCap), inside a caller-owned buffer (HBM
    // on the device: it is only touched by lane-parallel loops and
    // concealment)
    int32_t *mem;
    int32_t *mem_base;           // start of the history buffer (kMemCap samples per channel)
    struct CeltPlcScratch *plc;  // pitch-concealment buffers, caller-owned
    int32_t *freq;               // synthesis spectra (kMaxCh x 960), caller-owned: only lane-parallel loops touch them
    int32_t *etmp;               // prefilter_and_fold output (kOverlap), caller-owned
    int32_t start, end;
    int32_t channels;            // output channels (CC), fixed at init
    int32_t stream_channels;     // coded channels of the frame (C), set by the packet layer
    // ---- cleared by reset (OPUS_RESET_STATE) from here on
    uint32_t rng;
    int32_t error;
    int32_t last_pitch_index;
    int32_t loss_duration;
    int32_t skip_plc;
    int32_t pf_period, pf_period_old;
    int32_t pf_gain, pf_gain_old;   // opus_val16 values
    int32_t pf_tapset, pf_tapset_old;
    int32_t prefilter_and_fold;
    int32_t preemph_mem[kMaxCh];
    int16_t lpc[kMaxCh * kLpcOrder];
    int16_t oldBandE[2 * kNbBands];
    int16_t oldLogE[2 * kNbBands];
    int16_t oldLogE2[2 * kNbBands];
    int16_t backgroundLogE[2 * kNbBands];
};

// Buffers of the pitch-based concealment (rare): caller-owned, HBM on the
// device.
struct CeltPlcScratch {
    int16_t exc[kMaxPeriod + kLpcOrder];
    int16_t fir_tmp[kMaxPeriod];
    int16_t lp_pitch[kDecBuf >> 1];
    int32_t xcorr[kMaxPeriod];
    int16_t plc_a[kMaxPeriod + 8];                 // autocorrelation copy / pitch x4
    int16_t plc_b[960 + kOverlap + kLpcOrder];     // IIR history / pitch y4
};

struct CeltScratch {
    // frame decode and noise-based concealment (channel c of X at c * 960;
    // norm holds the folding source of each coded channel, 8 * 78 apart)
    int16_t X[kMaxCh * 960];
    int16_t norm[kMaxCh * 8 * 78];
    int16_t tmp[176];
    int32_t iy[176];
    int32_t pulses[kNbBands], fine_quant[kNbBands], fine_priority[kNbBands];
    int32_t cap[kNbBands], offsets[kNbBands], tf_res[kNbBands];
    int32_t bits1[kNbBands], bits2[kNbBands], thresh[kNbBands], trim_offset[kNbBands];
    uint8_t collapse[kMaxCh * kNbBands];
    RangeDec dec;  // decoder of a frame opened here (redundancy frames)
    // explicit stack of the band splitter (quant_partition), 6 frames
    alignas(8) unsigned char qp_stack[6 * 96];
};

// Output (CC) and coded (C) channel counts; constant 1 in the mono build.
CUO_FN int celt_cc(const CeltState *st) { return kMaxCh == 1 ? 1 : uniform(st->channels); }
CUO_FN int celt_c(const CeltState *st) { return kMaxCh == 1 ? 1 : uniform(st->stream_channels); }

// ------------------------------------------------------------------ reset
CUO_FN void celt_reset(CeltState *st)
{
    st->rng = 0; st->error = 0; st->last_pitch_index = 0; st->loss_duration = 0;
    st->pf_period = st->pf_period_old = 0; st->pf_gain = st->pf_gain_old = 0;
    st->pf_tapset = st->pf_tapset_old = 0; st->prefilter_and_fold = 0;
    for (int c = 0; c < kMaxCh; c++) st->preemph_mem[c] = 0;
    st->mem = st->mem_base;
    int32_t *mem = st->mem;
    for (int c = 0; c < celt_cc(st); c++) CUO_PFOR(i, 0, kMemStride) mem[kMemCap * c + i] = 0;
    for (int i = 0; i < kMaxCh * kLpcOrder; i++) st->lpc[i] = 0;
    for (int i = 0; i < 2 * kNbBands; i++) {
        st->oldBandE[i] = 0;
        st->oldLogE[i] = st->oldLogE2[i] = -(28 << kDbShift);
        st->backgroundLogE[i] = 0;
    }
    st->skip_plc = 1;
}

CUO_FN void celt_init(CeltState *st, int channels)
{
    st->start = 0;
    st->end = kNbBands;
    st->channels = st->stream_channels = kMaxCh == 1 ? 1 : channels;
    celt_reset(st);
}

// ------------------------------------------------------------------ 32-bit table copies
// Small tables read with wave-uniform indices in serial code, widened to
// 32-bit entries so each read is one scalar load.
template <size_t N>
struct I32Table {
    int32_t v[N];
};
template <class T, size_t N>
constexpr I32Table<N> widen(const T (&a)[N])
{
    I32Table<N> t{};
    for (size_t i = 0; i < N; i++) t.v[i] = (int32_t)a[i];
    return t;
}
alignas(16) static constexpr I32Table<22> kEB32 = widen(tab::kEBands);
alignas(16) static constexpr I32Table<21> kLogN32 = widen(tab::kLogN);
alignas(16) static constexpr I32Table<168> kCacheCaps32 = widen(tab::kCacheCaps);
alignas(16) static constexpr I32Table<25> kEMeans32 = widen(tab::kEMeans);
alignas(16) static constexpr I32Table<336> kEProb32 = widen(tab::kEProbModel);
alignas(16) static constexpr I32Table<32> kTfSelect32 = widen(tab::kTfSelect);

// ------------------------------------------------------------------ small helpers
// The history moves N samples on (OPUS_MOVE of the decode memory by one
// frame, celt_decoder.c:661-662): the window pointer advances; when the
// buffer has no room left, the kMemStride - N samples the next window keeps
// are copied to its start.  Lanes copy 64-element blocks in increasing order
// and the source lies N >= 120 samples above the destination, so no block
// reads what an earlier block wrote.  Only the window's last N samples
// differ from what a shift leaves there (stale values either way): the
// synthesis or the concealment overwrites each of them before anything reads
// it, so the decoded output is the same.
CUO_FN int32_t *advance_history(CeltState *st, int N)
{
    int32_t *mem = st->mem;
    int32_t *base = st->mem_base;
    if ((int)(mem - base) + N + kMemStride <= kMemCap) {
        st->mem = mem + N;
        return mem + N;
    }
    for (int c = 0; c < celt_cc(st); c++) {
        int32_t *dst = base + kMemCap * c;
        const int32_t *src = mem + kMemCap * c + N;
        CUO_PFOR(i, 0, kMemStride - N) dst[i] = src[i];
    }
    st->mem = base;
    return base;
}

CUO_FN int32_t inner_prod16(const int16_t *x, const int16_t *y, int n)
{
    int32_t s = 0;
    for (int i = 0; i < n; i++) s = mac16_16(s, x[i], y[i]);
    return s;
}

CUO_FN void renormalise_vector(int16_t *X, int N, int16_t gain)
{
    int32_t part = 0;
    CUO_PFOR(i, 0, N) part = mac16_16(part, X[i], X[i]);
    int32_t E = 1 + wave_sum(part);
    int k = celt_ilog2(E) >> 1;
    int32_t t = vshr32(E, 2 * (k - 7));
    int16_t g = (int16_t)mul16_16_p15(celt_rsqrt_norm(t), gain);
    CUO_PFOR(i, 0, N) X[i] = (int16_t)pshr32(mul16_16(g, X[i]), k + 1);
}

CUO_FN int16_t bitexact_cos(int16_t x)
{
    int32_t tmp = (4096 + (int32_t)x * x) >> 13;
    int16_t x2 = (int16_t)tmp;
    x2 = (int16_t)((32767 - x2) + frac_mul16(x2, (-7651 + frac_mul16(x2, (8277 + frac_mul16(-626, x2))))));
    return (int16_t)(1 + x2);
}

CUO_FN int bitexact_log2tan(int isin, int icos)
{
    int lc = ilog((uint32_t)icos), ls = ilog((uint32_t)isin);
    icos <<= 15 - lc;
    isin <<= 15 - ls;
    return (ls - lc) * (1 << 11) + frac_mul16(isin, frac_mul16(isin, -2597) + 7932)
           - frac_mul16(icos, frac_mul16(icos, -2597) + 7932);
}

// ------------------------------------------------------------------ rate helpers (rate.h)
CUO_FN const uint8_t *pulse_cache(int band, int LMp1)
{
    return tab::kCacheBits + ld_s16(tab::kCacheIndex, LMp1 * kNbBands + band);
}
CUO_FN int get_pulses(int i) { return i < 8 ? i : (8 + (i & 7)) << ((i >> 3) - 1); }
CUO_FN int bits2pulses(int band, int LM, int bits)
{
    const uint8_t *cache = pulse_cache(band, LM + 1);
    int lo = 0, hi = (int)ld_u8(cache, 0);
    bits--;
    for (int i = 0; i < 6; i++) {
        int mid = (lo + hi + 1) >> 1;
        if ((int)ld_u8(cache, mid) >= bits) hi = mid; else lo = mid;
    }
    if (bits - (lo == 0 ? -1 : (int)ld_u8(cache, lo)) <= (int)ld_u8(cache, hi) - bits) return lo;
    return hi;
}
CUO_FN int pulses2bits(int band, int LM, int pulses)
{
    const uint8_t *cache = pulse_cache(band, LM + 1);
    return pulses == 0 ? 0 : (int)ld_u8(cache, pulses) + 1;
}


// ------------------------------------------------------------------ coarse/fine energy
CUO_FI int laplace_decode(RangeDec *d, uint32_t fs, int decay)
{
    int val = 0;
    uint32_t fl = 0;
    uint32_t fm = ec_decode_bin(d, 15);
    if (fm >= fs) {
        val++;
        fl = fs;
        fs = ((32768 - 2 * 16 - fs) * (uint32_t)(int32_t)(16384 - decay) >> 15) + 1;
        while (fs > 1 && fm >= fl + 2 * fs) {
            fs *= 2;
            fl += fs;
            fs = ((fs - 2) * (uint32_t)(int32_t)decay) >> 15;
            fs += 1;
            val++;
        }
        if (fs <= 1) {
            int di = (int)((fm - fl) >> 1);
            val += di;
            fl += 2 * (uint32_t)di;
        }
        if (fm < fl + fs) val = -val; else fl += fs;
    }
    uint32_t fh = fl + fs < 32768 ? fl + fs : 32768;
    ec_update(d, fl, fh, 32768);
    return val;
}

// Energies of coded channel c live at oldE[c * kNbBands + i]; each band's
// channels are coded one after the other (quant_bands.c:428-543).
CUO_FI void unquant_coarse_energy(int start, int end, int16_t *oldE, int intra, RangeDec *d, int LM, int C)
{
    C = kMaxCh == 1 ? 1 : C;
    const int32_t *prob = kEProb32.v + (LM * 2 + intra) * 42;
    int32_t prev0 = 0, prev1 = 0;  // predictor state of each channel (kept in registers)
    int16_t coef, beta;
    if (intra) { coef = 0; beta = tab::kBetaIntra; }
    else { beta = (int16_t)ld_s16(tab::kBetaCoef, LM); coef = (int16_t)ld_s16(tab::kPredCoef, LM); }
    int32_t budget = (int32_t)uniform(d->storage) * 8;
    for (int i = start; i < end; i++) {
        for (int c = 0; c < C; c++) {
            int qi;
            int32_t tell = ec_tell(d);
            if (budget - tell >= 15) {
                int pi = 2 * imin(i, 20);
                qi = laplace_decode(d, (uint32_t)prob[pi] << 7, prob[pi + 1] << 6);
            } else if (budget - tell >= 2) {
                qi = ec_icdf(d, tab::kSmallEnergyIcdf, 2);
                qi = (qi >> 1) ^ -(qi & 1);
            } else if (budget - tell >= 1) {
                qi = -ec_bit_logp(d, 1);
            } else {
                qi = -1;
            }
            int32_t q = shl32(qi, kDbShift);
            int16_t &e = oldE[c * kNbBands + i];
            e = max16(-(9 << kDbShift), e);
            const int32_t prev = c ? prev1 : prev0;
            int32_t tmp = pshr32(mul16_16(coef, e), 8) + prev + shl32(q, 7);
            tmp = max32(-(28 << (kDbShift + 7)), tmp);
            e = (int16_t)pshr32(tmp, 7);
            const int32_t next = prev + shl32(q, 7) - mul16_16(beta, pshr32(q, 8));
            if (c) prev1 = next; else prev0 = next;
        }
    }
}

CUO_FI void unquant_fine_energy(int start, int end, int16_t *oldE, const int32_t *fine_quant, RangeDec *d, int C)
{
    C = kMaxCh == 1 ? 1 : C;
    for (int i = start; i < end; i++) {
        if (fine_quant[i] <= 0) continue;
        for (int c = 0; c < C; c++) {
            int q2 = (int)ec_bits(d, fine_quant[i]);
            int16_t offset =
                (int16_t)(((shl32(q2, kDbShift) + (1 << (kDbShift - 1))) >> fine_quant[i]) - (1 << (kDbShift - 1)));
            oldE[c * kNbBands + i] = (int16_t)(oldE[c * kNbBands + i] + offset);
        }
    }
}

CUO_FI void unquant_energy_finalise(int start, int end, int16_t *oldE, const int32_t *fine_quant,
                                    const int32_t *fine_priority, int bits_left, RangeDec *d, int C)
{
    C = kMaxCh == 1 ? 1 : C;
    for (int prio = 0; prio < 2; prio++) {
        for (int i = start; i < end && bits_left >= C; i++) {
            if (fine_quant[i] >= 8 || fine_priority[i] != prio) continue;
            for (int c = 0; c < C; c++) {
                int q2 = (int)ec_bits(d, 1);
                int16_t offset =
                    (int16_t)((int16_t)(shl16(q2, kDbShift) - (1 << (kDbShift - 1))) >> (fine_quant[i] + 1));
                oldE[c * kNbBands + i] = (int16_t)(oldE[c * kNbBands + i] + offset);
                bits_left--;
            }
        }
    }
}

// ------------------------------------------------------------------ tf / allocation
CUO_FI void tf_decode(int start, int end, int isTransient, int32_t *tf_res, int LM, RangeDec *d)
{
    uint32_t budget = d->storage * 8;
    uint32_t tell = (uint32_t)ec_tell(d);
    int logp = isTransient ? 2 : 4;
    int tf_select_rsv = LM > 0 && tell + logp + 1 <= budget;
    budget -= tf_select_rsv;
    int tf_changed = 0, curr = 0;
    for (int i = start; i < end; i++) {
        if (tell + logp <= budget) {
            curr ^= ec_bit_logp(d, logp);
            tell = (uint32_t)ec_tell(d);
            tf_changed |= curr;
        }
        tf_res[i] = curr;
        logp = isTransient ? 4 : 5;
    }
    int tf_select = 0;
    const int32_t *tsel = kTfSelect32.v + 8 * LM;
    if (tf_select_rsv && tsel[4 * isTransient + 0 + tf_changed] != tsel[4 * isTransient + 2 + tf_changed])
        tf_select = ec_bit_logp(d, 1);
    for (int i = start; i < end; i++) tf_res[i] = tsel[4 * isTransient + 2 * tf_select + tf_res[i]];
}

// clt_compute_allocation + interp_bits2pulses, decoder (rate.c:245-646).
// For two coded channels it also decodes the intensity band and the dual
// stereo flag.
CUO_FI int compute_allocation(int start, int end, const int32_t *offsets, const int32_t *cap, int alloc_trim,
                              int32_t total, int32_t *balance_out, int32_t *bits, int32_t *ebits,
                              int32_t *fine_priority, int LM, RangeDec *d, CeltScratch *sc, int Cin,
                              int *intensity, int *dual_stereo)
{
    const int32_t *eB = kEB32.v;
    int32_t *bits1 = sc->bits1, *bits2 = sc->bits2, *thresh = sc->thresh, *trim_offset = sc->trim_offset;
    const int C = kMaxCh == 1 ? 1 : Cin;
    const int stereo = C > 1;
    total = max32(total, 0);
    int skip_start = start;
    int skip_rsv = total >= 1 << kBitRes ? 1 << kBitRes : 0;
    total -= skip_rsv;
    int intensity_rsv = 0, dual_stereo_rsv = 0;
    if (C == 2) {
        intensity_rsv = ld_u8(tab::kLog2FracTable, end - start);
        if (intensity_rsv > total) {
            intensity_rsv = 0;
        } else {
            total -= intensity_rsv;
            dual_stereo_rsv = total >= 1 << kBitRes ? 1 << kBitRes : 0;
            total -= dual_stereo_rsv;
        }
    }
    CUO_PFOR(j, start, end) {
        thresh[j] = imax(C << kBitRes, (3 * (eB[j + 1] - eB[j]) << LM << kBitRes) >> 4);
        int t = C * (eB[j + 1] - eB[j]) * (alloc_trim - 5 - LM) * (end - j - 1) * (1 << (LM + kBitRes)) >> 6;
        if ((eB[j + 1] - eB[j]) << LM == 1) t -= C << kBitRes;
        trim_offset[j] = t;
    }
    // The band loops below run from the top band down with a `done` flag that
    // latches at the first band (from the top) whose bits reach its threshold:
    // every band at or below the highest such band counts min(bits, cap), every
    // band above it counts the floor or nothing.  The highest flagged band is a
    // max over lanes and the sum a sum over lanes, so each loop is one pass.
    auto vec_bits = [&](int v, int j) -> int {
        int bj = C * (eB[j + 1] - eB[j]) * tab::kAllocVectors[v * kNbBands + j] << LM >> 2;
        if (bj > 0) bj = imax(0, bj + trim_offset[j]);
        return bj + offsets[j];
    };
    int lo = 1, hi = tab::kNbAllocVectors - 1;
    CUO_PROF_SCOPE(kProfAllocBisect);
    do {
        int mid = (lo + hi) >> 1;
        int jtop = -1;
        CUO_PFOR(j, start, end) if (vec_bits(mid, j) >= thresh[j]) jtop = imax(jtop, j);
        jtop = wave_max(jtop);
        int32_t part = 0;
        CUO_PFOR(j, start, end) {
            int bitsj = vec_bits(mid, j);
            if (j <= jtop) part += imin(bitsj, cap[j]);
            else if (bitsj >= C << kBitRes) part += C << kBitRes;
        }
        int32_t psum = wave_sum(part);
        if (psum > total) hi = mid - 1; else lo = mid + 1;
    } while (lo <= hi);
    hi = lo--;
    int skip_top = -1;
    CUO_PFOR(j, start, end) {
        int N = eB[j + 1] - eB[j];
        int bits1j = C * N * tab::kAllocVectors[lo * kNbBands + j] << LM >> 2;
        int bits2j = hi >= tab::kNbAllocVectors ? cap[j] : C * N * tab::kAllocVectors[hi * kNbBands + j] << LM >> 2;
        if (bits1j > 0) bits1j = imax(0, bits1j + trim_offset[j]);
        if (bits2j > 0) bits2j = imax(0, bits2j + trim_offset[j]);
        if (lo > 0) bits1j += offsets[j];
        bits2j += offsets[j];
        if (offsets[j] > 0) skip_top = imax(skip_top, j);
        bits2j = imax(0, bits2j - bits1j);
        bits1[j] = bits1j;
        bits2[j] = bits2j;
    }
    skip_top = wave_max(skip_top);
    if (skip_top >= 0) skip_start = skip_top;
    // ---- interp_bits2pulses
    const int alloc_floor = C << kBitRes;
    const int logM = LM << kBitRes;
    lo = 0; hi = 1 << 6;
    for (int i = 0; i < 6; i++) {
        int mid = (lo + hi) >> 1;
        int jtop = -1;
        CUO_PFOR(j, start, end) if (bits1[j] + (mid * (int32_t)bits2[j] >> 6) >= thresh[j]) jtop = imax(jtop, j);
        jtop = wave_max(jtop);
        int32_t part = 0;
        CUO_PFOR(j, start, end) {
            int tmp = bits1[j] + (mid * (int32_t)bits2[j] >> 6);
            if (j <= jtop) part += imin(tmp, cap[j]);
            else if (tmp >= alloc_floor) part += alloc_floor;
        }
        int32_t psum = wave_sum(part);
        if (psum > total) hi = mid; else lo = mid;
    }
    int32_t psum;
    {
        int jtop = -1;
        CUO_PFOR(j, start, end) if (bits1[j] + ((int32_t)lo * bits2[j] >> 6) >= thresh[j]) jtop = imax(jtop, j);
        jtop = wave_max(jtop);
        int32_t part = 0;
        CUO_PFOR(j, start, end) {
            int tmp = bits1[j] + ((int32_t)lo * bits2[j] >> 6);
            if (j > jtop) tmp = tmp >= alloc_floor ? alloc_floor : 0;
            tmp = imin(tmp, cap[j]);
            bits[j] = tmp;
            part += tmp;
        }
        psum = wave_sum(part);
    }
    int codedBands;
    for (codedBands = end;; codedBands--) {
        int j = codedBands - 1;
        if (j <= skip_start) { total += skip_rsv; break; }
        int32_t left = total - psum;
        int32_t percoeff = left / (eB[codedBands] - eB[start]);
        left -= (eB[codedBands] - eB[start]) * percoeff;
        int rem = imax(left - (eB[j] - eB[start]), 0);
        int band_width = eB[codedBands] - eB[j];
        int band_bits = (int)(bits[j] + percoeff * band_width + rem);
        if (band_bits >= imax(thresh[j], alloc_floor + (1 << kBitRes))) {
            if (ec_bit_logp(d, 1)) break;
            psum += 1 << kBitRes;
            band_bits -= 1 << kBitRes;
        }
        psum -= bits[j] + intensity_rsv;
        if (intensity_rsv > 0) intensity_rsv = ld_u8(tab::kLog2FracTable, j - start);
        psum += intensity_rsv;
        if (band_bits >= alloc_floor) { psum += alloc_floor; bits[j] = alloc_floor; }
        else bits[j] = 0;
    }
    *intensity = intensity_rsv > 0 ? start + (int)ec_uint(d, codedBands + 1 - start) : 0;
    if (*intensity <= start) {
        total += dual_stereo_rsv;
        dual_stereo_rsv = 0;
    }
    *dual_stereo = dual_stereo_rsv > 0 ? ec_bit_logp(d, 1) : 0;
    int32_t left = total - psum;
    int32_t percoeff = left / (eB[codedBands] - eB[start]);
    left -= (eB[codedBands] - eB[start]) * percoeff;
    // the serial remainder hand-out gives band j min(max(left - (eB[j] -
    // eB[start]), 0), width_j) when left >= 0, and all of a negative left to
    // the first band
    CUO_PFOR(j, start, codedBands) {
        int w = eB[j + 1] - eB[j];
        int share = left >= 0 ? imin(imax(left - (eB[j] - eB[start]), 0), w) : (j == start ? left : 0);
        bits[j] += (int)percoeff * w + share;
    }
    int32_t balance = 0;
    int j;
    for (j = start; j < codedBands; j++) {
        int N0 = eB[j + 1] - eB[j];
        int N = N0 << LM;
        int32_t bit = (int32_t)bits[j] + balance;
        int32_t excess;
        if (N > 1) {
            excess = max32(bit - cap[j], 0);
            bits[j] = bit - excess;
            int den = C * N + (C == 2 && N > 2 && !*dual_stereo && j < *intensity ? 1 : 0);
            int NClogN = den * (kLogN32.v[j] + logM);
            int offset = (NClogN >> 1) - den * 21;
            if (N == 2) offset += den << kBitRes >> 2;
            if (bits[j] + offset < den * 2 << kBitRes) offset += NClogN >> 2;
            else if (bits[j] + offset < den * 3 << kBitRes) offset += NClogN >> 3;
            ebits[j] = imax(0, bits[j] + offset + (den << (kBitRes - 1)));
            ebits[j] = (ebits[j] / den) >> kBitRes;
            if (C * ebits[j] > (bits[j] >> kBitRes)) ebits[j] = bits[j] >> stereo >> kBitRes;
            ebits[j] = imin(ebits[j], 8);
            fine_priority[j] = ebits[j] * (den << kBitRes) >= bits[j] + offset;
            bits[j] -= C * ebits[j] << kBitRes;
        } else {
            excess = max32(0, bit - (C << kBitRes));
            bits[j] = bit - excess;
            ebits[j] = 0;
            fine_priority[j] = 1;
        }
        if (excess > 0) {
            int extra_fine = imin(excess >> (stereo + kBitRes), 8 - ebits[j]);
            ebits[j] += extra_fine;
            int extra_bits = extra_fine * C << kBitRes;
            fine_priority[j] = extra_bits >= excess - balance;
            excess -= extra_bits;
        }
        balance = excess;
    }
    *balance_out = balance;
    for (; j < end; j++) {
        ebits[j] = bits[j] >> stereo >> kBitRes;
        bits[j] = 0;
        fine_priority[j] = ebits[j] < 1;
    }
    return codedBands;
}

// ------------------------------------------------------------------ PVQ
CUO_FN uint32_t pvq_u(int n, int k)
{
    int r = imin(n, k), c = imax(n, k);
    return tab::kPvqU[ld_s16(tab::kPvqURow, r) + c];
}
CUO_FN uint32_t pvq_v(int n, int k) { return pvq_u(n, k) + pvq_u(n, k + 1); }

// cwrsi: codeword index -> pulse vector, returns sum of squares (cwrs.c:463-537)
CUO_FN int32_t cwrsi(int n, int k, uint32_t i, int32_t *y)
{
    n = uniform(n);
    k = uniform(k);
    i = uniform(i);
    int32_t yy = 0;
    uint32_t p;
    int s, k0;
    int16_t val;
    while (n > 2) {
        uint32_t q;
        if (k >= n) {
            const uint32_t *row = tab::kPvqU + ld_s16(tab::kPvqURow, n);
            p = row[k + 1];
            s = -(i >= p);
            i -= p & (uint32_t)s;
            k0 = k;
            q = row[n];
            if (q > i) {
                k = n;
                do p = tab::kPvqU[ld_s16(tab::kPvqURow, --k) + n]; while (p > i);
            } else {
                for (p = row[k]; p > i; p = row[k]) k--;
            }
            i -= p;
            val = (int16_t)((k0 - k + s) ^ s);
            *y++ = val;
            yy = mac16_16(yy, val, val);
        } else {
            p = tab::kPvqU[ld_s16(tab::kPvqURow, k) + n];
            q = tab::kPvqU[ld_s16(tab::kPvqURow, k + 1) + n];
            if (p <= i && i < q) {
                i -= p;
                *y++ = 0;
            } else {
                s = -(i >= q);
                i -= q & (uint32_t)s;
                k0 = k;
                do p = tab::kPvqU[ld_s16(tab::kPvqURow, --k) + n]; while (p > i);
                i -= p;
                val = (int16_t)((k0 - k + s) ^ s);
                *y++ = val;
                yy = mac16_16(yy, val, val);
            }
        }
        n--;
    }
    p = 2 * (uint32_t)k + 1;
    s = -(i >= p);
    i -= p & (uint32_t)s;
    k0 = k;
    k = (int)((i + 1) >> 1);
    if (k) i -= 2 * (uint32_t)k - 1;
    val = (int16_t)((k0 - k + s) ^ s);
    *y++ = val;
    yy = mac16_16(yy, val, val);
    s = -(int)i;
    val = (int16_t)((k + s) ^ s);
    *y = val;
    yy = mac16_16(yy, val, val);
    return yy;
}

#if CUO_LANE_PATHS
// Lane-parallel PVQ codeword decode (device).  Lane j holds U(n, j) for the
// current dimension n: each row is one vector gather from the U table
// (U(j, n) = U(n, j); the row of min(j, n) holds it), issued one dimension
// ahead of its use since the rows do not depend on the decoded values (a
// gather takes fewer instructions than deriving the row from the one above
// with a wave prefix sum).  The table walks of cwrs.c:463-537 --
// "largest k' below a bound with U(n, k') <= i" -- become one ballot over the
// lanes: U(n, .) is non-decreasing, so the lanes that pass form a prefix and
// the highest passing lane is the walk's result.  Needs k + 1 < 64; lanes
// beyond the ones a step reads are masked out of every ballot, and the
// gather index is clamped into the table.
using PvqTable = const __attribute__((address_space(1))) uint32_t *;
__device__ __forceinline__ uint32_t pvq_row_lanes(PvqTable U, int j, int n, int row_lane, int row_n)
{
    int idx = j < n ? row_lane + n : row_n + j;
    idx = idx < 0 ? 0 : (idx > 1271 ? 1271 : idx);
    return U[idx];
}
// The U table's global address in scalar registers, held there across a
// loop (rather than re-formed from the program counter at every gather).
__device__ __forceinline__ PvqTable pvq_u_table()
{
    uint64_t a = (uint64_t)(uintptr_t)tab::kPvqU;
    asm("" : "+s"(a));  // opaque: kept in its registers, not re-formed
    return (PvqTable)a;
}
__device__ __forceinline__ int pvq_row_off(int r) { return tab::kPvqURow[r < 14 ? r : 14]; }

// Inclusive prefix sum over all 64 lanes (wrapping): a 16-lane DPP scan per
// row plus the totals of the rows below, read from lanes 15, 31 and 47.
__device__ __forceinline__ uint32_t wave_incl_sum(uint32_t x)
{
    const int l = lane_id();
    const uint32_t v = (uint32_t)row16_prefix_sum((int32_t)x);
    const uint32_t t0 = (uint32_t)lane_get((int32_t)v, LaneK<15>());
    const uint32_t t1 = (uint32_t)lane_get((int32_t)v, LaneK<31>());
    const uint32_t t2 = (uint32_t)lane_get((int32_t)v, LaneK<47>());
    const uint32_t below = l < 16 ? 0u : (l < 32 ? t0 : (l < 48 ? t0 + t1 : t0 + t1 + t2));
    return v + below;
}
__device__ __forceinline__ int highest_lane(uint64_t m) { return m ? 63 - __builtin_clzll(m) : 0; }

// Runs of zeros (cwrs.c:513-520 taken for several dimensions in a row): with
// k < n, dimension d decodes to 0 while U(d, k) <= i_d < U(d, k + 1), where
// i_d is i less the U(d', k) of the dimensions d' > d before it.  Lane m takes
// dimension n - m: it gathers U(n - m, k) and U(n - m, k + 1) (row k and k + 1
// of the table, column n - m > k), an exclusive prefix sum gives its i, and
// the lowest lane that fails the test ends the run.  Lanes are limited to
// dimensions above max(k, 2), which the row-by-row steps handle.  The sums are
// exact up to the first failing lane (each is at most i there).  Writes the
// zeros, lowers i and returns the run length (>= 1 when dimension n is a zero).
__device__ __forceinline__ int pvq_zero_run(int n, int k, uint32_t &i, int32_t *y)
{
    const int m = lane_id();
    const int d = n - m;
    const int lim = k > 2 ? k : 2;
    const bool valid = d > lim;
    const int ia = valid ? ld_s16(tab::kPvqURow, k) + d : 0;
    const int ib = valid ? ld_s16(tab::kPvqURow, k + 1) + d : 0;
    const uint32_t a = valid ? tab::kPvqU[ia] : 0u;
    const uint32_t b = tab::kPvqU[ib];
    const uint32_t incl = wave_incl_sum(a);
    const uint32_t im = i - (incl - a);
    const uint64_t fail = lanes_s<kCmpSgt>(d, lim) & (lanes_u<kCmpUgt>(a, im) | lanes_u<kCmpUge>(im, b));
    const int run = fail ? __builtin_ctzll(fail) : imin(n - lim, 64);
    i -= (uint32_t)lane_get((int32_t)incl, run - 1);
    if (m < run) y[m] = 0;
    return run;
}
// The run step is taken when n - k is at least this; shorter stretches are
// cheaper row by row.  k + 1 <= 14 there: the allocation only yields (N, K)
// with V(N, K) < 2^32, and k < n >= 15 with k >= 14 would exceed it.
constexpr int kPvqRunMin = 8;

// cwrsi with the row of the first dimension already gathered (row = U(n, .);
// row_lane = the table offset of row min(lane, 14))
__device__ __noinline__ int32_t cwrsi_lanes(int n, int k, uint32_t i, int32_t *y, uint32_t row, int row_lane)
{
    n = uniform(n);
    k = uniform(k);
    i = uniform(i);
    y = lds_uniform(y);
    const int l = lane_id();
    const PvqTable U = pvq_u_table();
    int32_t yy = 0;
    while (n > 2) {
        // the next dimension's row, consumed one step later
        const uint32_t next = n > 3 ? pvq_row_lanes(U, l, n - 1, row_lane, lane_get(row_lane, imin(n - 1, 14))) : 0u;
        uint32_t p;
        int s, k0;
        int16_t val;
        if (k >= n) {
            p = (uint32_t)lane_get((int32_t)row, k + 1);
            s = -(i >= p);
            i -= p & (uint32_t)s;
            k0 = k;
            const uint32_t q = (uint32_t)lane_get((int32_t)row, n);
            const int bound = q > i ? n - 1 : k;  // largest k' <= bound with U(n, k') <= i
            k = highest_lane(lanes_s<kCmpSle>(l, bound) & lanes_u<kCmpUle>(row, i));
            p = (uint32_t)lane_get((int32_t)row, k);
            i -= p;
            val = (int16_t)((k0 - k + s) ^ s);
            *y++ = val;
            yy = mac16_16(yy, val, val);
        } else {
            p = (uint32_t)lane_get((int32_t)row, k);
            const uint32_t q = (uint32_t)lane_get((int32_t)row, k + 1);
            if (p <= i && i < q) {
                if (n - k >= kPvqRunMin) {
                    const int run = pvq_zero_run(n, k, i, y);
                    y += run;
                    n -= run;
                    if (n > 2) row = pvq_row_lanes(U, l, n, row_lane, lane_get(row_lane, imin(n, 14)));
                    continue;
                }
                i -= p;
                *y++ = 0;
            } else {
                s = -(i >= q);
                i -= q & (uint32_t)s;
                k0 = k;
                k = highest_lane(lanes_s<kCmpSlt>(l, k) & lanes_u<kCmpUle>(row, i));
                p = (uint32_t)lane_get((int32_t)row, k);
                i -= p;
                val = (int16_t)((k0 - k + s) ^ s);
                *y++ = val;
                yy = mac16_16(yy, val, val);
            }
        }
        n--;
        row = next;
    }
    uint32_t p = 2 * (uint32_t)k + 1;
    int s = -(i >= p);
    i -= p & (uint32_t)s;
    int k0 = k;
    k = (int)((i + 1) >> 1);
    if (k) i -= 2 * (uint32_t)k - 1;
    int16_t val = (int16_t)((k0 - k + s) ^ s);
    *y++ = val;
    yy = mac16_16(yy, val, val);
    s = -(int)i;
    val = (int16_t)((k + s) ^ s);
    *y = val;
    yy = mac16_16(yy, val, val);
    return yy;
}
#endif

// exp_rotation1 (vq.c:74-95): a forward then a backward pass of 2x2
// rotations between X[i] and X[i+stride].  Residue classes mod stride are
// independent chains; stride 1 is one serial chain, run register-blocked.
CUO_FN void exp_rotation1(int16_t *X, int len, int stride, int16_t c, int16_t s)
{
    c = uniform(c);
    s = uniform(s);
    len = uniform(len);
    const int16_t ms = (int16_t)-s;
    if (stride == 1) {
        // Only the carried value is serial: step i's carry is
        // (c*X[i+1] + s*carry + 2^14) >> 15, so c*X[i+1] + 2^14 is formed for
        // the whole block first, the serial pass records the carry entering
        // each step in that step's lane, and the block's outputs are then
        // formed lane-parallel from it.
        const int l = lane_id();
        // forward: carry = X[i] as updated by step i-1
        if (len >= 2) {
            int32_t carry = uniform((int32_t)X[0]);
            for (int i0 = 0; i0 < len - 1; i0 += kLanes) {
                const int nb = uniform(imin(kLanes, len - 1 - i0));
                const int32_t x2 = l < nb ? X[i0 + l + 1] : 0;
                const int32_t p = mul16_16(c, x2) + 16384;
                int32_t x1 = 0;
                serial_lanes(nb, [&](auto k) {
                    x1 = lane_set_s(x1, k, carry);
                    carry = shr15_to16(lane_get(p, k) + mul16_16(s, carry));
                });
                if (l < nb) X[i0 + l] = (int16_t)pshr32(mac16_16(mul16_16(c, x1), ms, x2), 15);
            }
            X[len - 1] = (int16_t)carry;
        }
        // backward: carry = X[i+1] as updated by step i+1
        if (len >= 3) {
            int32_t carry = uniform((int32_t)X[len - 2]);
            for (int ih = len - 3; ih >= 0; ih -= kLanes) {
                const int nb = uniform(imin(kLanes, ih + 1));
                const int32_t x1 = l < nb ? X[ih - l] : 0;
                const int32_t q = mul16_16(c, x1) + 16384;
                int32_t x2 = 0;
                serial_lanes(nb, [&](auto k) {
                    x2 = lane_set_s(x2, k, carry);
                    carry = shr15_to16(lane_get(q, k) + mul16_16(ms, carry));
                });
                if (l < nb) X[ih - l + 1] = (int16_t)pshr32(mac16_16(mul16_16(c, x2), s, x1), 15);
            }
            X[0] = (int16_t)carry;
        }
        return;
    }
    CUO_PFOR(r, 0, stride) {
        for (int i = r; i < len - stride; i += stride) {
            int16_t x1 = X[i], x2 = X[i + stride];
            X[i + stride] = (int16_t)pshr32(mac16_16(mul16_16(c, x2), s, x1), 15);
            X[i] = (int16_t)pshr32(mac16_16(mul16_16(c, x1), ms, x2), 15);
        }
    }
    const int top = len - 2 * stride - 1;
    CUO_PFOR(r, 0, stride) {
        if (top >= r) {
            for (int i = top - (top - r) % stride; i >= 0; i -= stride) {
                int16_t x1 = X[i], x2 = X[i + stride];
                X[i + stride] = (int16_t)pshr32(mac16_16(mul16_16(c, x2), s, x1), 15);
                X[i] = (int16_t)pshr32(mac16_16(mul16_16(c, x1), ms, x2), 15);
            }
        }
    }
}

CUO_FN void exp_rotation(int16_t *X, int len, int dir, int stride, int K, int spread)
{
    if (2 * K >= len || spread == SPREAD_NONE) return;
    const int factor = spread == 1 ? 15 : (spread == 2 ? 10 : 5);
    int16_t gain = (int16_t)celt_div((int32_t)mul16_16(kQ15One, len), (int32_t)(len + factor * K));
    int16_t theta = (int16_t)(mul16_16_q15(gain, gain) >> 1);
    int16_t c = celt_cos_norm((int32_t)theta);
    int16_t s = celt_cos_norm((int32_t)(int16_t)sub16(kQ15One, theta));
    int stride2 = 0;
    if (len >= 8 * stride) {
        stride2 = 1;
        while ((stride2 * stride2 + stride2) * stride + (stride >> 2) < len) stride2++;
    }
    len = len / stride;
    for (int i = 0; i < stride; i++) {
        if (dir < 0) {
            if (stride2) exp_rotation1(X + i * len, len, stride2, s, c);
            exp_rotation1(X + i * len, len, 1, c, s);
        } else {
            exp_rotation1(X + i * len, len, 1, c, (int16_t)-s);
            if (stride2) exp_rotation1(X + i * len, len, stride2, s, (int16_t)-c);
        }
    }
}

CUO_FI unsigned alg_unquant(int16_t *X, int N, int K, int spread, int B, RangeDec *d, int16_t gain, int32_t *iy)
{
    CUO_PROF_SCOPE(kProfPvq);
    uint32_t idx;
    int32_t Ryy;
#if CUO_LANE_PATHS
    N = uniform(N);
    K = uniform(K);
    if (K + 1 < 64 && N > 2) {
        // V(N, K) = U(N, K) + U(N, K + 1) from the first dimension's row
        const int row_lane = pvq_row_off(lane_id());
        const uint32_t row = pvq_row_lanes(pvq_u_table(), lane_id(), N, row_lane, lane_get(row_lane, imin(N, 14)));
        const uint32_t v = (uint32_t)lane_get((int32_t)row, K) + (uint32_t)lane_get((int32_t)row, K + 1);
        {
            CUO_PROF_SCOPE(kProfEcUint);
            idx = ec_uint(d, v);
        }
        CUO_PROF_SCOPE(kProfCwrsi);
        Ryy = cwrsi_lanes(N, K, idx, iy, row, row_lane);
    } else
#endif
    {
        {
            CUO_PROF_SCOPE(kProfEcUint);
            idx = ec_uint(d, pvq_v(N, K));
        }
        CUO_PROF_SCOPE(kProfCwrsi);
        Ryy = cwrsi(N, K, idx, iy);
    }
    // normalise_residual
    int k = celt_ilog2(Ryy) >> 1;
    int32_t t = vshr32(Ryy, 2 * (k - 7));
    int16_t g = (int16_t)mul16_16_p15(celt_rsqrt_norm(t), gain);
    CUO_PFOR(i, 0, N) X[i] = (int16_t)pshr32(mul16_16(g, iy[i]), k + 1);
    {
        CUO_PROF_SCOPE(kProfExpRot);
        exp_rotation(X, N, -1, B, K, spread);
    }
    // extract_collapse_mask
    if (B <= 1) return 1;
    int N0 = N / B;
    unsigned cm = 0;
    CUO_PFOR(j, 0, N) if (iy[j]) cm |= 1u << (j / N0);
    return wave_or(cm);
}

// haar1 and the (de)interleavers (bands.c:560-640): every (i, j) step reads
// and writes its own elements, so the steps are spread over the lanes; the
// interleavers gather into tmp in one pass and copy back in a second (LDS
// accesses of a wavefront complete in program order).
CUO_FN void haar1(int16_t *X, int N0, int stride)
{
    N0 >>= 1;
    CUO_PFOR(ij, 0, stride * N0) {
        const int i = ij / N0, j = ij - i * N0;
        int32_t t1 = mul16_16(23170, X[stride * 2 * j + i]);
        int32_t t2 = mul16_16(23170, X[stride * (2 * j + 1) + i]);
        X[stride * 2 * j + i] = (int16_t)pshr32(t1 + t2, 15);
        X[stride * (2 * j + 1) + i] = (int16_t)pshr32(t1 - t2, 15);
    }
}

CUO_FN void deinterleave_hadamard(int16_t *X, int N0, int stride, int hadamard, int16_t *tmp)
{
    int N = N0 * stride;
    const int16_t *ordery = hadamard ? tab::kOrderyTable + stride - 2 : nullptr;
    CUO_PFOR(ij, 0, N) {
        const int i = ij / N0, j = ij - i * N0;
        tmp[(hadamard ? ordery[i] : i) * N0 + j] = X[j * stride + i];
    }
    CUO_PFOR(i, 0, N) X[i] = tmp[i];
}

CUO_FN void interleave_hadamard(int16_t *X, int N0, int stride, int hadamard, int16_t *tmp)
{
    int N = N0 * stride;
    const int16_t *ordery = hadamard ? tab::kOrderyTable + stride - 2 : nullptr;
    CUO_PFOR(ij, 0, N) {
        const int i = ij / N0, j = ij - i * N0;
        tmp[j * stride + i] = X[(hadamard ? ordery[i] : i) * N0 + j];
    }
    CUO_PFOR(i, 0, N) X[i] = tmp[i];
}

struct BandCtx {
    unsigned char *qp_stack;
    RangeDec *ec;
    int band;
    int spread;
    int tf_change;
    int32_t remaining_bits;
    uint32_t seed;
    int32_t *iy;
    int intensity;    // first intensity-stereo band (two coded channels)
    int disable_inv;  // phase inversion off: a stereo stream decoded to one channel
};

static constexpr int16_t kExp2Table8[8] = {16384, 17866, 19483, 21247, 23170, 25267, 27554, 30048};

CUO_FN int compute_qn(int N, int b, int offset, int pulse_cap, int stereo)
{
    int N2 = 2 * N - 1;
    if (stereo && N == 2) N2--;
    int qb = (b + N2 * offset) / N2;
    qb = imin(b - pulse_cap - (4 << kBitRes), qb);
    qb = imin(8 << kBitRes, qb);
    if (qb < (1 << kBitRes >> 1)) return 1;
    int qn = kExp2Table8[qb & 0x7] >> (14 - (qb >> kBitRes));
    return (qn + 1) >> 1 << 1;
}

// One level of the mono theta split (bands.c:700-903 with stereo = 0).
struct SplitOut { int imid, iside, delta, itheta, qalloc; };

CUO_FI SplitOut compute_theta_mono(BandCtx *ctx, int N, int *b, int B, int B0, int LM, int *fill)
{
    SplitOut o;
    int i = ctx->band;
    int pulse_cap = kLogN32.v[i] + LM * (1 << kBitRes);
    int offset = (pulse_cap >> 1) - 4;
    int qn = compute_qn(N, *b, offset, pulse_cap, 0);
    int32_t tell = (int32_t)ec_tell_frac(ctx->ec);
    int itheta = 0;
    if (qn != 1) {
        if (B0 > 1) {
            itheta = (int)ec_uint(ctx->ec, qn + 1);
        } else {
            int fs = 1, fl = 0;
            int ft = ((qn >> 1) + 1) * ((qn >> 1) + 1);
            int fm = (int)ec_decode(ctx->ec, ft);
            if (fm < ((qn >> 1) * ((qn >> 1) + 1) >> 1)) {
                itheta = (int)(isqrt32(8 * (uint32_t)fm + 1) - 1) >> 1;
                fs = itheta + 1;
                fl = itheta * (itheta + 1) >> 1;
            } else {
                itheta = (2 * (qn + 1) - (int)isqrt32(8 * (uint32_t)(ft - fm - 1) + 1)) >> 1;
                fs = qn + 1 - itheta;
                fl = ft - ((qn + 1 - itheta) * (qn + 2 - itheta) >> 1);
            }
            ec_update(ctx->ec, fl, fl + fs, ft);
        }
        itheta = (int)((int32_t)itheta * 16384 / qn);
    }
    int qalloc = (int32_t)ec_tell_frac(ctx->ec) - tell;
    *b -= qalloc;
    if (itheta == 0) {
        o.imid = 32767; o.iside = 0; *fill &= (1 << B) - 1; o.delta = -16384;
    } else if (itheta == 16384) {
        o.imid = 0; o.iside = 32767; *fill &= ((1 << B) - 1) << B; o.delta = 16384;
    } else {
        o.imid = bitexact_cos((int16_t)itheta);
        o.iside = bitexact_cos((int16_t)(16384 - itheta));
        o.delta = frac_mul16((N - 1) << 7, bitexact_log2tan(o.iside, o.imid));
    }
    o.itheta = itheta;
    o.qalloc = qalloc;
    return o;
}

// quant_partition (decoder, mono) as an explicit-stack iteration: each frame
// holds the arguments of one recursive call plus what the call keeps between
// its two children (bands.c:943-1108).
struct QpFrame {
    int16_t *X, *lowband, *Y, *lowband2;
    int N, b, B, B0, LM, fill;
    int16_t gain, mid, side;
    int stage, mfirst, mbits, sbits, itheta, rebalance;
    unsigned cm;
};
static_assert(sizeof(QpFrame) <= 96, "QpFrame exceeds its stack slot");

// Leaf of the band splitter: PVQ decode of the pulses the bits buy, or
// folding / noise when there are none (bands.c:1046-1105).
CUO_FI unsigned qp_leaf(BandCtx *ctx, int16_t *fX, int N, int b, int B, int16_t *flow, int LM, int16_t gain,
                        int fill_in)
{
    int q;
    {
        CUO_PROF_SCOPE(kProfQpLeafBits);
#if CUO_LANE_PATHS
        // The cache row (at most 40 entries after its count) one entry per
        // lane, read once: bits2pulses' six-step bisection ends at the lowest
        // entry index >= bits (its count when there is none) and the one
        // below it -- the same pair for every row and every bits value
        // (checked over all rows and bit counts) -- and pulses2bits is a
        // lane read.
        const uint8_t *cache = pulse_cache(ctx->band, LM + 1);
        const int mx = (int)ld_u8(cache, 0);
        const int j = lane_id();
        const int cj = (int)cache[j <= mx ? j : mx];
        const int bits = b - 1;
        const uint64_t m = lanes_s<kCmpSge>(cj, bits) & lanes_s<kCmpSge>(j, 1) & lanes_s<kCmpSle>(j, mx);
        const int hi = m ? __builtin_ctzll(m) : mx;
        const int lo = hi - 1;
        q = bits - (lo == 0 ? -1 : lane_get(cj, lo)) <= lane_get(cj, hi) - bits ? lo : hi;
        int curr_bits = q ? lane_get(cj, q) + 1 : 0;
        ctx->remaining_bits -= curr_bits;
        while (ctx->remaining_bits < 0 && q > 0) {
            ctx->remaining_bits += curr_bits;
            q--;
            curr_bits = q ? lane_get(cj, q) + 1 : 0;
            ctx->remaining_bits -= curr_bits;
        }
#else
        q = bits2pulses(ctx->band, LM, b);
        int curr_bits = pulses2bits(ctx->band, LM, q);
        ctx->remaining_bits -= curr_bits;
        while (ctx->remaining_bits < 0 && q > 0) {
            ctx->remaining_bits += curr_bits;
            q--;
            curr_bits = pulses2bits(ctx->band, LM, q);
            ctx->remaining_bits -= curr_bits;
        }
#endif
    }
    unsigned cm = 0;
    if (q != 0) {
        cm = alg_unquant(fX, N, get_pulses(q), ctx->spread, B, ctx->ec, gain, ctx->iy);
    } else {
        CUO_PROF_SCOPE(kProfLeafFold);
        unsigned cm_mask = (unsigned)(1UL << B) - 1;
        int fill = fill_in & (int)cm_mask;
        const int n = N;
        if (!fill) {
            CUO_PFOR(j, 0, n) fX[j] = 0;
        } else {
            // LCG noise: lane j draws seed_{j+1} = A^(j+1) seed + C_(j+1)
            // (jump-ahead of the serial recurrence, exact mod 2^32)
            const uint32_t seed0 = ctx->seed;
            if (flow == nullptr) {
                CUO_PFOR(j, 0, n) fX[j] = (int16_t)((int32_t)lcg_jump(seed0, j + 1) >> 20);
                cm = cm_mask;
            } else {
                CUO_PFOR(j, 0, n) {
                    int16_t tmp = (lcg_jump(seed0, j + 1) & 0x8000) ? 4 : -4;  // QCONST16(1/256, 10)
                    fX[j] = (int16_t)(flow[j] + tmp);
                }
                cm = (unsigned)fill;
            }
            ctx->seed = lcg_jump(seed0, n);
            renormalise_vector(fX, n, gain);
        }
    }
    return cm;
}

// Split test of quant_partition (bands.c:970-975).
CUO_FN bool qp_splits(const BandCtx *ctx, int N, int b, int LM)
{
    const uint8_t *cache = pulse_cache(ctx->band, LM + 1);
    return LM != -1 && b > (int)ld_u8(cache, (int)ld_u8(cache, 0)) + 12 && N > 2;
}

CUO_FI unsigned quant_partition(BandCtx *ctx, int16_t *X0, int N0, int b0, int B0, int16_t *lowband0,
                                int LM0, int16_t gain0, int fill0)
{
    // most partitions are leaves: decode those without the explicit stack
    if (!qp_splits(ctx, N0, b0, LM0)) return qp_leaf(ctx, X0, N0, b0, B0, lowband0, LM0, gain0, fill0);
    QpFrame *stk = lds_uniform((QpFrame *)ctx->qp_stack);
    int sp = 0;
    unsigned ret = 0;
    {
        QpFrame &f = stk[sp++];
        f.X = X0; f.N = N0; f.b = b0; f.B = B0; f.lowband = lowband0; f.LM = LM0; f.gain = gain0;
        f.fill = fill0; f.stage = 0;
    }
    while (sp > 0) {
        CUO_PROF_SCOPE(kProfQpStage);
        QpFrame &f = stk[sp - 1];
        if (f.stage == 0) {
            if (qp_splits(ctx, f.N, f.b, f.LM)) {
                f.B0 = f.B;
                f.N >>= 1;
                f.Y = f.X + f.N;
                f.LM -= 1;
                if (f.B == 1) f.fill = (f.fill & 1) | (f.fill << 1);
                f.B = (f.B + 1) >> 1;
                SplitOut so;
                {
                    CUO_PROF_SCOPE(kProfTheta);
                    so = compute_theta_mono(ctx, f.N, &f.b, f.B, f.B0, f.LM, &f.fill);
                }
                f.mid = (int16_t)so.imid;
                f.side = (int16_t)so.iside;
                int delta = so.delta;
                f.itheta = so.itheta;
                if (f.B0 > 1 && (f.itheta & 0x3fff)) {
                    if (f.itheta > 8192) delta -= delta >> (4 - f.LM);
                    else delta = imin(0, delta + (f.N << kBitRes >> (5 - f.LM)));
                }
                f.mbits = imax(0, imin(f.b, (f.b - delta) / 2));
                f.sbits = f.b - f.mbits;
                ctx->remaining_bits -= so.qalloc;
                f.lowband2 = f.lowband ? f.lowband + f.N : nullptr;
                f.rebalance = ctx->remaining_bits;
                f.mfirst = f.mbits >= f.sbits;
                f.stage = 1;
                QpFrame &c = stk[sp++];
                c.stage = 0; c.N = f.N; c.B = f.B; c.LM = f.LM;
                if (f.mfirst) {
                    c.X = f.X; c.b = f.mbits; c.lowband = f.lowband;
                    c.gain = (int16_t)mul16_16_p15(f.gain, f.mid); c.fill = f.fill;
                } else {
                    c.X = f.Y; c.b = f.sbits; c.lowband = f.lowband2;
                    c.gain = (int16_t)mul16_16_p15(f.gain, f.side); c.fill = f.fill >> f.B;
                }
                continue;
            }
            unsigned cm;
            {
                CUO_PROF_SCOPE(kProfQpLeaf);
                cm = qp_leaf(ctx, lds_uniform(f.X), f.N, f.b, f.B, f.lowband ? lds_uniform(f.lowband) : nullptr,
                             f.LM, f.gain, f.fill);
            }
            ret = cm;
            sp--;
            continue;
        }
        if (f.stage == 1) {
            if (f.mfirst) {
                f.cm = ret;
                int reb = f.mbits - (f.rebalance - ctx->remaining_bits);
                if (reb > 3 << kBitRes && f.itheta != 0) f.sbits += reb - (3 << kBitRes);
            } else {
                f.cm = ret << (f.B0 >> 1);
                int reb = f.sbits - (f.rebalance - ctx->remaining_bits);
                if (reb > 3 << kBitRes && f.itheta != 16384) f.mbits += reb - (3 << kBitRes);
            }
            f.stage = 2;
            QpFrame &c = stk[sp++];
            c.stage = 0; c.N = f.N; c.B = f.B; c.LM = f.LM;
            if (f.mfirst) {
                c.X = f.Y; c.b = f.sbits; c.lowband = f.lowband2;
                c.gain = (int16_t)mul16_16_p15(f.gain, f.side); c.fill = f.fill >> f.B;
            } else {
                c.X = f.X; c.b = f.mbits; c.lowband = f.lowband;
                c.gain = (int16_t)mul16_16_p15(f.gain, f.mid); c.fill = f.fill;
            }
            continue;
        }
        // stage 2: both children done
        if (f.mfirst) f.cm |= ret << (f.B0 >> 1); else f.cm |= ret;
        ret = f.cm;
        sp--;
    }
    return ret;
}

static constexpr uint8_t kBitInterleave[16] = {0, 1, 1, 1, 2, 3, 3, 3, 2, 3, 3, 3, 2, 3, 3, 3};
static constexpr uint8_t kBitDeinterleave[16] = {0x00, 0x03, 0x0C, 0x0F, 0x30, 0x33, 0x3C, 0x3F,
                                                 0xC0, 0xC3, 0xCC, 0xCF, 0xF0, 0xF3, 0xFC, 0xFF};

// quant_band, decoder, mono (bands.c:1109-1234)
CUO_FI unsigned quant_band(BandCtx *ctx, int16_t *X, int N, int b, int B, int16_t *lowband, int LM,
                           int16_t *lowband_out, int16_t gain, int16_t *lowband_scratch, int fill, int16_t *tmp)
{
    int N0 = N, N_B = N, B0 = B, time_divide = 0, recombine = 0;
    int longBlocks = B0 == 1;
    int tf_change = ctx->tf_change;
    N_B = N_B / B;
    if (N == 1) {
        int sign = 0;
        if (ctx->remaining_bits >= 1 << kBitRes) {
            sign = (int)ec_bits(ctx->ec, 1);
            ctx->remaining_bits -= 1 << kBitRes;
        }
        X[0] = sign ? -16384 : 16384;
        if (lowband_out) lowband_out[0] = (int16_t)(X[0] >> 4);
        return 1;
    }
    if (tf_change > 0) recombine = tf_change;
    unsigned cm;
    int N_B0;
    {
    CUO_PROF_SCOPE(kProfBandPre);
    if (lowband_scratch && lowband && (recombine || ((N_B & 1) == 0 && tf_change < 0) || B0 > 1)) {
        CUO_PFOR(j, 0, N) lowband_scratch[j] = lowband[j];
        lowband = lowband_scratch;
    }
    for (int k = 0; k < recombine; k++) {
        if (lowband) haar1(lowband, N >> k, 1 << k);
        fill = kBitInterleave[fill & 0xF] | kBitInterleave[fill >> 4] << 2;
    }
    B >>= recombine;
    N_B <<= recombine;
    while ((N_B & 1) == 0 && tf_change < 0) {
        if (lowband) haar1(lowband, N_B, B);
        fill |= fill << B;
        B <<= 1;
        N_B >>= 1;
        time_divide++;
        tf_change++;
    }
    B0 = B;
    N_B0 = N_B;
    if (B0 > 1 && lowband) deinterleave_hadamard(lowband, N_B >> recombine, B0 << recombine, longBlocks, tmp);
    }
    cm = quant_partition(ctx, X, N, b, B, lowband, LM, gain, fill);
    CUO_PROF_SCOPE(kProfBandPost);
    if (B0 > 1) interleave_hadamard(X, N_B >> recombine, B0 << recombine, longBlocks, tmp);
    N_B = N_B0;
    B = B0;
    for (int k = 0; k < time_divide; k++) {
        B >>= 1;
        N_B <<= 1;
        cm |= cm >> B;
        haar1(X, N_B, B);
    }
    for (int k = 0; k < recombine; k++) {
        cm = kBitDeinterleave[cm];
        haar1(X, N0 >> k, 1 << k);
    }
    B <<= recombine;
    if (lowband_out) {
        CUO_PROF_SCOPE(kProfLowbandOut);
        int16_t n = (int16_t)celt_sqrt(shl32(N0, 22));
        CUO_PFOR(j, 0, N0) lowband_out[j] = (int16_t)mul16_16_q15(n, X[j]);
    }
    cm &= (1u << B) - 1;
    return cm;
}

// ------------------------------------------------------------------ stereo bands
// The top split of a band coded as mid/side (compute_theta with stereo = 1,
// bands.c:647-903): a stepped pdf for N > 2, a uniform one for N == 2, no
// angle above the intensity band (then an optional phase-inversion flag).
struct StereoSplit { int imid, iside, delta, itheta, qalloc, inv; };

CUO_FI StereoSplit compute_theta_stereo(BandCtx *ctx, int N, int *b, int B, int LM, int *fill)
{
    StereoSplit o;
    const int i = ctx->band;
    const int pulse_cap = kLogN32.v[i] + LM * (1 << kBitRes);
    const int offset = (pulse_cap >> 1) - (N == 2 ? 16 : 4);  // QTHETA_OFFSET_TWOPHASE, QTHETA_OFFSET
    int qn = compute_qn(N, *b, offset, pulse_cap, 1);
    if (i >= ctx->intensity) qn = 1;
    const int32_t tell = (int32_t)ec_tell_frac(ctx->ec);
    int itheta = 0, inv = 0;
    if (qn != 1) {
        if (N > 2) {
            // probability 3 up to itheta = qn / 2, then 1
            const int p0 = 3, x0 = qn / 2, ft = p0 * (x0 + 1) + x0;
            const int fs = (int)ec_decode(ctx->ec, (uint32_t)ft);
            const int x = fs < (x0 + 1) * p0 ? fs / p0 : x0 + 1 + (fs - (x0 + 1) * p0);
            const int fl = x <= x0 ? p0 * x : (x - 1 - x0) + (x0 + 1) * p0;
            const int fh = x <= x0 ? p0 * (x + 1) : (x - x0) + (x0 + 1) * p0;
            ec_update(ctx->ec, (uint32_t)fl, (uint32_t)fh, (uint32_t)ft);
            itheta = x;
        } else {
            itheta = (int)ec_uint(ctx->ec, (uint32_t)(qn + 1));
        }
        itheta = (int)((int32_t)itheta * 16384 / qn);
    } else {
        if (*b > 2 << kBitRes && ctx->remaining_bits > 2 << kBitRes) inv = ec_bit_logp(ctx->ec, 2);
        if (ctx->disable_inv) inv = 0;
    }
    const int qalloc = (int32_t)ec_tell_frac(ctx->ec) - tell;
    *b -= qalloc;
    if (itheta == 0) {
        o.imid = 32767; o.iside = 0; *fill &= (1 << B) - 1; o.delta = -16384;
    } else if (itheta == 16384) {
        o.imid = 0; o.iside = 32767; *fill &= ((1 << B) - 1) << B; o.delta = 16384;
    } else {
        o.imid = bitexact_cos((int16_t)itheta);
        o.iside = bitexact_cos((int16_t)(16384 - itheta));
        o.delta = frac_mul16((N - 1) << 7, bitexact_log2tan(o.iside, o.imid));
    }
    o.itheta = itheta;
    o.qalloc = qalloc;
    o.inv = inv;
    return o;
}

// stereo_merge (bands.c:426-475): mid/side back to left/right with each
// side renormalised; X is the mid (unit norm), Y the side already scaled.
CUO_FN void stereo_merge(int16_t *X, int16_t *Y, int16_t mid, int N)
{
    int32_t xp = 0, side = 0;
    CUO_PFOR(j, 0, N) {
        xp = mac16_16(xp, Y[j], X[j]);
        side = mac16_16(side, Y[j], Y[j]);
    }
    xp = wave_sum(xp);
    side = wave_sum(side);
    xp = mul16_32_q15(mid, xp);
    const int16_t mid2 = (int16_t)(mid >> 1);
    const int32_t El = mul16_16(mid2, mid2) + side - 2 * xp;
    const int32_t Er = mul16_16(mid2, mid2) + side + 2 * xp;
    if (Er < 161061 || El < 161061) {  // QCONST32(6e-4f, 28)
        CUO_PFOR(j, 0, N) Y[j] = X[j];
        return;
    }
    int kl = celt_ilog2(El) >> 1, kr = celt_ilog2(Er) >> 1;
    const int16_t lgain = celt_rsqrt_norm(vshr32(El, (kl - 7) << 1));
    const int16_t rgain = celt_rsqrt_norm(vshr32(Er, (kr - 7) << 1));
    kl = imax(kl, 7);
    kr = imax(kr, 7);
    CUO_PFOR(j, 0, N) {
        const int16_t l = (int16_t)mul16_16_p15(mid, X[j]);
        const int16_t r = Y[j];
        X[j] = (int16_t)pshr32(mul16_16(lgain, sub16(l, r)), kl + 1);
        Y[j] = (int16_t)pshr32(mul16_16(rgain, add16(l, r)), kr + 1);
    }
}

CUO_FI unsigned quant_band(BandCtx *ctx, int16_t *X, int N, int b, int B, int16_t *lowband, int LM,
                           int16_t *lowband_out, int16_t gain, int16_t *lowband_scratch, int fill, int16_t *tmp);

// quant_band_stereo, decoder (bands.c:905-937 for N == 1, 1235-1382)
CUO_FN unsigned quant_band_stereo(BandCtx *ctx, int16_t *X, int16_t *Y, int N, int b, int B, int16_t *lowband,
                                  int LM, int16_t *lowband_out, int16_t *lowband_scratch, int fill, int16_t *tmp)
{
    if (N == 1) {
        for (int c = 0; c < 2; c++) {
            int sign = 0;
            if (ctx->remaining_bits >= 1 << kBitRes) {
                sign = (int)ec_bits(ctx->ec, 1);
                ctx->remaining_bits -= 1 << kBitRes;
            }
            (c ? Y : X)[0] = sign ? -16384 : 16384;
        }
        if (lowband_out) lowband_out[0] = (int16_t)(X[0] >> 4);
        return 1;
    }
    const int orig_fill = fill;
    StereoSplit sp;
    {
        CUO_PROF_SCOPE(kProfTheta);
        sp = compute_theta_stereo(ctx, N, &b, B, LM, &fill);
    }
    const int16_t mid = (int16_t)sp.imid, side = (int16_t)sp.iside;
    unsigned cm;
    if (N == 2) {
        // the side is orthogonal to the mid: one sign bit codes it
        const int sbits = sp.itheta != 0 && sp.itheta != 16384 ? 1 << kBitRes : 0;
        const int mbits = b - sbits;
        const int c = sp.itheta > 8192;
        ctx->remaining_bits -= sp.qalloc + sbits;
        int16_t *x2 = c ? Y : X, *y2 = c ? X : Y;
        int sign = 0;
        if (sbits) sign = (int)ec_bits(ctx->ec, 1);
        sign = 1 - 2 * sign;
        cm = quant_band(ctx, x2, N, mbits, B, lowband, LM, lowband_out, kQ15One, lowband_scratch, orig_fill, tmp);
        const int16_t a0 = x2[0], a1 = x2[1];
        y2[0] = (int16_t)(-sign * a1);
        y2[1] = (int16_t)(sign * a0);
        const int16_t x0 = (int16_t)mul16_16_q15(mid, X[0]), x1 = (int16_t)mul16_16_q15(mid, X[1]);
        const int16_t s0 = (int16_t)mul16_16_q15(side, Y[0]), s1 = (int16_t)mul16_16_q15(side, Y[1]);
        X[0] = (int16_t)sub16(x0, s0);
        Y[0] = add16(x0, s0);
        X[1] = (int16_t)sub16(x1, s1);
        Y[1] = add16(x1, s1);
    } else {
        int mbits = imax(0, imin(b, (b - sp.delta) / 2));
        int sbits = b - mbits;
        ctx->remaining_bits -= sp.qalloc;
        int32_t rebalance = ctx->remaining_bits;
        if (mbits >= sbits) {
            // the mid keeps unit gain: it is the folding source of later bands
            cm = quant_band(ctx, X, N, mbits, B, lowband, LM, lowband_out, kQ15One, lowband_scratch, fill, tmp);
            rebalance = mbits - (rebalance - ctx->remaining_bits);
            if (rebalance > 3 << kBitRes && sp.itheta != 0) sbits += rebalance - (3 << kBitRes);
            cm |= quant_band(ctx, Y, N, sbits, B, nullptr, LM, nullptr, side, nullptr, fill >> B, tmp);
        } else {
            cm = quant_band(ctx, Y, N, sbits, B, nullptr, LM, nullptr, side, nullptr, fill >> B, tmp);
            rebalance = sbits - (rebalance - ctx->remaining_bits);
            if (rebalance > 3 << kBitRes && sp.itheta != 16384) mbits += rebalance - (3 << kBitRes);
            cm |= quant_band(ctx, X, N, mbits, B, lowband, LM, lowband_out, kQ15One, lowband_scratch, fill, tmp);
        }
        stereo_merge(X, Y, mid, N);
    }
    if (sp.inv) CUO_PFOR(j, 0, N) Y[j] = (int16_t)-Y[j];
    return cm;
}

// quant_all_bands, decoder (bands.c:1398-1672).  Y_ is the second coded
// channel's spectrum (nullptr for one channel).
CUO_FI void quant_all_bands(int start, int end, int16_t *X_, int16_t *Y_, uint8_t *collapse, const int32_t *pulses,
                            int shortBlocks, int spread, int dual_stereo, int intensity, const int32_t *tf_res,
                            int32_t total_bits, int32_t balance, RangeDec *ec, int LM, int codedBands,
                            uint32_t *seed, int disable_inv, CeltScratch *sc)
{
    const int32_t *eB = kEB32.v;
    const int C = kMaxCh == 1 || Y_ == nullptr ? 1 : 2;
    const int M = 1 << LM;
    const int B = shortBlocks ? M : 1;
    const int norm_offset = M * eB[start];
    int16_t *norm = sc->norm;
    int16_t *norm2 = norm + M * eB[kNbBands - 1] - norm_offset;
    int16_t *lowband_scratch = X_ + M * eB[kNbBands - 1];
    int lowband_offset = 0, update_lowband = 1;
    BandCtx ctx;
    ctx.ec = ec; ctx.spread = spread; ctx.seed = *seed; ctx.iy = sc->iy; ctx.qp_stack = sc->qp_stack;
    ctx.intensity = intensity; ctx.disable_inv = disable_inv;
    for (int i = start; i < end; i++) {
        int effective_lowband = -1;
        ctx.band = i;
        int last = i == end - 1;
        CUO_PROF_SCOPE(kProfBandsPrologue);
        int16_t *X = X_ + M * eB[i];
        int16_t *Y = C == 2 ? Y_ + M * eB[i] : nullptr;
        int N = M * eB[i + 1] - M * eB[i];
        int32_t tell = (int32_t)ec_tell_frac(ec);
        if (i != start) balance -= tell;
        int32_t remaining_bits = total_bits - tell - 1;
        ctx.remaining_bits = remaining_bits;
        int b;
        if (i <= codedBands - 1) {
            // balance / imin(3, codedBands - i), the three divisors spelled out
            // so each division is a constant one
            const int d = imin(3, codedBands - i);
            int32_t curr_balance = d == 3 ? balance / 3 : (d == 2 ? balance / 2 : balance);
            b = imax(0, imin(16383, imin(remaining_bits + 1, pulses[i] + curr_balance)));
        } else {
            b = 0;
        }
        if ((M * eB[i] - N >= M * eB[start] || i == start + 1) && (update_lowband || lowband_offset == 0))
            lowband_offset = i;
        if (i == start + 1) {
            // special_hybrid_folding: the source ends where the target starts
            int n1 = M * (eB[start + 1] - eB[start]);
            int n2 = M * (eB[start + 2] - eB[start + 1]);
            CUO_PFOR(j, 0, n2 - n1) norm[n1 + j] = norm[2 * n1 - n2 + j];
            if (dual_stereo) CUO_PFOR(j, 0, n2 - n1) norm2[n1 + j] = norm2[2 * n1 - n2 + j];
        }
        int tf_change = tf_res[i];
        ctx.tf_change = tf_change;
        if (last) lowband_scratch = nullptr;
        unsigned x_cm, y_cm;
        if (lowband_offset != 0 && (spread != SPREAD_AGGRESSIVE || B > 1 || tf_change < 0)) {
            effective_lowband = imax(0, M * eB[lowband_offset] - norm_offset - N);
            int fold_start = lowband_offset;
            while (M * eB[--fold_start] > effective_lowband + norm_offset) {}
            int fold_end = lowband_offset - 1;
            while (++fold_end < i && M * eB[fold_end] < effective_lowband + norm_offset + N) {}
            x_cm = y_cm = 0;
            int fold_i = fold_start;
            do {
                x_cm |= collapse[fold_i * C];
                y_cm |= collapse[fold_i * C + C - 1];
            } while (++fold_i < fold_end);
        } else {
            x_cm = y_cm = (1u << B) - 1;
        }
        if (dual_stereo && i == intensity) {
            // from the intensity band on, both channels fold from their average
            dual_stereo = 0;
            CUO_PFOR(j, 0, M * eB[i] - norm_offset) norm[j] = (int16_t)(((int32_t)norm[j] + norm2[j]) >> 1);
        }
        int16_t *lowband = effective_lowband != -1 ? norm + effective_lowband : nullptr;
        int16_t *lowband_out = last ? nullptr : norm + M * eB[i] - norm_offset;
        if (dual_stereo) {
            x_cm = quant_band(&ctx, X, N, b / 2, B, lowband, LM, lowband_out, kQ15One, lowband_scratch, (int)x_cm,
                              sc->tmp);
            y_cm = quant_band(&ctx, Y, N, b / 2, B, effective_lowband != -1 ? norm2 + effective_lowband : nullptr,
                              LM, last ? nullptr : norm2 + M * eB[i] - norm_offset, kQ15One, lowband_scratch,
                              (int)y_cm, sc->tmp);
        } else {
            if (C == 2)
                x_cm = quant_band_stereo(&ctx, X, Y, N, b, B, lowband, LM, lowband_out, lowband_scratch,
                                         (int)(x_cm | y_cm), sc->tmp);
            else
                x_cm = quant_band(&ctx, X, N, b, B, lowband, LM, lowband_out, kQ15One, lowband_scratch,
                                  (int)(x_cm | y_cm), sc->tmp);
            y_cm = x_cm;
        }
        collapse[i * C] = (uint8_t)x_cm;
        collapse[i * C + C - 1] = (uint8_t)y_cm;
        balance += pulses[i] + tell;
        update_lowband = b > (N << kBitRes);
    }
    *seed = ctx.seed;
}

// anti_collapse (bands.c:268-370); channel c's spectrum at X_ + c * size
CUO_FN void anti_collapse(int16_t *X_, const uint8_t *collapse, int LM, int C, int size, int start, int end,
                          const int16_t *logE, const int16_t *prev1logE, const int16_t *prev2logE,
                          const int32_t *pulses, uint32_t seed)
{
    const int32_t *eB = kEB32.v;
    C = kMaxCh == 1 ? 1 : C;
    for (int i = start; i < end; i++) {
        int N0 = eB[i + 1] - eB[i];
        int depth = ((1 + pulses[i]) / N0) >> LM;
        int32_t thresh32 = celt_exp2((int16_t)-shl16(depth, 10 - kBitRes)) >> 1;
        int16_t thresh = (int16_t)mul16_32_q15(16384, min32(32767, thresh32));
        int32_t t = N0 << LM;
        int shift = celt_ilog2(t) >> 1;
        t = shl32(t, (7 - shift) << 1);
        int16_t sqrt_1 = celt_rsqrt_norm(t);
        for (int c = 0; c < C; c++) {
            int16_t prev1 = prev1logE[c * kNbBands + i];
            int16_t prev2 = prev2logE[c * kNbBands + i];
            if (C == 1) {
                prev1 = max16(prev1, prev1logE[kNbBands + i]);
                prev2 = max16(prev2, prev2logE[kNbBands + i]);
            }
            int32_t Ediff = (int32_t)logE[c * kNbBands + i] - (int32_t)min16(prev1, prev2);
            Ediff = max32(0, Ediff);
            int16_t r;
            if (Ediff < 16384) {
                int32_t r32 = celt_exp2((int16_t)-(int16_t)Ediff) >> 1;
                r = (int16_t)(2 * min32(16383, r32));
            } else {
                r = 0;
            }
            if (LM == 3) r = (int16_t)mul16_16_q14(23170, min32(23169, r));
            r = (int16_t)(min16(thresh, r) >> 1);
            r = (int16_t)(mul16_16_q15(sqrt_1, r) >> shift);
            int16_t *X = X_ + c * size + (eB[i] << LM);
            int renormalize = 0;
            for (int k = 0; k < 1 << LM; k++) {
                if (!(collapse[i * C + c] & 1 << k)) {
                    for (int j = 0; j < N0; j++) {
                        seed = lcg_rand(seed);
                        X[(j << LM) + k] = (seed & 0x8000) ? r : (int16_t)-r;
                    }
                    renormalize = 1;
                }
            }
            if (renormalize) renormalise_vector(X, N0 << LM, kQ15One);
        }
    }
}

// ------------------------------------------------------------------ synthesis
CUO_FN void denormalise_bands(const int16_t *X, int32_t *freq, const int16_t *bandLogE, int start, int end,
                              int M, int silence)
{
    const int32_t *eB = kEB32.v;
    int N = M * 120;
    int bound = M * eB[end];
    if (silence) { bound = 0; start = end = 0; }
    CUO_PFOR(j, 0, M * eB[start]) freq[j] = 0;
    for (int i = start; i < end; i++) {
        const int lo = M * eB[i], hi = M * eB[i + 1];
        int16_t lg = sat16(bandLogE[i] + shl32(kEMeans32.v[i], 6));
        int shift = 16 - (lg >> kDbShift);
        int16_t g;
        if (shift > 31) { shift = 0; g = 0; }
        else g = (int16_t)celt_exp2_frac((int16_t)(lg & ((1 << kDbShift) - 1)));
        if (shift < 0) {
            if (shift <= -2) { g = 16384; shift = -2; }
            CUO_PFOR(j, lo, hi) freq[j] = shl32(mul16_16(X[j], g), -shift);
        } else {
            CUO_PFOR(j, lo, hi) freq[j] = shr32(mul16_16(X[j], g), shift);
        }
    }
    CUO_PFOR(j, bound, N) freq[j] = 0;
}

// FFT butterflies on interleaved int32 complex (kiss_fft.c:48-317); twiddles
// are int16 pairs of the 480-point table, strided for smaller sizes.
struct Cpx { int32_t r, i; };
CUO_FN int32_t smul(int32_t a, int16_t b) { return mul16_32_q15(b, a); }

// Butterflies of one radix stage are independent: each lane takes whole
// butterflies (index i over the N sub-transforms, j over the m positions).
CUO_FN void bfly2(Cpx *F0, int N)
{
    const int16_t tw = 23170;
    CUO_PFOR(i, 0, N) {
        Cpx *F = F0 + 8 * i;
        Cpx *F2 = F + 4;
        Cpx t = F2[0];
        F2[0].r = sub32_ovf(F[0].r, t.r); F2[0].i = sub32_ovf(F[0].i, t.i);
        F[0].r = add32_ovf(F[0].r, t.r); F[0].i = add32_ovf(F[0].i, t.i);
        t.r = smul(add32_ovf(F2[1].r, F2[1].i), tw);
        t.i = smul(sub32_ovf(F2[1].i, F2[1].r), tw);
        F2[1].r = sub32_ovf(F[1].r, t.r); F2[1].i = sub32_ovf(F[1].i, t.i);
        F[1].r = add32_ovf(F[1].r, t.r); F[1].i = add32_ovf(F[1].i, t.i);
        t.r = F2[2].i; t.i = neg32_ovf(F2[2].r);
        F2[2].r = sub32_ovf(F[2].r, t.r); F2[2].i = sub32_ovf(F[2].i, t.i);
        F[2].r = add32_ovf(F[2].r, t.r); F[2].i = add32_ovf(F[2].i, t.i);
        t.r = smul(sub32_ovf(F2[3].i, F2[3].r), tw);
        t.i = smul(neg32_ovf(add32_ovf(F2[3].i, F2[3].r)), tw);
        F2[3].r = sub32_ovf(F[3].r, t.r); F2[3].i = sub32_ovf(F[3].i, t.i);
        F[3].r = add32_ovf(F[3].r, t.r); F[3].i = add32_ovf(F[3].i, t.i);
    }
}

CUO_FN Cpx cmul(Cpx a, int16_t br, int16_t bi)
{
    Cpx m;
    m.r = sub32_ovf(smul(a.r, br), smul(a.i, bi));
    m.i = add32_ovf(smul(a.r, bi), smul(a.i, br));
    return m;
}

CUO_FN void bfly4(Cpx *Fbeg, int fstride, int m, int N, int mm)
{
    if (m == 1) {
        CUO_PFOR(i, 0, N) {
            Cpx *Fout = Fbeg + 4 * i;
            Cpx s0, s1;
            s0.r = sub32_ovf(Fout[0].r, Fout[2].r); s0.i = sub32_ovf(Fout[0].i, Fout[2].i);
            Fout[0].r = add32_ovf(Fout[0].r, Fout[2].r); Fout[0].i = add32_ovf(Fout[0].i, Fout[2].i);
            s1.r = add32_ovf(Fout[1].r, Fout[3].r); s1.i = add32_ovf(Fout[1].i, Fout[3].i);
            Fout[2].r = sub32_ovf(Fout[0].r, s1.r); Fout[2].i = sub32_ovf(Fout[0].i, s1.i);
            Fout[0].r = add32_ovf(Fout[0].r, s1.r); Fout[0].i = add32_ovf(Fout[0].i, s1.i);
            s1.r = sub32_ovf(Fout[1].r, Fout[3].r); s1.i = sub32_ovf(Fout[1].i, Fout[3].i);
            Fout[1].r = add32_ovf(s0.r, s1.i); Fout[1].i = sub32_ovf(s0.i, s1.r);
            Fout[3].r = sub32_ovf(s0.r, s1.i); Fout[3].i = add32_ovf(s0.i, s1.r);
        }
        return;
    }
    const int16_t *tw = tab::kFftTwiddles;
    CUO_PFOR(k, 0, N * m) {
        const int i = k / m, j = k - i * m;
        Cpx *Fout = Fbeg + i * mm + j;
        const int t1 = j * fstride, t2 = 2 * j * fstride, t3 = 3 * j * fstride;
        Cpx s0 = cmul(Fout[m], tw[2 * t1], tw[2 * t1 + 1]);
        Cpx s1 = cmul(Fout[2 * m], tw[2 * t2], tw[2 * t2 + 1]);
        Cpx s2 = cmul(Fout[3 * m], tw[2 * t3], tw[2 * t3 + 1]);
        Cpx s5, s3, s4;
        s5.r = sub32_ovf(Fout[0].r, s1.r); s5.i = sub32_ovf(Fout[0].i, s1.i);
        Fout[0].r = add32_ovf(Fout[0].r, s1.r); Fout[0].i = add32_ovf(Fout[0].i, s1.i);
        s3.r = add32_ovf(s0.r, s2.r); s3.i = add32_ovf(s0.i, s2.i);
        s4.r = sub32_ovf(s0.r, s2.r); s4.i = sub32_ovf(s0.i, s2.i);
        Fout[2 * m].r = sub32_ovf(Fout[0].r, s3.r); Fout[2 * m].i = sub32_ovf(Fout[0].i, s3.i);
        Fout[0].r = add32_ovf(Fout[0].r, s3.r); Fout[0].i = add32_ovf(Fout[0].i, s3.i);
        Fout[m].r = add32_ovf(s5.r, s4.i); Fout[m].i = sub32_ovf(s5.i, s4.r);
        Fout[3 * m].r = sub32_ovf(s5.r, s4.i); Fout[3 * m].i = add32_ovf(s5.i, s4.r);
    }
}

CUO_FN void bfly3(Cpx *Fbeg, int fstride, int m, int N, int mm)
{
    const int16_t *tw = tab::kFftTwiddles;
    const int16_t epi3i = -28378;
    CUO_PFOR(k, 0, N * m) {
        const int i = k / m, j = k - i * m;
        Cpx *Fout = Fbeg + i * mm + j;
        const int t1 = j * fstride, t2 = 2 * j * fstride;
        Cpx s1 = cmul(Fout[m], tw[2 * t1], tw[2 * t1 + 1]);
        Cpx s2 = cmul(Fout[2 * m], tw[2 * t2], tw[2 * t2 + 1]);
        Cpx s3, s0;
        s3.r = add32_ovf(s1.r, s2.r); s3.i = add32_ovf(s1.i, s2.i);
        s0.r = sub32_ovf(s1.r, s2.r); s0.i = sub32_ovf(s1.i, s2.i);
        Cpx fm;
        fm.r = sub32_ovf(Fout[0].r, s3.r >> 1);
        fm.i = sub32_ovf(Fout[0].i, s3.i >> 1);
        s0.r = smul(s0.r, epi3i); s0.i = smul(s0.i, epi3i);
        Fout[0].r = add32_ovf(Fout[0].r, s3.r); Fout[0].i = add32_ovf(Fout[0].i, s3.i);
        Fout[2 * m].r = add32_ovf(fm.r, s0.i);
        Fout[2 * m].i = sub32_ovf(fm.i, s0.r);
        Fout[m].r = sub32_ovf(fm.r, s0.i);
        Fout[m].i = add32_ovf(fm.i, s0.r);
    }
}

CUO_FN void bfly5(Cpx *Fbeg, int fstride, int m, int N, int mm)
{
    const int16_t *tw = tab::kFftTwiddles;
    const int16_t yar = 10126, yai = -31164, ybr = -26510, ybi = -19261;
    CUO_PFOR(k, 0, N * m) {
        const int i = k / m, u = k - i * m;
        Cpx *F0 = Fbeg + i * mm + u, *F1 = F0 + m, *F2 = F0 + 2 * m, *F3 = F0 + 3 * m, *F4 = F0 + 4 * m;
        Cpx s0 = *F0;
        int a1 = u * fstride, a2 = 2 * u * fstride, a3 = 3 * u * fstride, a4 = 4 * u * fstride;
        Cpx s1 = cmul(*F1, tw[2 * a1], tw[2 * a1 + 1]);
        Cpx s2 = cmul(*F2, tw[2 * a2], tw[2 * a2 + 1]);
        Cpx s3 = cmul(*F3, tw[2 * a3], tw[2 * a3 + 1]);
        Cpx s4 = cmul(*F4, tw[2 * a4], tw[2 * a4 + 1]);
        Cpx s7, s10, s8, s9, s5, s6, s11, s12;
        s7.r = add32_ovf(s1.r, s4.r); s7.i = add32_ovf(s1.i, s4.i);
        s10.r = sub32_ovf(s1.r, s4.r); s10.i = sub32_ovf(s1.i, s4.i);
        s8.r = add32_ovf(s2.r, s3.r); s8.i = add32_ovf(s2.i, s3.i);
        s9.r = sub32_ovf(s2.r, s3.r); s9.i = sub32_ovf(s2.i, s3.i);
        F0->r = add32_ovf(F0->r, add32_ovf(s7.r, s8.r));
        F0->i = add32_ovf(F0->i, add32_ovf(s7.i, s8.i));
        s5.r = add32_ovf(s0.r, add32_ovf(smul(s7.r, yar), smul(s8.r, ybr)));
        s5.i = add32_ovf(s0.i, add32_ovf(smul(s7.i, yar), smul(s8.i, ybr)));
        s6.r = add32_ovf(smul(s10.i, yai), smul(s9.i, ybi));
        s6.i = neg32_ovf(add32_ovf(smul(s10.r, yai), smul(s9.r, ybi)));
        F1->r = sub32_ovf(s5.r, s6.r); F1->i = sub32_ovf(s5.i, s6.i);
        F4->r = add32_ovf(s5.r, s6.r); F4->i = add32_ovf(s5.i, s6.i);
        s11.r = add32_ovf(s0.r, add32_ovf(smul(s7.r, ybr), smul(s8.r, yar)));
        s11.i = add32_ovf(s0.i, add32_ovf(smul(s7.i, ybr), smul(s8.i, yar)));
        s12.r = sub32_ovf(smul(s9.i, yai), smul(s10.i, ybi));
        s12.i = sub32_ovf(smul(s10.r, ybi), smul(s9.r, yai));
        F2->r = add32_ovf(s11.r, s12.r); F2->i = add32_ovf(s11.i, s12.i);
        F3->r = sub32_ovf(s11.r, s12.r); F3->i = sub32_ovf(s11.i, s12.i);
    }
}

// opus_fft_impl for FFT index k (0:480, 1:240, 2:120, 3:60) (kiss_fft.c:521-567)
CUO_FN void fft_impl(int k, Cpx *fout)
{
    const int16_t *fac = tab::kFftFactors + 16 * k;
    int shift = tab::kFftShift[k] > 0 ? tab::kFftShift[k] : 0;
    int fstride[8];
    fstride[0] = 1;
    int L = 0, m;
    do {
        int p = fac[2 * L];
        m = fac[2 * L + 1];
        fstride[L + 1] = fstride[L] * p;
        L++;
    } while (m != 1);
    m = fac[2 * L - 1];
    for (int i = L - 1; i >= 0; i--) {
        int m2 = i != 0 ? fac[2 * i - 1] : 1;
        switch (fac[2 * i]) {
        case 2: bfly2(fout, fstride[i]); break;
        case 4: bfly4(fout, fstride[i] << shift, m, fstride[i], m2); break;
        case 3: bfly3(fout, fstride[i] << shift, m, fstride[i], m2); break;
        case 5: bfly5(fout, fstride[i] << shift, m, fstride[i], m2); break;
        }
        m = m2;
    }
}

// clt_mdct_backward (mdct.c:242-342): in has stride `stride`, out gets the
// overlap-added time signal.  The N/4-point FFT runs in fbuf (LDS, N/4
// complex values) rather than in place in the history buffer: the
// pre-rotation writes fbuf, the butterflies work on LDS, the post-rotation
// reads fbuf and writes the history.  Outputs below sat_n are stored
// saturated to +-kSigSat (the synthesis' saturation pass folded into the
// stores: with one block per frame nothing reads them in between).
CUO_FN void mdct_backward(const int32_t *in, int32_t *out, int shift, int stride, Cpx *fbuf, int sat_n)
{
    int N = tab::kMdctN;
    const int16_t *trig = tab::kMdctTrig;
    for (int i = 0; i < shift; i++) { N >>= 1; trig += N; }
    const int N2 = N >> 1, N4 = N >> 2;
    int32_t *yp = out + (kOverlap >> 1);
    {
        // pre-rotation into bit-reversed order (independent per i)
        const int16_t *bitrev = tab::kFftBitrev + tab::kFftBitrevOff[shift];
        CUO_PFOR(i, 0, N4) {
            int rev = bitrev[i];
            int32_t x1 = in[2 * stride * i], x2 = in[stride * (N2 - 1) - 2 * stride * i];
            int32_t yr = add32_ovf(smul(x2, trig[i]), smul(x1, trig[N4 + i]));
            int32_t yi = sub32_ovf(smul(x1, trig[i]), smul(x2, trig[N4 + i]));
            fbuf[rev].i = yr;
            fbuf[rev].r = yi;
        }
    }
    fft_impl(shift, fbuf);
    {
        // post-rotation: iteration i owns the pairs at 2i and N2-2-2i
        const int32_t *fb = (const int32_t *)fbuf;
        CUO_PFOR(i, 0, (N4 + 1) >> 1) {
            const int32_t *fb0 = fb + 2 * i, *fb1 = fb + N2 - 2 - 2 * i;
            int32_t *yp0 = yp + 2 * i, *yp1 = yp + N2 - 2 - 2 * i;
            int32_t re = fb0[1], im = fb0[0];
            int16_t t0 = trig[i], t1 = trig[N4 + i];
            int32_t yr = add32_ovf(smul(re, t0), smul(im, t1));
            int32_t yi = sub32_ovf(smul(re, t1), smul(im, t0));
            re = fb1[1];
            im = fb1[0];
            // TDAC rereads the outputs below kOverlap: those are saturated there
            const int o0 = (kOverlap >> 1) + 2 * i, o1 = (kOverlap >> 1) + N2 - 2 - 2 * i;
            yp0[0] = o0 >= kOverlap && o0 < sat_n ? saturate(yr, kSigSat) : yr;
            yp1[1] = o1 + 1 >= kOverlap && o1 + 1 < sat_n ? saturate(yi, kSigSat) : yi;
            t0 = trig[N4 - i - 1];
            t1 = trig[N2 - i - 1];
            yr = add32_ovf(smul(re, t0), smul(im, t1));
            yi = sub32_ovf(smul(re, t1), smul(im, t0));
            yp1[0] = o1 >= kOverlap && o1 < sat_n ? saturate(yr, kSigSat) : yr;
            yp0[1] = o0 + 1 >= kOverlap && o0 + 1 < sat_n ? saturate(yi, kSigSat) : yi;
        }
    }
    {
        // TDAC mirror over the overlap: iteration i owns out[i] and out[119-i]
        const int16_t *win = tab::kWindow;
        CUO_PFOR(i, 0, kOverlap / 2) {
            int32_t x1 = out[kOverlap - 1 - i], x2 = out[i];
            int16_t w1 = win[i], w2 = win[kOverlap - 1 - i];
            const int32_t a = sub32_ovf(mul16_32_q15(w2, x2), mul16_32_q15(w1, x1));
            const int32_t b = add32_ovf(mul16_32_q15(w1, x2), mul16_32_q15(w2, x1));
            out[i] = sat_n > 0 ? saturate(a, kSigSat) : a;
            out[kOverlap - 1 - i] = sat_n > 0 ? saturate(b, kSigSat) : b;
        }
    }
}

// The inverse MDCTs run their FFTs in the band decoder's scratch (X, norm,
// tmp, iy: dead once denormalise_bands has read X), which holds the largest
// FFT (480 complex values).
static_assert(offsetof(CeltScratch, iy) + sizeof(CeltScratch::iy) - offsetof(CeltScratch, X) >= 480 * 8,
              "band scratch too small for the synthesis FFT");
// celt_synthesis (celt_decoder.c:382-458).  C coded channels (spectra at X
// and X + N) to CC output channels (histories kMemCap apart from out_syn):
// a mono stream is synthesised into both outputs, a stereo stream to one
// output is the average of the two denormalised spectra.  Every spectrum is
// denormalised before the first inverse MDCT, whose FFT reuses X's storage.
CUO_FI void celt_synthesis(const int16_t *X, int32_t *out_syn, const int16_t *oldBandE, int start, int effEnd,
                           int C, int CC, int isTransient, int LM, int silence, int32_t *freq, CeltScratch *sc)
{
    CUO_PROF_SCOPE(kProfCeltSynth);
    C = kMaxCh == 1 ? 1 : C;
    CC = kMaxCh == 1 ? 1 : CC;
    X = lds_uniform(X);
    oldBandE = lds_uniform(oldBandE);
    Cpx *fbuf = lds_uniform((Cpx *)sc->X);
    const int N = 120 << LM, M = 1 << LM;
    int B, NB, shift;
    if (isTransient) { B = M; NB = 120; shift = 3; }
    else { B = 1; NB = 120 << LM; shift = 3 - LM; }
    denormalise_bands(X, freq, oldBandE, start, effEnd, M, silence);
    if (C == 2) {
        int32_t *freq2 = freq + 960;
        denormalise_bands(X + N, freq2, oldBandE + kNbBands, start, effEnd, M, silence);
        if (CC == 1) CUO_PFOR(i, 0, N) freq[i] = (freq[i] >> 1) + (freq2[i] >> 1);
    }
    for (int c = 0; c < CC; c++) {
        const int32_t *fc = C == 2 ? freq + 960 * c : freq;
        // one block: saturated as stored; several: each block's TDAC reads
        // the previous block's unsaturated output, so saturate afterwards
        for (int b = 0; b < B; b++)
            mdct_backward(&fc[b], out_syn + kMemCap * c + NB * b, shift, B, fbuf, B == 1 ? N : 0);
    }
    if (B > 1) {
        for (int c = 0; c < CC; c++) {
            int32_t *o = out_syn + kMemCap * c;
            CUO_PFOR(i, 0, N) o[i] = saturate(o[i], kSigSat);
        }
    }
}

// ------------------------------------------------------------------ post-filter
static constexpr int16_t kCombGains[3][3] = {{10048, 7112, 4248}, {15200, 8784, 0}, {26208, 3280, 0}};

// y[i] = x[i] + g10 x[i-T] + g11 (x[i-T+1] + x[i-T-1]) + g12 (x[i-T+2] + x[i-T-2]),
// computed in blocks of `blk` consecutive outputs spread over the lanes.  When
// y == x the filter is recursive with lag >= T - 2, so blk <= T - 2 keeps every
// input of a block already final, exactly as in the serial loop.
CUO_FN void comb_filter_const(int32_t *y, const int32_t *x, int T, int N, int16_t g10, int16_t g11, int16_t g12,
                              int blk)
{
    for (int i0 = 0; i0 < N; i0 += blk) {
        const int i1 = imin(N, i0 + blk);
        CUO_PFOR(i, i0, i1) {
            int32_t v = x[i] + mul16_32_q15(g10, x[i - T]) + mul16_32_q15(g11, x[i - T + 1] + x[i - T - 1])
                      + mul16_32_q15(g12, x[i - T + 2] + x[i - T - 2]);
            y[i] = saturate(v, kSigSat);
        }
    }
}

// comb_filter (celt.c:190-256); x may equal y (in-place); window may be null
CUO_FN void comb_filter(int32_t *y, int32_t *x, int T0, int T1, int N, int16_t g0, int16_t g1, int tapset0,
                        int tapset1, const int16_t *window, int overlap)
{
    if (g0 == 0 && g1 == 0) {
        if (x != y) CUO_PFOR(i, 0, N) y[i] = x[i];
        return;
    }
    T0 = imax(T0, kCombMinPeriod);
    T1 = imax(T1, kCombMinPeriod);
    int16_t g00 = (int16_t)mul16_16_p15(g0, kCombGains[tapset0][0]);
    int16_t g01 = (int16_t)mul16_16_p15(g0, kCombGains[tapset0][1]);
    int16_t g02 = (int16_t)mul16_16_p15(g0, kCombGains[tapset0][2]);
    int16_t g10 = (int16_t)mul16_16_p15(g1, kCombGains[tapset1][0]);
    int16_t g11 = (int16_t)mul16_16_p15(g1, kCombGains[tapset1][1]);
    int16_t g12 = (int16_t)mul16_16_p15(g1, kCombGains[tapset1][2]);
    if (g0 == g1 && T0 == T1 && tapset0 == tapset1) overlap = 0;
    const bool inplace = x == y;
    int blk = inplace ? imin(T0, T1) - 2 : kLanes;
    if (blk > kLanes) blk = kLanes;
    for (int i0 = 0; i0 < overlap; i0 += blk) {
        const int i1 = imin(overlap, i0 + blk);
        CUO_PFOR(i, i0, i1) {
            int16_t f = (int16_t)mul16_16_q15(window[i], window[i]);
            int32_t v = x[i]
                + mul16_32_q15(mul16_16_q15((kQ15One - f), g00), x[i - T0])
                + mul16_32_q15(mul16_16_q15((kQ15One - f), g01), x[i - T0 + 1] + x[i - T0 - 1])
                + mul16_32_q15(mul16_16_q15((kQ15One - f), g02), x[i - T0 + 2] + x[i - T0 - 2])
                + mul16_32_q15(mul16_16_q15(f, g10), x[i - T1])
                + mul16_32_q15(mul16_16_q15(f, g11), x[i - T1 + 1] + x[i - T1 - 1])
                + mul16_32_q15(mul16_16_q15(f, g12), x[i - T1 + 2] + x[i - T1 - 2]);
            y[i] = saturate(v, kSigSat);
        }
    }
    if (g1 == 0) {
        if (x != y) CUO_PFOR(j, overlap, N) y[j] = x[j];
        return;
    }
    int blk1 = inplace ? imin(kLanes, T1 - 2) : kLanes;
    comb_filter_const(y + overlap, x + overlap, T1, N - overlap, g10, g11, g12, blk1);
}

// deemphasis of one channel, no downsampling (celt_decoder.c:279-377); the
// output is written every `stride` samples (interleaved channels).
// The history (x) and the frame output live in HBM: the one-pole recurrence
// runs register-blocked, lanes loading and storing 64 samples at a time.
CUO_FN void deemphasis(const int32_t *x, int16_t *pcm, int N, int32_t *mem, int accum, int stride)
{
    stride = kMaxCh == 1 ? 1 : stride;
    CUO_PROF_SCOPE(kProfCeltDeemph);
    const int16_t coef0 = tab::kPreemph[0];
    int32_t m = uniform(*mem);
    N = uniform(N);
    for (int j0 = 0; j0 < N; j0 += kLanes) {
        const int nb = uniform(imin(kLanes, N - j0));
        const int l = lane_id();
        int32_t v = l < nb ? x[j0 + l] : 0;
        int32_t t = 0;
        serial_lanes(nb, [&](auto k) {
            int32_t tmp = lane_get(v, k) + m;
            m = mul16_32_q15(coef0, tmp);
            t = lane_set_s(t, k, tmp);
        });
        if (l < nb) {
            int16_t y = sig2word16(t);
            int16_t *p = pcm + (j0 + l) * stride;
            *p = accum ? sat16(*p + y) : y;
        }
    }
    *mem = m;
}

// ------------------------------------------------------------------ loss concealment
// prefilter_and_fold (celt_decoder.c:515-549), each output channel
CUO_FN void prefilter_and_fold(CeltState *st, int N, int32_t *etmp)
{
    const int CC = celt_cc(st);
    for (int c = 0; c < CC; c++) {
        int32_t *mem = st->mem + kMemCap * c;
        comb_filter(etmp, mem + kDecBuf - N, st->pf_period_old, st->pf_period, kOverlap, (int16_t)-st->pf_gain_old,
                    (int16_t)-st->pf_gain, st->pf_tapset_old, st->pf_tapset, nullptr, 0);
        CUO_PFOR(i, 0, kOverlap / 2)
            mem[kDecBuf - N + i] = mul16_32_q15(tab::kWindow[i], etmp[kOverlap - 1 - i])
                                 + mul16_32_q15(tab::kWindow[kOverlap - i - 1], etmp[i]);
    }
}

// Pitch-based concealment (CELT-only frames after a loss); defined in celt_plc.h.
CUO_FN int celt_plc_periodic(CeltState *st, int N, CeltScratch *sc);

// celt_decode_lost (celt_decoder.c:604-969). Returns 0, or -1 if the
// concealment path is not available.
CUO_FN int celt_decode_lost(CeltState *st, int N, int LM, CeltScratch *sc)
{
    CUO_PROF_SCOPE(kProfCeltLost);
    const int CC = celt_cc(st);
    int32_t *mem = st->mem;
    int32_t *out_syn = mem + kDecBuf - N;
    int loss_duration = st->loss_duration;
    int start = st->start;
    int noise_based = loss_duration >= 40 || start != 0 || st->skip_plc;
    if (noise_based) {
        int end = st->end;
        int effEnd = imax(start, imin(end, kNbBands));
        mem = advance_history(st, N);
        out_syn = mem + kDecBuf - N;
        if (st->prefilter_and_fold) prefilter_and_fold(st, N, st->etmp);
        int16_t decay = loss_duration == 0 ? 1536 : 512;
        for (int c = 0; c < CC; c++)
            for (int i = start; i < end; i++) {
                const int k = c * kNbBands + i;
                st->oldBandE[k] = max16(st->backgroundLogE[k], (int16_t)(st->oldBandE[k] - decay));
            }
        uint32_t seed = st->rng;
        for (int c = 0; c < CC; c++) {
            int16_t *X = sc->X + N * c;
            for (int i = start; i < effEnd; i++) {
                int boffs = kEB32.v[i] << LM;
                int blen = (kEB32.v[i + 1] - kEB32.v[i]) << LM;
                for (int j = 0; j < blen; j++) {
                    seed = lcg_rand(seed);
                    X[boffs + j] = (int16_t)((int32_t)seed >> 20);
                }
                renormalise_vector(X + boffs, blen, kQ15One);
            }
        }
        st->rng = seed;
        celt_synthesis(sc->X, out_syn, st->oldBandE, start, effEnd, CC, CC, 0, LM, 0, st->freq, sc);
        st->prefilter_and_fold = 0;
        st->skip_plc = 1;
    } else {
        if (celt_plc_periodic(st, N, sc) < 0) return -1;
        st->prefilter_and_fold = 1;
    }
    st->loss_duration = imin(10000, loss_duration + (1 << LM));
    return 0;
}

// ------------------------------------------------------------------ frame decode
// celt_decode_with_ec (celt_decoder.c:970-1369).  data == nullptr or len <= 1
// conceals.  ec may be nullptr (then a decoder is opened on data).  Returns the
// number of samples (frame_size) or a negative error.  celt_decode_frame is
// the body, inlined where a caller decodes runs of frames; celt_decode is its
// out-of-line form for every other call site.
// MayConceal false: the caller guarantees data != nullptr and len > 1.
template <bool MayConceal>
CUO_FI int celt_decode_frame(CeltState *st, const uint8_t *data, int len, int16_t *pcm, int frame_size,
                             RangeDec *dec, int accum, CeltScratch *sc)
{
    CUO_PROF_SCOPE(kProfCelt);
    st = lds_uniform(st);
    sc = lds_uniform(sc);
    if (dec) dec = lds_uniform(dec);
    int LM;
    for (LM = 0; LM <= 3; LM++)
        if (120 << LM == frame_size) break;
    if (LM > 3) return -1;  // OPUS_BAD_ARG
    const int M = 1 << LM;
    if (len < 0 || len > 1275 || pcm == nullptr) return -1;
    const int N = M * 120;
    int32_t *mem = st->mem;
    int32_t *out_syn = mem + kDecBuf - N;
    int start = st->start, end = st->end;
    int effEnd = end > kNbBands ? kNbBands : end;
    int16_t *oldBandE = st->oldBandE, *oldLogE = st->oldLogE, *oldLogE2 = st->oldLogE2;

    const int CC = celt_cc(st);
    if (MayConceal && (data == nullptr || len <= 1)) {
        if (celt_decode_lost(st, N, LM, sc) < 0) return -3;
        out_syn = st->mem + kDecBuf - N;
        for (int c = 0; c < CC; c++)
            deemphasis(out_syn + kMemCap * c, pcm + c, N, &st->preemph_mem[c], accum, CC);
        return frame_size;
    }
    const int C = celt_c(st);
    if (st->loss_duration == 0) st->skip_plc = 0;
    if (dec == nullptr) { ec_init(&sc->dec, data, (uint32_t)len); dec = &sc->dec; }
    // The frame decodes on a register copy of the range decoder (every
    // function below that takes it is inlined, so its fields stay in scalar
    // registers); the copy goes back to its home once the frame is decoded,
    // where the caller reads the final state.
    RangeDec *const dec_home = dec;
    RangeDec dec_regs = ec_regs(dec_home);
    dec = &dec_regs;
    if (C == 1) CUO_PFOR(i, 0, kNbBands) oldBandE[i] = max16(oldBandE[i], oldBandE[kNbBands + i]);

    int32_t total_bits = len * 8;
    int32_t tell = ec_tell(dec);
    int silence;
    if (tell >= total_bits) silence = 1;
    else if (tell == 1) silence = ec_bit_logp(dec, 15);
    else silence = 0;
    if (silence) {
        tell = len * 8;
        dec->nbits_total += tell - ec_tell(dec);
    }
    int16_t postfilter_gain = 0;
    int postfilter_pitch = 0, postfilter_tapset = 0;
    if (start == 0 && tell + 16 <= total_bits) {
        if (ec_bit_logp(dec, 1)) {
            int octave = (int)ec_uint(dec, 6);
            postfilter_pitch = (16 << octave) + (int)ec_bits(dec, 4 + octave) - 1;
            int qg = (int)ec_bits(dec, 3);
            if (ec_tell(dec) + 2 <= total_bits) postfilter_tapset = ec_icdf(dec, tab::kTapsetIcdf, 2);
            postfilter_gain = (int16_t)(3072 * (qg + 1));
        }
        tell = ec_tell(dec);
    }
    int isTransient;
    if (LM > 0 && tell + 3 <= total_bits) {
        isTransient = ec_bit_logp(dec, 3);
        tell = ec_tell(dec);
    } else {
        isTransient = 0;
    }
    int shortBlocks = isTransient ? M : 0;
    int intra_ener = tell + 3 <= total_bits ? ec_bit_logp(dec, 3) : 0;
    if (!intra_ener && st->loss_duration != 0) {
        for (int c = 0; c < 2; c++) {
            int16_t safety = 0;
            int missing = imin(10, st->loss_duration >> LM);
            if (LM == 0) safety = 1536;
            else if (LM == 1) safety = 512;
            for (int i = start; i < end; i++) {
                int k = c * kNbBands + i;
                if (oldBandE[k] < max16(oldLogE[k], oldLogE2[k])) {
                    int32_t E0 = oldBandE[k], E1 = oldLogE[k], E2 = oldLogE2[k];
                    int32_t slope = max32(E1 - E0, (E2 - E0) >> 1);
                    E0 -= max32(0, (1 + missing) * slope);
                    oldBandE[k] = (int16_t)max32(-(20 << kDbShift), E0);
                } else {
                    oldBandE[k] = min16(min16(oldBandE[k], oldLogE[k]), oldLogE2[k]);
                }
                oldBandE[k] = (int16_t)(oldBandE[k] - safety);
            }
        }
    }
    {
        CUO_PROF_SCOPE(kProfCeltEnergy);
        unquant_coarse_energy(start, end, oldBandE, intra_ener, dec, LM, C);
    }
    int32_t *tf_res = sc->tf_res;
    tf_decode(start, end, isTransient, tf_res, LM, dec);
    tell = ec_tell(dec);
    int spread_decision = SPREAD_NORMAL;
    if (tell + 4 <= total_bits) spread_decision = ec_icdf(dec, tab::kSpreadIcdf, 5);
    int32_t *cap = sc->cap, *offsets = sc->offsets;
    CUO_PFOR(i, 0, kNbBands) {
        int Nb = (kEB32.v[i + 1] - kEB32.v[i]) << LM;
        cap[i] = (kCacheCaps32.v[kNbBands * (2 * LM + C - 1) + i] + 64) * C * Nb >> 2;
    }
    int dynalloc_logp = 6;
    total_bits <<= kBitRes;
    tell = (int32_t)ec_tell_frac(dec);
    for (int i = start; i < end; i++) {
        int width = C * (kEB32.v[i + 1] - kEB32.v[i]) << LM;
        int quanta = imin(width << kBitRes, imax(6 << kBitRes, width));
        int loop_logp = dynalloc_logp;
        int boost = 0;
        while (tell + (loop_logp << kBitRes) < total_bits && boost < cap[i]) {
            int flag = ec_bit_logp(dec, loop_logp);
            tell = (int32_t)ec_tell_frac(dec);
            if (!flag) break;
            boost += quanta;
            total_bits -= quanta;
            loop_logp = 1;
        }
        offsets[i] = boost;
        if (boost > 0) dynalloc_logp = imax(2, dynalloc_logp - 1);
    }
    int alloc_trim = tell + (6 << kBitRes) <= total_bits ? ec_icdf(dec, tab::kTrimIcdf, 7) : 5;
    int32_t bits = (((int32_t)len * 8) << kBitRes) - (int32_t)ec_tell_frac(dec) - 1;
    int anti_collapse_rsv = isTransient && LM >= 2 && bits >= ((LM + 2) << kBitRes) ? (1 << kBitRes) : 0;
    bits -= anti_collapse_rsv;
    int32_t balance;
    int32_t *pulses = sc->pulses, *fine_quant = sc->fine_quant, *fine_priority = sc->fine_priority;
    int codedBands, intensity = 0, dual_stereo = 0;
    {
        CUO_PROF_SCOPE(kProfCeltAlloc);
        codedBands = compute_allocation(start, end, offsets, cap, alloc_trim, bits, &balance, pulses, fine_quant,
                                        fine_priority, LM, dec, sc, C, &intensity, &dual_stereo);
    }
    unquant_fine_energy(start, end, oldBandE, fine_quant, dec, C);
    mem = advance_history(st, N);
    out_syn = mem + kDecBuf - N;
    int16_t *X = sc->X;
    {
        CUO_PROF_SCOPE(kProfCeltBands);
        quant_all_bands(start, end, X, C == 2 ? X + N : nullptr, sc->collapse, pulses, shortBlocks, spread_decision,
                        dual_stereo, intensity, tf_res, len * (8 << kBitRes) - anti_collapse_rsv, balance, dec, LM,
                        codedBands, &st->rng, CC == 1, sc);
    }
    int anti_collapse_on = 0;
    if (anti_collapse_rsv > 0) anti_collapse_on = (int)ec_bits(dec, 1);
    unquant_energy_finalise(start, end, oldBandE, fine_quant, fine_priority, len * 8 - ec_tell(dec), dec, C);
    if (anti_collapse_on)
        anti_collapse(X, sc->collapse, LM, C, N, start, end, oldBandE, oldLogE, oldLogE2, pulses, st->rng);
    if (silence)
        for (int i = 0; i < C * kNbBands; i++) oldBandE[i] = -(28 << kDbShift);
    if (st->prefilter_and_fold) prefilter_and_fold(st, N, st->etmp);
    celt_synthesis(X, out_syn, oldBandE, start, effEnd, C, CC, isTransient, LM, silence, st->freq, sc);

    st->pf_period = imax(st->pf_period, kCombMinPeriod);
    st->pf_period_old = imax(st->pf_period_old, kCombMinPeriod);
    for (int c = 0; c < CC; c++) {
        CUO_PROF_SCOPE(kProfCeltComb);
        int32_t *os = out_syn + kMemCap * c;
        comb_filter(os, os, st->pf_period_old, st->pf_period, 120, (int16_t)st->pf_gain_old,
                    (int16_t)st->pf_gain, st->pf_tapset_old, st->pf_tapset, tab::kWindow, kOverlap);
        if (LM != 0)
            comb_filter(os + 120, os + 120, st->pf_period, postfilter_pitch, N - 120,
                        (int16_t)st->pf_gain, postfilter_gain, st->pf_tapset, postfilter_tapset, tab::kWindow,
                        kOverlap);
    }
    st->pf_period_old = st->pf_period;
    st->pf_gain_old = st->pf_gain;
    st->pf_tapset_old = st->pf_tapset;
    st->pf_period = postfilter_pitch;
    st->pf_gain = postfilter_gain;
    st->pf_tapset = postfilter_tapset;
    if (LM != 0) {
        st->pf_period_old = st->pf_period;
        st->pf_gain_old = st->pf_gain;
        st->pf_tapset_old = st->pf_tapset;
    }
    if (C == 1) CUO_PFOR(i, 0, kNbBands) oldBandE[kNbBands + i] = oldBandE[i];
    if (!isTransient) {
        CUO_PFOR(i, 0, 2 * kNbBands) { oldLogE2[i] = oldLogE[i]; oldLogE[i] = oldBandE[i]; }
    } else {
        CUO_PFOR(i, 0, 2 * kNbBands) oldLogE[i] = min16(oldLogE[i], oldBandE[i]);
    }
    int16_t max_bg_inc = (int16_t)(imin(160, st->loss_duration + M) * 1);  // QCONST16(0.001, 10) == 1
    int16_t *bgE = st->backgroundLogE;
    CUO_PFOR(i, 0, 2 * kNbBands) bgE[i] = min16((int16_t)(bgE[i] + max_bg_inc), oldBandE[i]);
    CUO_PFOR(k, 0, 2 * kNbBands) {
        int i = k < kNbBands ? k : k - kNbBands;
        if (i < start || i >= end) {
            oldBandE[k] = 0;
            oldLogE[k] = oldLogE2[k] = -(28 << kDbShift);
        }
    }
    st->rng = dec->rng;
    *dec_home = dec_regs;
    for (int c = 0; c < CC; c++) deemphasis(out_syn + kMemCap * c, pcm + c, N, &st->preemph_mem[c], accum, CC);
    st->loss_duration = 0;
    st->prefilter_and_fold = 0;
    if (ec_tell(dec) > 8 * len) return -3;  // OPUS_INTERNAL_ERROR
    if (dec->error) st->error = 1;
    return frame_size;
}

CUO_FN int celt_decode(CeltState *st, const uint8_t *data, int len, int16_t *pcm, int frame_size, RangeDec *dec,
                       int accum, CeltScratch *sc)
{
    return celt_decode_frame<true>(st, data, len, pcm, frame_size, dec, accum, sc);
}

}  // namespace CUO_NS
