// silk.h -- SILK fixed-point decoder, mono.
//
// Restates libopus 1.5.2's silk_Decode (vendor/silk/dec_API.c:132-452) and its
// closure: silk_decode_frame (decode_frame.c:43-169), indices
// (decode_indices.c), pulses/shell/signs (decode_pulses.c, shell_coder.c,
// code_signs.c), parameters (decode_parameters.c, gain_quant.c:94-127,
// NLSF_decode.c, NLSF_unpack.c, NLSF_stabilize.c, NLSF2A.c, LPC_fit.c,
// LPC_inv_pred_gain.c, bwexpander*.c, decode_pitch.c), excitation + LTP/LPC
// synthesis (decode_core.c:38-237), loss concealment (PLC.c), comfort noise
// (CNG.c) and the decoder's internal -> 48 kHz resampler (resampler.c,
// resampler_private_IIR_FIR.c, resampler_private_up2_HQ.c), including the LBRR
// (in-band FEC) frame path (dec_API.c:252-289, decode_frame.c:70-72).
#pragma once
#include <type_traits>
#include "entdec.h"
#include "tables_gen.h"

namespace CUO_NS {

constexpr int kSilkMaxFrame = 320;   // 20 ms at 16 kHz
constexpr int kSilkMaxOrder = 16;
constexpr int kLtpOrder = 5;
enum { SIG_NO_VOICE = 0, SIG_UNVOICED = 1, SIG_VOICED = 2 };
enum { CODE_INDEP = 0, CODE_INDEP_NO_LTP_SCALING = 1, CODE_COND = 2 };
enum { FLAG_NORMAL = 0, FLAG_LOST = 1, FLAG_LBRR = 2 };

CUO_FN constexpr int32_t silk_fix(double x, int q) { return (int32_t)(x * (double)((int64_t)1 << q) + 0.5); }
CUO_FN int32_t silk_rand(int32_t seed) { return (int32_t)(907633515u + (uint32_t)seed * 196314165u); }

struct SilkIndices {
    int8_t GainsIndices[4];
    int8_t LTPIndex[4];
    int8_t NLSFIndices[kSilkMaxOrder + 1];
    int16_t lagIndex;
    int8_t contourIndex, signalType, quantOffsetType, NLSFInterpCoef_Q2, PERIndex, LTP_scaleIndex, Seed;
};

struct SilkResampler {
    int32_t sIIR[6];
    int16_t sFIR[8];
    int16_t delayBuf[48];
    int32_t batchSize, invRatio_Q16, Fs_in_kHz, Fs_out_kHz, inputDelay;
};

struct SilkPlcState {
    int32_t pitchL_Q8;
    int16_t LTPCoef_Q14[kLtpOrder];
    int16_t prevLPC_Q12[kSilkMaxOrder];
    int32_t last_frame_lost;
    int32_t rand_seed;
    int16_t randScale_Q14;
    int32_t conc_energy, conc_energy_shift;
    int16_t prevLTP_scale_Q14;
    int32_t prevGain_Q16[2];
    int32_t fs_kHz, nb_subfr, subfr_length;
};

struct SilkCngState {
    // CNG_exc_buf_Q14 lives outside (SilkDecoder::cng_exc, caller-owned)
    int16_t smth_NLSF_Q15[kSilkMaxOrder];
    int32_t synth_state[kSilkMaxOrder];
    int32_t smth_Gain_Q16;
    int32_t rand_seed;
    int32_t fs_kHz;
};

// A channel's excitation buffer (kSilkMaxFrame int32).  The mono build keeps
// it in the channel state (LDS); the stereo build keeps it in the caller's
// heap (HBM), which keeps two channel states inside an LDS block that allows
// three wavefronts per SIMD.  Both index and convert like the array.
struct ExcInState {
    int32_t v[kSilkMaxFrame];
    CUO_FN int32_t &operator[](int i) { return v[i]; }
    CUO_FN operator int32_t *() { return v; }
    CUO_FN operator const int32_t *() const { return v; }
    CUO_FN int32_t *heap() const { return nullptr; }
    CUO_FN void bind(int32_t *) {}
};
struct ExcInHeap {
    int32_t *v;
    CUO_FN int32_t &operator[](int i) { return v[i]; }
    CUO_FN operator int32_t *() { return v; }
    CUO_FN operator const int32_t *() const { return v; }
    CUO_FN int32_t *heap() const { return v; }
    CUO_FN void bind(int32_t *p) { v = p; }
};
using SilkExc = std::conditional_t<kMaxCh == 1, ExcInState, ExcInHeap>;

// silk_decoder_state for one channel, fields in libopus order from the reset
// point (prev_gain_Q16) on (vendor/silk/structs.h).
struct SilkChannel {
    int32_t prev_gain_Q16;
    SilkExc exc_Q14;
    int32_t sLPC_Q14_buf[kSilkMaxOrder];
    int16_t outBuf[kSilkMaxFrame + 2 * 80];
    int32_t lagPrev;
    int8_t LastGainIndex;
    int32_t fs_kHz, fs_API_hz, nb_subfr, frame_length, subfr_length, ltp_mem_length, LPC_order;
    int16_t prevNLSF_Q15[kSilkMaxOrder];
    int32_t first_frame_after_reset;
    int32_t nFramesDecoded, nFramesPerPacket;
    int32_t ec_prevSignalType;
    int16_t ec_prevLagIndex;
    int32_t VAD_flags[3];
    int32_t LBRR_flag;
    int32_t LBRR_flags[3];
    SilkResampler rs;
    SilkIndices indices;
    SilkCngState cng;
    int32_t lossCnt, prevSignalType;
    SilkPlcState plc;
};

// silk_decoder (vendor/silk/structs.h): the mid (and side) channel states
// and the stereo unmixing state.
struct SilkDecoder {
    SilkChannel ch[kMaxCh];
    int16_t sMid[2], sSide[2];     // stereo_dec_state: last two samples of mid and side
    int32_t pred_prev_Q13[2];      // stereo_dec_state: previous mid/side predictors
    int32_t nChannelsAPI, nChannelsInternal, prev_decode_only_middle;
    int32_t *cng_sig;  // caller-owned comfort-noise buffer (kept by init and reset)
    int32_t *cng_exc;  // caller-owned CNG excitation histories (kMaxCh x kSilkMaxFrame), part of the channel states
};

struct SilkCtrl {
    int32_t pitchL[4];
    int32_t Gains_Q16[4];
    int16_t PredCoef_Q12[2][kSilkMaxOrder];
    int16_t LTPCoef_Q14[kLtpOrder * 4];
    int32_t LTP_scale_Q14;
};

struct SilkScratch {
    SilkCtrl ctrl;
    int16_t pulses[kSilkMaxFrame];
    int16_t sLTP[kSilkMaxFrame];
    int32_t sLTP_Q15[2 * kSilkMaxFrame];
    int32_t res_Q14[80];
    int32_t sLPC_Q14[80 + kSilkMaxOrder];
    int16_t frame_out[kMaxCh][kSilkMaxFrame + 2];  // internal-rate frame + 2-sample history, per channel
    int16_t exc_buf[160];
    int16_t rs_buf[2 * 160 + 8];
    int32_t *cng_sig;  // comfort-noise synthesis buffer (kSilkMaxFrame + kSilkMaxOrder), caller-owned
    int32_t *cng_exc;  // CNG excitation history (kSilkMaxFrame), caller-owned
    int16_t nlsf_tmp[kSilkMaxOrder];
};

// ------------------------------------------------------------------ codebook access
struct NlsfCb {
    int nVectors, order, quantStep_Q16, invQuantStep_Q6;
    const uint8_t *CB1_Q8;
    const int16_t *CB1_Wght_Q9;
    const uint8_t *CB1_iCDF, *pred_Q8, *ec_sel, *ec_iCDF;
    const int16_t *deltaMin_Q15;
};
CUO_FN NlsfCb nlsf_cb(int wb)
{
    NlsfCb c;
    if (wb) {
        c.nVectors = tab::kNlsfWB_nVectors; c.order = tab::kNlsfWB_order;
        c.quantStep_Q16 = tab::kNlsfWB_quantStep_Q16; c.invQuantStep_Q6 = tab::kNlsfWB_invQuantStep_Q6;
        c.CB1_Q8 = tab::kNlsfWB_CB1_Q8; c.CB1_Wght_Q9 = tab::kNlsfWB_CB1_Wght_Q9; c.CB1_iCDF = tab::kNlsfWB_CB1_iCDF;
        c.pred_Q8 = tab::kNlsfWB_pred_Q8; c.ec_sel = tab::kNlsfWB_ec_sel; c.ec_iCDF = tab::kNlsfWB_ec_iCDF;
        c.deltaMin_Q15 = tab::kNlsfWB_deltaMin_Q15;
    } else {
        c.nVectors = tab::kNlsfNBMB_nVectors; c.order = tab::kNlsfNBMB_order;
        c.quantStep_Q16 = tab::kNlsfNBMB_quantStep_Q16; c.invQuantStep_Q6 = tab::kNlsfNBMB_invQuantStep_Q6;
        c.CB1_Q8 = tab::kNlsfNBMB_CB1_Q8; c.CB1_Wght_Q9 = tab::kNlsfNBMB_CB1_Wght_Q9; c.CB1_iCDF = tab::kNlsfNBMB_CB1_iCDF;
        c.pred_Q8 = tab::kNlsfNBMB_pred_Q8; c.ec_sel = tab::kNlsfNBMB_ec_sel; c.ec_iCDF = tab::kNlsfNBMB_ec_iCDF;
        c.deltaMin_Q15 = tab::kNlsfNBMB_deltaMin_Q15;
    }
    return c;
}
CUO_FN const uint8_t *ltp_gain_icdf(int per) { return per == 0 ? tab::kLtpGainIcdf0 : (per == 1 ? tab::kLtpGainIcdf1 : tab::kLtpGainIcdf2); }
CUO_FN const int8_t *ltp_vq(int per) { return per == 0 ? tab::kLtpVq0_Q7 : (per == 1 ? tab::kLtpVq1_Q7 : tab::kLtpVq2_Q7); }
CUO_FN const uint8_t *shell_table(int k)
{
    return k == 0 ? tab::kShellTable0 : (k == 1 ? tab::kShellTable1 : (k == 2 ? tab::kShellTable2 : tab::kShellTable3));
}
CUO_FN const uint8_t *pitch_contour_icdf(const SilkChannel *ch)
{
    if (ch->fs_kHz == 8) return ch->nb_subfr == 4 ? tab::kPitchContourNBIcdf : tab::kPitchContour10msNBIcdf;
    return ch->nb_subfr == 4 ? tab::kPitchContourIcdf : tab::kPitchContour10msIcdf;
}
// entries of pitch_contour_icdf(ch)
CUO_FN int pitch_contour_n(const SilkChannel *ch)
{
    if (ch->fs_kHz == 8) return ch->nb_subfr == 4 ? 11 : 3;
    return ch->nb_subfr == 4 ? 34 : 12;
}
CUO_FN const uint8_t *pitch_lag_low_icdf(const SilkChannel *ch)
{
    return ch->fs_kHz == 16 ? tab::kUniform8Icdf : (ch->fs_kHz == 12 ? tab::kUniform6Icdf : tab::kUniform4Icdf);
}

// ------------------------------------------------------------------ small DSP helpers
CUO_FN void silk_bwexpander(int16_t *ar, int d, int32_t chirp_Q16)
{
    int32_t cm1 = chirp_Q16 - 65536;
    for (int i = 0; i < d - 1; i++) {
        ar[i] = (int16_t)silk_rshift_round(silk_mul(chirp_Q16, ar[i]), 16);
        chirp_Q16 += silk_rshift_round(silk_mul(chirp_Q16, cm1), 16);
    }
    ar[d - 1] = (int16_t)silk_rshift_round(silk_mul(chirp_Q16, ar[d - 1]), 16);
}
CUO_FN void silk_bwexpander_32(int32_t *ar, int d, int32_t chirp_Q16)
{
    int32_t cm1 = chirp_Q16 - 65536;
    for (int i = 0; i < d - 1; i++) {
        ar[i] = silk_smulww(chirp_Q16, ar[i]);
        chirp_Q16 += silk_rshift_round(silk_mul(chirp_Q16, cm1), 16);
    }
    ar[d - 1] = silk_smulww(chirp_Q16, ar[d - 1]);
}

CUO_FN int32_t silk_smmul(int32_t a, int32_t b) { return (int32_t)(silk_smull(a, b) >> 32); }

CUO_FN int32_t lpc_inv_pred_gain_qa(int32_t *A, int order)
{
    const int32_t A_LIMIT = silk_fix(0.99975, 24);
    const int32_t MIN_INV = silk_fix(1.0f / 1e4f, 30);
    int32_t invGain = 1 << 30;
    int k;
    for (k = order - 1; k > 0; k--) {
        if (A[k] > A_LIMIT || A[k] < -A_LIMIT) return 0;
        int32_t rc_Q31 = -silk_lshift(A[k], 31 - 24);
        int32_t rc_mult1_Q30 = (1 << 30) - silk_smmul(rc_Q31, rc_Q31);
        invGain = silk_lshift(silk_smmul(invGain, rc_mult1_Q30), 2);
        if (invGain < MIN_INV) return 0;
        int mult2Q = 32 - silk_clz32(rc_mult1_Q30 < 0 ? -rc_mult1_Q30 : rc_mult1_Q30);
        int32_t rc_mult2 = silk_inverse32_varq(rc_mult1_Q30, mult2Q + 30);
        for (int n = 0; n < (k + 1) >> 1; n++) {
            int32_t tmp1 = A[n], tmp2 = A[k - n - 1];
            int64_t t64 = silk_rshift_round64(silk_smull(silk_sub_sat32(tmp1,
                              (int32_t)silk_rshift_round64(silk_smull(tmp2, rc_Q31), 31)), rc_mult2), mult2Q);
            if (t64 > INT32_MAX || t64 < INT32_MIN) return 0;
            A[n] = (int32_t)t64;
            t64 = silk_rshift_round64(silk_smull(silk_sub_sat32(tmp2,
                      (int32_t)silk_rshift_round64(silk_smull(tmp1, rc_Q31), 31)), rc_mult2), mult2Q);
            if (t64 > INT32_MAX || t64 < INT32_MIN) return 0;
            A[k - n - 1] = (int32_t)t64;
        }
    }
    if (A[k] > A_LIMIT || A[k] < -A_LIMIT) return 0;
    int32_t rc_Q31 = -silk_lshift(A[0], 31 - 24);
    int32_t rc_mult1_Q30 = (1 << 30) - silk_smmul(rc_Q31, rc_Q31);
    invGain = silk_lshift(silk_smmul(invGain, rc_mult1_Q30), 2);
    if (invGain < MIN_INV) return 0;
    return invGain;
}

CUO_FN int32_t silk_lpc_inverse_pred_gain(const int16_t *A_Q12, int order)
{
    int32_t Atmp[kSilkMaxOrder];
    int32_t dc = 0;
    for (int k = 0; k < order; k++) {
        dc += A_Q12[k];
        Atmp[k] = silk_lshift((int32_t)A_Q12[k], 24 - 12);
    }
    if (dc >= 4096) return 0;
    return lpc_inv_pred_gain_qa(Atmp, order);
}

CUO_FN void silk_lpc_fit(int16_t *a_QOUT, int32_t *a_QIN, int QOUT, int QIN, int d)
{
    int i, k, idx = 0;
    for (i = 0; i < 10; i++) {
        int32_t maxabs = 0;
        for (k = 0; k < d; k++) {
            int32_t absval = a_QIN[k] > 0 ? a_QIN[k] : -a_QIN[k];
            if (absval > maxabs) { maxabs = absval; idx = k; }
        }
        maxabs = silk_rshift_round(maxabs, QIN - QOUT);
        if (maxabs > 32767) {
            maxabs = min32(maxabs, 163838);
            int32_t chirp_Q16 = silk_fix(0.999, 16) - silk_lshift(maxabs - 32767, 14) / ((maxabs * (idx + 1)) >> 2);
            silk_bwexpander_32(a_QIN, d, chirp_Q16);
        } else {
            break;
        }
    }
    if (i == 10) {
        for (k = 0; k < d; k++) {
            a_QOUT[k] = (int16_t)silk_sat16(silk_rshift_round(a_QIN[k], QIN - QOUT));
            a_QIN[k] = silk_lshift((int32_t)a_QOUT[k], QIN - QOUT);
        }
    } else {
        for (k = 0; k < d; k++) a_QOUT[k] = (int16_t)silk_rshift_round(a_QIN[k], QIN - QOUT);
    }
}

CUO_FN void nlsf2a_find_poly(int32_t *out, const int32_t *cLSF, int dd)
{
    out[0] = 1 << 16;
    out[1] = -cLSF[0];
    for (int k = 1; k < dd; k++) {
        int32_t ftmp = cLSF[2 * k];
        out[k + 1] = silk_lshift(out[k - 1], 1) - (int32_t)silk_rshift_round64(silk_smull(ftmp, out[k]), 16);
        for (int n = k; n > 1; n--) out[n] += out[n - 2] - (int32_t)silk_rshift_round64(silk_smull(ftmp, out[n - 1]), 16);
        out[1] -= ftmp;
    }
}

static constexpr uint8_t kNlsfOrdering16[16] = {0, 15, 8, 7, 4, 11, 12, 3, 2, 13, 10, 5, 6, 9, 14, 1};
static constexpr uint8_t kNlsfOrdering10[10] = {0, 9, 6, 3, 4, 5, 8, 1, 2, 7};
// inverse permutations: position j of cos_LSF_QA holds NLSF kNlsfInv*[j]
static constexpr uint8_t kNlsfInv16[16] = {0, 15, 8, 7, 4, 11, 12, 3, 2, 13, 10, 5, 6, 9, 14, 1};
static constexpr uint8_t kNlsfInv10[10] = {0, 7, 8, 3, 4, 5, 2, 9, 6, 1};

// silk_NLSF2A (NLSF2A.c:66-141)
CUO_FN void silk_nlsf2a(int16_t *a_Q12, const int16_t *NLSF, int d)
{
    CUO_PROF_SCOPE(kProfSilkNlsf2a);
    const uint8_t *ordering = d == 16 ? kNlsfOrdering16 : kNlsfOrdering10;
    int32_t cos_LSF_QA[kSilkMaxOrder];
    int32_t P[kSilkMaxOrder / 2 + 1], Q[kSilkMaxOrder / 2 + 1];
    int32_t a32_QA1[kSilkMaxOrder];
    for (int k = 0; k < d; k++) {
        int32_t f_int = NLSF[k] >> (15 - 7);
        int32_t f_frac = NLSF[k] - silk_lshift(f_int, 15 - 7);
        int32_t cos_val = tab::kLsfCos_Q12[f_int];
        int32_t delta = tab::kLsfCos_Q12[f_int + 1] - cos_val;
        cos_LSF_QA[ordering[k]] = silk_rshift_round(silk_lshift(cos_val, 8) + silk_mul(delta, f_frac), 20 - 16);
    }
    int dd = d >> 1;
    nlsf2a_find_poly(P, &cos_LSF_QA[0], dd);
    nlsf2a_find_poly(Q, &cos_LSF_QA[1], dd);
    for (int k = 0; k < dd; k++) {
        int32_t Ptmp = P[k + 1] + P[k];
        int32_t Qtmp = Q[k + 1] - Q[k];
        a32_QA1[k] = -Qtmp - Ptmp;
        a32_QA1[d - k - 1] = Qtmp - Ptmp;
    }
    silk_lpc_fit(a_Q12, a32_QA1, 12, 17, d);
    for (int i = 0; silk_lpc_inverse_pred_gain(a_Q12, d) == 0 && i < 16; i++) {
        silk_bwexpander_32(a32_QA1, d, 65536 - silk_lshift(2, i));
        for (int k = 0; k < d; k++) a_Q12[k] = (int16_t)silk_rshift_round(a32_QA1[k], 17 - 12);
    }
}

CUO_FN void silk_nlsf_unpack(int16_t *ec_ix, uint8_t *pred_Q8, const NlsfCb &cb, int CB1_index)
{
    const uint8_t *sel = &cb.ec_sel[CB1_index * cb.order / 2];
    for (int i = 0; i < cb.order; i += 2) {
        uint8_t entry = *sel++;
        ec_ix[i] = (int16_t)silk_smulbb((entry >> 1) & 7, 2 * 4 + 1);
        pred_Q8[i] = cb.pred_Q8[i + (entry & 1) * (cb.order - 1)];
        ec_ix[i + 1] = (int16_t)silk_smulbb((entry >> 5) & 7, 2 * 4 + 1);
        pred_Q8[i + 1] = cb.pred_Q8[i + ((entry >> 4) & 1) * (cb.order - 1) + 1];
    }
}

CUO_FN void silk_nlsf_stabilize(int16_t *NLSF_Q15, const int16_t *dmin, int L)
{
    int loops, I = 0;
    for (loops = 0; loops < 20; loops++) {
        int32_t min_diff = NLSF_Q15[0] - dmin[0];
        I = 0;
        for (int i = 1; i <= L - 1; i++) {
            int32_t diff = NLSF_Q15[i] - (NLSF_Q15[i - 1] + dmin[i]);
            if (diff < min_diff) { min_diff = diff; I = i; }
        }
        int32_t diff = (1 << 15) - (NLSF_Q15[L - 1] + dmin[L]);
        if (diff < min_diff) { min_diff = diff; I = L; }
        if (min_diff >= 0) return;
        if (I == 0) {
            NLSF_Q15[0] = dmin[0];
        } else if (I == L) {
            NLSF_Q15[L - 1] = (int16_t)((1 << 15) - dmin[L]);
        } else {
            int32_t min_center = 0;
            for (int k = 0; k < I; k++) min_center += dmin[k];
            min_center += dmin[I] >> 1;
            int32_t max_center = 1 << 15;
            for (int k = L; k > I; k--) max_center -= dmin[k];
            max_center -= dmin[I] >> 1;
            int16_t center = (int16_t)silk_limit(silk_rshift_round((int32_t)NLSF_Q15[I - 1] + (int32_t)NLSF_Q15[I], 1),
                                                 min_center, max_center);
            NLSF_Q15[I - 1] = (int16_t)(center - (dmin[I] >> 1));
            NLSF_Q15[I] = (int16_t)(NLSF_Q15[I - 1] + dmin[I]);
        }
    }
    if (loops == 20) {
        for (int i = 1; i < L; i++) {  // insertion sort, increasing
            int16_t v = NLSF_Q15[i];
            int j;
            for (j = i - 1; j >= 0 && v < NLSF_Q15[j]; j--) NLSF_Q15[j + 1] = NLSF_Q15[j];
            NLSF_Q15[j + 1] = v;
        }
        NLSF_Q15[0] = (int16_t)imax(NLSF_Q15[0], dmin[0]);
        for (int i = 1; i < L; i++)
            NLSF_Q15[i] = (int16_t)imax(NLSF_Q15[i], silk_sat16((int32_t)NLSF_Q15[i - 1] + dmin[i]));
        NLSF_Q15[L - 1] = (int16_t)imin(NLSF_Q15[L - 1], (1 << 15) - dmin[L]);
        for (int i = L - 2; i >= 0; i--) NLSF_Q15[i] = (int16_t)imin(NLSF_Q15[i], NLSF_Q15[i + 1] - dmin[i + 1]);
    }
}

CUO_FN void silk_nlsf_decode(int16_t *pNLSF_Q15, const int8_t *NLSFIndices, const NlsfCb &cb)
{
    CUO_PROF_SCOPE(kProfSilkNlsf);
    uint8_t pred_Q8[kSilkMaxOrder];
    int16_t ec_ix[kSilkMaxOrder];
    int16_t res_Q10[kSilkMaxOrder];
    silk_nlsf_unpack(ec_ix, pred_Q8, cb, NLSFIndices[0]);
    // residual dequantisation
    int32_t out_Q10 = 0;
    for (int i = cb.order - 1; i >= 0; i--) {
        int32_t pred_Q10 = silk_smulbb(out_Q10, (int16_t)pred_Q8[i]) >> 8;
        out_Q10 = silk_lshift(NLSFIndices[1 + i], 10);
        if (out_Q10 > 0) out_Q10 = out_Q10 - silk_fix(0.1, 10);
        else if (out_Q10 < 0) out_Q10 = out_Q10 + silk_fix(0.1, 10);
        out_Q10 = silk_smlawb(pred_Q10, out_Q10, cb.quantStep_Q16);
        res_Q10[i] = (int16_t)out_Q10;
    }
    const uint8_t *pCB = &cb.CB1_Q8[NLSFIndices[0] * cb.order];
    const int16_t *pW = &cb.CB1_Wght_Q9[NLSFIndices[0] * cb.order];
    for (int i = 0; i < cb.order; i++) {
        int32_t t = silk_lshift((int32_t)res_Q10[i], 14) / pW[i] + silk_lshift((int16_t)pCB[i], 7);
        pNLSF_Q15[i] = (int16_t)silk_limit(t, 0, 32767);
    }
    silk_nlsf_stabilize(pNLSF_Q15, cb.deltaMin_Q15, cb.order);
}

CUO_FN int32_t silk_log2lin(int32_t inLog_Q7)
{
    if (inLog_Q7 < 0) return 0;
    if (inLog_Q7 >= 3967) return INT32_MAX;
    int32_t out = silk_lshift(1, inLog_Q7 >> 7);
    int32_t frac = inLog_Q7 & 0x7F;
    if (inLog_Q7 < 2048)
        out = out + (silk_mul(out, silk_smlawb(frac, silk_smulbb(frac, 128 - frac), -174)) >> 7);
    else
        out = silk_mla(out, out >> 7, silk_smlawb(frac, silk_smulbb(frac, 128 - frac), -174));
    return out;
}

CUO_FN void silk_gains_dequant(int32_t *gain_Q16, const int8_t *ind, int8_t *prev_ind, int conditional, int nb_subfr)
{
    const int32_t OFFSET = (2 * 128) / 6 + 16 * 128;
    const int32_t INV_SCALE_Q16 = (65536 * (((88 - 2) * 128) / 6)) / (64 - 1);
    for (int k = 0; k < nb_subfr; k++) {
        if (k == 0 && conditional == 0) {
            *prev_ind = (int8_t)imax(ind[k], *prev_ind - 16);
        } else {
            int ind_tmp = ind[k] + -4;
            int thr = 2 * 36 - 64 + *prev_ind;
            if (ind_tmp > thr) *prev_ind = (int8_t)(*prev_ind + silk_lshift(ind_tmp, 1) - thr);
            else *prev_ind = (int8_t)(*prev_ind + ind_tmp);
        }
        *prev_ind = (int8_t)silk_limit(*prev_ind, 0, 63);
        gain_Q16[k] = silk_log2lin(min32(silk_smulwb(INV_SCALE_Q16, *prev_ind) + OFFSET, 3967));
    }
}

CUO_FN void silk_decode_pitch(int lagIndex, int contourIndex, int32_t *pitch_lags, int Fs_kHz, int nb_subfr)
{
    const int8_t *cb;
    int cbk_size;
    if (Fs_kHz == 8) {
        if (nb_subfr == 4) { cb = tab::kCbLagsStage2; cbk_size = 11; }
        else { cb = tab::kCbLagsStage2_10ms; cbk_size = 3; }
    } else {
        if (nb_subfr == 4) { cb = tab::kCbLagsStage3; cbk_size = 34; }
        else { cb = tab::kCbLagsStage3_10ms; cbk_size = 12; }
    }
    int min_lag = silk_smulbb(2, Fs_kHz), max_lag = silk_smulbb(18, Fs_kHz);
    int lag = min_lag + lagIndex;
    for (int k = 0; k < nb_subfr; k++) {
        pitch_lags[k] = lag + cb[k * cbk_size + contourIndex];
        pitch_lags[k] = silk_limit(pitch_lags[k], min_lag, max_lag);
    }
}

CUO_FN void silk_lpc_analysis_filter(int16_t *out, const int16_t *in, const int16_t *B, int len, int d)
{
    CUO_PFOR(ix, d, len) {
        const int16_t *p = &in[ix - 1];
        int32_t acc = silk_smulbb(p[0], B[0]);
        for (int j = 1; j < d; j++) acc = add32_ovf(acc, (int32_t)p[-j] * (int32_t)B[j]);
        acc = sub32_ovf(silk_lshift((int32_t)p[1], 12), acc);
        out[ix] = (int16_t)silk_sat16(silk_rshift_round(acc, 12));
    }
    CUO_PFOR(j, 0, d) out[j] = 0;
}

CUO_FN void silk_sum_sqr_shift(int32_t *energy, int32_t *shift, const int16_t *x, int len)
{
    int shft = 31 - silk_clz32(len);
    int32_t nrg = len;
    int i;
    for (i = 0; i < len - 1; i += 2) {
        uint32_t t = (uint32_t)silk_smulbb(x[i], x[i]);
        t = t + (uint32_t)((int32_t)x[i + 1] * (int32_t)x[i + 1]);
        nrg = (int32_t)((uint32_t)nrg + (t >> shft));
    }
    if (i < len) {
        uint32_t t = (uint32_t)silk_smulbb(x[i], x[i]);
        nrg = (int32_t)((uint32_t)nrg + (t >> shft));
    }
    shft = max32(0, shft + 3 - silk_clz32(nrg));
    nrg = 0;
    for (i = 0; i < len - 1; i += 2) {
        uint32_t t = (uint32_t)silk_smulbb(x[i], x[i]);
        t = t + (uint32_t)((int32_t)x[i + 1] * (int32_t)x[i + 1]);
        nrg = (int32_t)((uint32_t)nrg + (t >> shft));
    }
    if (i < len) {
        uint32_t t = (uint32_t)silk_smulbb(x[i], x[i]);
        nrg = (int32_t)((uint32_t)nrg + (t >> shft));
    }
    *shift = shft;
    *energy = nrg;
}

// ------------------------------------------------------------------ resampler (decoder: 8/12/16 -> 48 kHz)
CUO_FN void silk_resampler_init(SilkResampler *S, int32_t Fs_in, int32_t Fs_out)
{
    __builtin_memset(S, 0, sizeof(SilkResampler));  // memset: ordered before the field stores below
    const int in_id = Fs_in == 8000 ? 0 : (Fs_in == 12000 ? 1 : 2);
    static constexpr int8_t kDelayTo48[3] = {0, 4, 7};  // delay_matrix_dec[.][48 kHz]
    S->inputDelay = kDelayTo48[in_id];
    S->Fs_in_kHz = Fs_in / 1000;
    S->Fs_out_kHz = Fs_out / 1000;
    S->batchSize = S->Fs_in_kHz * 10;
    // 48 kHz is never 2x of 8/12/16 kHz: always the IIR_FIR path (up2x = 1)
    S->invRatio_Q16 = silk_lshift(silk_lshift(Fs_in, 14 + 1) / Fs_out, 2);
    while (silk_smulww(S->invRatio_Q16, Fs_out) < silk_lshift(Fs_in, 1)) S->invRatio_Q16++;
}

// 2x allpass upsampler (resampler_private_up2_HQ.c:38-113).  The recursion is
// serial; inputs are read a block of kLanes at a time (one per lane) and the
// even/odd outputs collected the same way, so the loop touches no memory.
CUO_FN void silk_up2_hq(int32_t *S, int16_t *out, const int16_t *in, int len)
{
    const int32_t h00 = tab::kResUp2HQ0[0], h01 = tab::kResUp2HQ0[1], h02 = tab::kResUp2HQ0[2];
    const int32_t h10 = tab::kResUp2HQ1[0], h11 = tab::kResUp2HQ1[1], h12 = tab::kResUp2HQ1[2];
    int32_t s0 = uniform(S[0]), s1 = uniform(S[1]), s2 = uniform(S[2]);
    int32_t s3 = uniform(S[3]), s4 = uniform(S[4]), s5 = uniform(S[5]);
    for (int k0 = 0; k0 < len; k0 += kLanes) {
        const int cnt = min32(kLanes, len - k0);
        const int32_t iv = lane_id() < cnt ? (int32_t)in[k0 + lane_id()] : 0;
        int32_t oe = 0, oo = 0;
        for (int k = 0; k < cnt; k++) {
            const int32_t in32 = silk_lshift(lane_get(iv, k), 10);
            int32_t Y = in32 - s0;
            int32_t X = silk_smulwb(Y, h00);
            int32_t o1 = s0 + X;
            s0 = in32 + X;
            Y = o1 - s1;
            X = silk_smulwb(Y, h01);
            int32_t o2 = s1 + X;
            s1 = o1 + X;
            Y = o2 - s2;
            X = silk_smlawb(Y, Y, h02);
            o1 = s2 + X;
            s2 = o2 + X;
            oe = lane_set_s(oe, k, silk_sat16(silk_rshift_round(o1, 10)));
            Y = in32 - s3;
            X = silk_smulwb(Y, h10);
            o1 = s3 + X;
            s3 = in32 + X;
            Y = o1 - s4;
            X = silk_smulwb(Y, h11);
            o2 = s4 + X;
            s4 = o1 + X;
            Y = o2 - s5;
            X = silk_smlawb(Y, Y, h12);
            o1 = s5 + X;
            s5 = o2 + X;
            oo = lane_set_s(oo, k, silk_sat16(silk_rshift_round(o1, 10)));
        }
        CUO_PFOR(k, 0, cnt) {
            out[2 * (k0 + k)] = (int16_t)oe;
            out[2 * (k0 + k) + 1] = (int16_t)oo;
        }
    }
    S[0] = s0; S[1] = s1; S[2] = s2;
    S[3] = s3; S[4] = s4; S[5] = s5;
}

// Fractional FIR interpolation (resampler_private_IIR_FIR.c:36-61): every
// output is independent, one per lane.
CUO_FN int16_t *silk_iir_fir_interpol(int16_t *out, const int16_t *buf, int32_t max_index_Q16, int32_t inc_Q16)
{
    const int16_t *F = tab::kResFracFir12;
    const int n = max_index_Q16 > 0 ? (max_index_Q16 + inc_Q16 - 1) / inc_Q16 : 0;
    CUO_PFOR(j, 0, n) {
        const int32_t index_Q16 = j * inc_Q16;
        int32_t ti = silk_smulwb(index_Q16 & 0xFFFF, 12);
        const int16_t *b = &buf[index_Q16 >> 16];
        int32_t r = silk_smulbb(b[0], F[ti * 4 + 0]);
        r = silk_smlabb(r, b[1], F[ti * 4 + 1]);
        r = silk_smlabb(r, b[2], F[ti * 4 + 2]);
        r = silk_smlabb(r, b[3], F[ti * 4 + 3]);
        r = silk_smlabb(r, b[4], F[(11 - ti) * 4 + 3]);
        r = silk_smlabb(r, b[5], F[(11 - ti) * 4 + 2]);
        r = silk_smlabb(r, b[6], F[(11 - ti) * 4 + 1]);
        r = silk_smlabb(r, b[7], F[(11 - ti) * 4 + 0]);
        out[j] = (int16_t)silk_sat16(silk_rshift_round(r, 15));
    }
    return out + n;
}

CUO_FN void silk_resampler_iir_fir(SilkResampler *S, int16_t *out, const int16_t *in, int32_t inLen, int16_t *buf)
{
    CUO_PROF_SCOPE(kProfSilkResamp);
    CUO_PFOR(i, 0, 8) buf[i] = S->sFIR[i];
    int32_t inc = S->invRatio_Q16;
    int32_t nIn;
    while (1) {
        nIn = min32(inLen, S->batchSize);
        silk_up2_hq(S->sIIR, &buf[8], in, nIn);
        int32_t max_index_Q16 = silk_lshift(nIn, 16 + 1);
        out = silk_iir_fir_interpol(out, buf, max_index_Q16, inc);
        in += nIn;
        inLen -= nIn;
        if (inLen > 0) { CUO_PFOR(i, 0, 8) buf[i] = buf[(nIn << 1) + i]; }
        else break;
    }
    CUO_PFOR(i, 0, 8) S->sFIR[i] = buf[(nIn << 1) + i];
}

CUO_FN void silk_resample(SilkResampler *S, int16_t *out, const int16_t *in, int32_t inLen, int16_t *buf)
{
    int nSamples = S->Fs_in_kHz - S->inputDelay;
    CUO_PFOR(i, 0, nSamples) S->delayBuf[S->inputDelay + i] = in[i];
    silk_resampler_iir_fir(S, out, S->delayBuf, S->Fs_in_kHz, buf);
    silk_resampler_iir_fir(S, &out[S->Fs_out_kHz], &in[nSamples], inLen - S->Fs_in_kHz, buf);
    CUO_PFOR(i, 0, S->inputDelay) S->delayBuf[i] = in[inLen - S->inputDelay + i];
}

// ------------------------------------------------------------------ state management
CUO_FN void silk_cng_reset(SilkChannel *ch)
{
    int32_t step = 32767 / (ch->LPC_order + 1);
    int32_t acc = 0;
    for (int i = 0; i < ch->LPC_order; i++) { acc += step; ch->cng.smth_NLSF_Q15[i] = (int16_t)acc; }
    ch->cng.smth_Gain_Q16 = 0;
    ch->cng.rand_seed = 3176576;
}

CUO_FN void silk_plc_reset(SilkChannel *ch)
{
    ch->plc.pitchL_Q8 = silk_lshift(ch->frame_length, 8 - 1);
    ch->plc.prevGain_Q16[0] = 1 << 16;
    ch->plc.prevGain_Q16[1] = 1 << 16;
    ch->plc.subfr_length = 20;
    ch->plc.nb_subfr = 2;
}

CUO_FN void silk_reset_channel(SilkChannel *ch)
{
    int32_t *exc = ch->exc_Q14.heap();  // a heap excitation buffer survives the wipe, cleared
    // memset, not a loop of int32 stores: those could be ordered after the
    // field stores below, which type-based alias analysis lets the compiler
    // assume they do not overlap
    __builtin_memset(ch, 0, sizeof(SilkChannel));
    if (exc) {
        ch->exc_Q14.bind(exc);
        CUO_PFOR(i, 0, kSilkMaxFrame) exc[i] = 0;
    }
    ch->first_frame_after_reset = 1;
    ch->prev_gain_Q16 = 65536;
    silk_cng_reset(ch);
    silk_plc_reset(ch);
}

// The CNG excitation history is channel state kept outside the channel
// struct: a channel reset clears it too.
CUO_FN void silk_reset_channel_of(SilkDecoder *d, int n)
{
    silk_reset_channel(&d->ch[n]);
    int32_t *e = d->cng_exc + kSilkMaxFrame * n;
    CUO_PFOR(i, 0, kSilkMaxFrame) e[i] = 0;
}

CUO_FN void silk_reset_stereo(SilkDecoder *d)
{
    d->sMid[0] = d->sMid[1] = 0;
    d->sSide[0] = d->sSide[1] = 0;
    d->pred_prev_Q13[0] = d->pred_prev_Q13[1] = 0;
    d->prev_decode_only_middle = 0;
}

// silk_InitDecoder (dec_API.c:59-75), on a zeroed decoder
CUO_FN void silk_init_decoder(SilkDecoder *d)
{
    for (int n = 0; n < kMaxCh; n++) silk_reset_channel_of(d, n);
    silk_reset_stereo(d);
    d->nChannelsAPI = 0;
    d->nChannelsInternal = 0;
}

// silk_ResetDecoder (dec_API.c:77-92): resets channel states and the stereo
// memory but keeps nChannelsAPI / nChannelsInternal.
CUO_FN void silk_reset_decoder(SilkDecoder *d)
{
    for (int n = 0; n < kMaxCh; n++) silk_reset_channel_of(d, n);
    silk_reset_stereo(d);
}

CUO_FN void silk_set_fs(SilkChannel *ch, int fs_kHz, int32_t fs_API_Hz)
{
    ch->subfr_length = silk_smulbb(5, fs_kHz);
    int frame_length = silk_smulbb(ch->nb_subfr, ch->subfr_length);
    if (ch->fs_kHz != fs_kHz || ch->fs_API_hz != fs_API_Hz) {
        silk_resampler_init(&ch->rs, silk_smulbb(fs_kHz, 1000), fs_API_Hz);
        ch->fs_API_hz = fs_API_Hz;
    }
    if (ch->fs_kHz != fs_kHz || frame_length != ch->frame_length) {
        if (ch->fs_kHz != fs_kHz) {
            ch->ltp_mem_length = silk_smulbb(20, fs_kHz);
            ch->LPC_order = (fs_kHz == 8 || fs_kHz == 12) ? 10 : 16;
            ch->first_frame_after_reset = 1;
            ch->lagPrev = 100;
            ch->LastGainIndex = 10;
            ch->prevSignalType = SIG_NO_VOICE;
            for (int i = 0; i < kSilkMaxFrame + 2 * 80; i++) ch->outBuf[i] = 0;
            for (int i = 0; i < kSilkMaxOrder; i++) ch->sLPC_Q14_buf[i] = 0;
        }
        ch->fs_kHz = fs_kHz;
        ch->frame_length = frame_length;
    }
}

// ------------------------------------------------------------------ bitstream
// ec_icdf for a table of n entries (its last one 0).  On the device a table
// of more than four entries is searched in one step: lane j reads entry j
// (one vector load), forms r * icdf[j], and the symbol is the lowest lane with
// val >= r * icdf[j] -- where the serial search stops.
#if CUO_LANE_PATHS
__device__ __forceinline__ int silk_icdf(RangeDec *d, const uint8_t *icdf, int n)
{
    if (n <= 4) return ec_icdf(d, icdf, 8);
    const uint32_t s = uniform(d->rng), v = uniform(d->val);
    const uint32_t r = s >> 8;
    const int j = lane_id();
    const uint32_t tj = r * (uint32_t)icdf[j < n ? j : n - 1];
    const uint64_t m = lanes_u<kCmpUge>(v, tj) & lanes_s<kCmpSlt>(j, n);
    const int ret = __builtin_ctzll(m);
    const uint32_t sv = (uint32_t)lane_get((int32_t)tj, ret);
    const uint32_t tv = ret ? (uint32_t)lane_get((int32_t)tj, ret - 1) : s;
    ec_normalize_rv(d, tv - sv, v - sv);
    return ret;
}
#else
CUO_FN int silk_icdf(RangeDec *d, const uint8_t *icdf, int) { return ec_icdf(d, icdf, 8); }
#endif

CUO_FN void silk_decode_indices(SilkChannel *ch, RangeDec *dp, int FrameIndex, int decode_LBRR, int condCoding)
{
    CUO_PROF_SCOPE(kProfSilkIndices);
    ch = lds_uniform(ch);
    dp = lds_uniform(dp);
    RangeDec dl = ec_regs(dp);  // decode on a register copy of the range decoder
    RangeDec *d = &dl;
    SilkIndices &ix = ch->indices;
    int Ix;
    if (decode_LBRR || ch->VAD_flags[FrameIndex]) Ix = ec_icdf(d, tab::kTypeOffsetVadIcdf, 8) + 2;
    else Ix = ec_icdf(d, tab::kTypeOffsetNoVadIcdf, 8);
    ix.signalType = (int8_t)(Ix >> 1);
    ix.quantOffsetType = (int8_t)(Ix & 1);
    if (condCoding == CODE_COND) {
        ix.GainsIndices[0] = (int8_t)silk_icdf(d, tab::kDeltaGainIcdf, 41);
    } else {
        ix.GainsIndices[0] = (int8_t)silk_lshift(silk_icdf(d, tab::kGainIcdf + 8 * ix.signalType, 8), 3);
        ix.GainsIndices[0] = (int8_t)(ix.GainsIndices[0] + silk_icdf(d, tab::kUniform8Icdf, 8));
    }
    for (int i = 1; i < ch->nb_subfr; i++) ix.GainsIndices[i] = (int8_t)silk_icdf(d, tab::kDeltaGainIcdf, 41);
    NlsfCb cb = nlsf_cb(ch->fs_kHz == 16);
    const int cb1 = silk_icdf(d, &cb.CB1_iCDF[(ix.signalType >> 1) * cb.nVectors], cb.nVectors);
    ix.NLSFIndices[0] = (int8_t)cb1;
    // entropy-table selector of coefficient i (silk_NLSF_unpack, NLSF_unpack.c:35-55)
    const uint8_t *sel = &cb.ec_sel[cb1 * cb.order / 2];
    for (int i = 0; i < cb.order; i++) {
        const int entry = (int)ld_u8(sel, i >> 1);
        const int ec_ix = silk_smulbb((entry >> ((i & 1) ? 5 : 1)) & 7, 2 * 4 + 1);
        Ix = silk_icdf(d, &cb.ec_iCDF[ec_ix], 2 * 4 + 1);
        if (Ix == 0) Ix -= silk_icdf(d, tab::kNlsfExtIcdf, 7);
        else if (Ix == 2 * 4) Ix += silk_icdf(d, tab::kNlsfExtIcdf, 7);
        ix.NLSFIndices[i + 1] = (int8_t)(Ix - 4);
    }
    if (ch->nb_subfr == 4) ix.NLSFInterpCoef_Q2 = (int8_t)silk_icdf(d, tab::kNlsfInterpFactorIcdf, 5);
    else ix.NLSFInterpCoef_Q2 = 4;
    if (ix.signalType == SIG_VOICED) {
        int decode_absolute = 1;
        if (condCoding == CODE_COND && ch->ec_prevSignalType == SIG_VOICED) {
            int delta = (int16_t)silk_icdf(d, tab::kPitchDeltaIcdf, 21);
            if (delta > 0) {
                delta = delta - 9;
                ix.lagIndex = (int16_t)(ch->ec_prevLagIndex + delta);
                decode_absolute = 0;
            }
        }
        if (decode_absolute) {
            ix.lagIndex = (int16_t)(silk_icdf(d, tab::kPitchLagIcdf, 32) * (ch->fs_kHz >> 1));
            ix.lagIndex = (int16_t)(ix.lagIndex + silk_icdf(d, pitch_lag_low_icdf(ch), ch->fs_kHz >> 1));
        }
        ch->ec_prevLagIndex = ix.lagIndex;
        ix.contourIndex = (int8_t)silk_icdf(d, pitch_contour_icdf(ch), pitch_contour_n(ch));
        ix.PERIndex = (int8_t)ec_icdf(d, tab::kLtpPerIndexIcdf, 8);
        for (int k = 0; k < ch->nb_subfr; k++)
            ix.LTPIndex[k] = (int8_t)silk_icdf(d, ltp_gain_icdf(ix.PERIndex), 8 << ix.PERIndex);
        if (condCoding == CODE_INDEP) ix.LTP_scaleIndex = (int8_t)ec_icdf(d, tab::kLtpScaleIcdf, 8);
        else ix.LTP_scaleIndex = 0;
    }
    ch->ec_prevSignalType = ix.signalType;
    ix.Seed = (int8_t)ec_icdf(d, tab::kUniform4Icdf, 8);
    *dp = dl;
}

CUO_FN void silk_shell_split(int16_t *c1, int16_t *c2, RangeDec *d, int p, const uint8_t *tabl)
{
    if (p > 0) {
        c1[0] = (int16_t)ec_icdf(d, &tabl[tab::kShellTableOffsets[p]], 8);
        c2[0] = (int16_t)(p - c1[0]);
    } else {
        c1[0] = 0;
        c2[0] = 0;
    }
}

CUO_FN void silk_shell_decoder(int16_t *p0, RangeDec *d, int pulses4)
{
    int16_t p3[2], p2[4], p1[8];
    const uint8_t *t0 = tab::kShellTable0, *t1 = tab::kShellTable1, *t2 = tab::kShellTable2, *t3 = tab::kShellTable3;
    silk_shell_split(&p3[0], &p3[1], d, pulses4, t3);
    silk_shell_split(&p2[0], &p2[1], d, p3[0], t2);
    silk_shell_split(&p1[0], &p1[1], d, p2[0], t1);
    silk_shell_split(&p0[0], &p0[1], d, p1[0], t0);
    silk_shell_split(&p0[2], &p0[3], d, p1[1], t0);
    silk_shell_split(&p1[2], &p1[3], d, p2[1], t1);
    silk_shell_split(&p0[4], &p0[5], d, p1[2], t0);
    silk_shell_split(&p0[6], &p0[7], d, p1[3], t0);
    silk_shell_split(&p2[2], &p2[3], d, p3[1], t2);
    silk_shell_split(&p1[4], &p1[5], d, p2[2], t1);
    silk_shell_split(&p0[8], &p0[9], d, p1[4], t0);
    silk_shell_split(&p0[10], &p0[11], d, p1[5], t0);
    silk_shell_split(&p1[6], &p1[7], d, p2[3], t1);
    silk_shell_split(&p0[12], &p0[13], d, p1[6], t0);
    silk_shell_split(&p0[14], &p0[15], d, p1[7], t0);
}

#if CUO_LANE_PATHS
// ec_icdf over a lane-held table starting at byte off: the same search as
// ec_icdf_t, its words read with v_readlane instead of scalar loads.
__device__ __forceinline__ int ec_icdf_lt(RangeDec *d, int32_t tabv, int off, int ftb)
{
    uint32_t s = uniform(d->rng), v = uniform(d->val);
    const uint32_t r = s >> ftb;
    int ret = 0;
    uint32_t t;
    int w = off >> 2;
    uint32_t word = (uint32_t)lane_get(tabv, w) >> (8 * (off & 3));
    int left = 4 - (off & 3);
    for (;;) {
        t = s;
        s = r * (word & 0xFFu);
        if (v >= s) break;
        ret++;
        if (--left == 0) {
            word = (uint32_t)lane_get(tabv, ++w);
            left = 4;
        } else {
            word >>= 8;
        }
    }
    ec_normalize_rv(d, t - s, v - s);
    return ret;
}
// ec_icdf over a lane-held table in one step: lane j takes entry j (byte
// off + j, fetched from the lane holding its word), forms r * icdf[j], and the
// symbol is the lowest lane j < n with val >= r * icdf[j] (the serial search's
// stopping point; the table's entry n - 1 is 0, so one always exists).
__device__ __forceinline__ int ec_icdf_ballot(RangeDec *d, int32_t tabv, int off, int n, int ftb)
{
    const uint32_t s = uniform(d->rng), v = uniform(d->val);
    const uint32_t r = s >> ftb;
    const int j = lane_id();
    const int b = off + j;
    const uint32_t w = (uint32_t)__builtin_amdgcn_ds_bpermute((b >> 2) << 2, tabv);
    const uint32_t tj = r * ((w >> ((b & 3) << 3)) & 0xFFu);
    const uint64_t m = lanes_u<kCmpUge>(v, tj) & lanes_s<kCmpSlt>(j, n);
    const int ret = __builtin_ctzll(m);
    const uint32_t sv = (uint32_t)lane_get((int32_t)tj, ret);
    const uint32_t tv = ret ? (uint32_t)lane_get((int32_t)tj, ret - 1) : s;
    ec_normalize_rv(d, tv - sv, v - sv);
    return ret;
}
// silk_shell_decoder (shell_coder.c:119-151) over lane-held tables: split of
// p pulses uses the p + 1 table bytes from offset p(p+1)/2 - 1, decoded in one
// ballot step (against a serial search: 16.4 -> 16.0 ms on a 619 s SILK-WB
// stream; the pulse counts' ballot took it from 18.0 to 16.8 ms).
__device__ __forceinline__ void shell_split_lt(int16_t *c1, int16_t *c2, RangeDec *d, int p, int32_t tabv)
{
    if (p > 0) {
        c1[0] = (int16_t)ec_icdf_ballot(d, tabv, (p * (p + 1) >> 1) - 1, p + 1, 8);
        c2[0] = (int16_t)(p - c1[0]);
    } else {
        c1[0] = 0;
        c2[0] = 0;
    }
}
__device__ __forceinline__ void silk_shell_decoder_lt(int16_t *p0, RangeDec *d, int pulses4, int32_t t0, int32_t t1,
                                                      int32_t t2, int32_t t3)
{
    int16_t p3[2], p2[4], p1[8];
    shell_split_lt(&p3[0], &p3[1], d, pulses4, t3);
    shell_split_lt(&p2[0], &p2[1], d, p3[0], t2);
    shell_split_lt(&p1[0], &p1[1], d, p2[0], t1);
    shell_split_lt(&p0[0], &p0[1], d, p1[0], t0);
    shell_split_lt(&p0[2], &p0[3], d, p1[1], t0);
    shell_split_lt(&p1[2], &p1[3], d, p2[1], t1);
    shell_split_lt(&p0[4], &p0[5], d, p1[2], t0);
    shell_split_lt(&p0[6], &p0[7], d, p1[3], t0);
    shell_split_lt(&p2[2], &p2[3], d, p3[1], t2);
    shell_split_lt(&p1[4], &p1[5], d, p2[2], t1);
    shell_split_lt(&p0[8], &p0[9], d, p1[4], t0);
    shell_split_lt(&p0[10], &p0[11], d, p1[5], t0);
    shell_split_lt(&p1[6], &p1[7], d, p2[3], t1);
    shell_split_lt(&p0[12], &p0[13], d, p1[6], t0);
    shell_split_lt(&p0[14], &p0[15], d, p1[7], t0);
}

// silk_decode_pulses on the wavefront (decode_pulses.c:37-115, code_signs.c:
// 75-115).  The range decoder runs on a register copy of its state; per-block
// pulse counts live one per lane; a block's 16 magnitudes are read and written
// with one access per lane, and only the nonzero ones (a ballot) cost a sign
// symbol, in ascending position order as in the serial loop.
__device__ inline void silk_decode_pulses_lanes(RangeDec *dp, int16_t *pulses, int signalType, int quantOffsetType,
                                                int frame_length)
{
    CUO_PROF_SCOPE(kProfSilkPulses);
    dp = lds_uniform(dp);
    pulses = lds_uniform(pulses);
    RangeDec dl = ec_regs(dp);
    RangeDec *d = &dl;
    const int lane = lane_id();
    // the pulse-count and shell tables, one VGPR each
    const int32_t tpb = lane_table(tab::kPulsesPerBlockIcdf, (int)sizeof(tab::kPulsesPerBlockIcdf));
    const int32_t ts0 = lane_table(tab::kShellTable0, (int)sizeof(tab::kShellTable0));
    const int32_t ts1 = lane_table(tab::kShellTable1, (int)sizeof(tab::kShellTable1));
    const int32_t ts2 = lane_table(tab::kShellTable2, (int)sizeof(tab::kShellTable2));
    const int32_t ts3 = lane_table(tab::kShellTable3, (int)sizeof(tab::kShellTable3));
    const int RateLevelIndex = ec_icdf(d, tab::kRateLevelsIcdf + 9 * (signalType >> 1), 8);
    const int iter = (frame_length + 15) >> 4;
    const int cdf = 18 * RateLevelIndex;
    int32_t sumv = 0, lsv = 0;
    int any_ls = 0;
    {
    CUO_PROF_SCOPE(kProfSilkPulseCounts);
    for (int i = 0; i < iter; i++) {
        int nls = 0;
        int sp = ec_icdf_ballot(d, tpb, cdf, 18, 8);
        while (sp == 16 + 1) {
            nls++;
            sp = ec_icdf_lt(d, tpb, 18 * 9 + (nls == 10), 8);
        }
        sumv = lane_set(sumv, i, sp);
        lsv = lane_set(lsv, i, nls);
        any_ls |= nls;
    }
    }
    {
    CUO_PROF_SCOPE(kProfSilkShell);
    for (int i = 0; i < iter; i++) {
        const int sp = lane_get(sumv, i);
        if (sp > 0) silk_shell_decoder_lt(&pulses[i * 16], d, sp, ts0, ts1, ts2, ts3);
        else if (lane < 16) pulses[i * 16 + lane] = 0;
    }
    }
    if (any_ls) {
        for (int i = 0; i < iter; i++) {
            const int nLS = lane_get(lsv, i);
            if (nLS > 0) {
                int16_t *pp = &pulses[i * 16];
                int32_t qv = lane < 16 ? (int32_t)pp[lane] : 0;
                for (int k = 0; k < 16; k++) {
                    int32_t abs_q = lane_get(qv, k);
                    for (int j = 0; j < nLS; j++) {
                        abs_q = silk_lshift(abs_q, 1);
                        abs_q += ec_icdf(d, tab::kLsbIcdf, 8);
                    }
                    qv = lane_set(qv, k, abs_q);
                }
                if (lane < 16) pp[lane] = (int16_t)qv;
                sumv = lane == i ? (sumv | (nLS << 5)) : sumv;
            }
        }
    }
    const uint8_t *icdf_ptr = &tab::kSignIcdf[silk_smulbb(7, quantOffsetType + (signalType << 1))];
    const int length = (frame_length + 16 / 2) >> 4;
    CUO_PROF_SCOPE(kProfSilkSigns);
    for (int i = 0; i < length; i++) {
        const int p = lane_get(sumv, i);
        if (p > 0) {
            const uint32_t icdf0 = ld_u8(icdf_ptr, imin(p & 0x1F, 6));
            int16_t *q = &pulses[i * 16];
            const int32_t qv = lane < 16 ? (int32_t)q[lane] : 0;
            uint64_t nz = lanes_s<kCmpSgt>(qv, 0);
            uint32_t neg = 0;
            while (nz) {
                const int j = __builtin_ctzll(nz);
                nz &= nz - 1;
                if (!ec_icdf2(d, icdf0, 8)) neg |= 1u << j;
            }
            if (lane < 16 && ((neg >> lane) & 1u)) q[lane] = (int16_t)-qv;
        }
    }
    *dp = dl;
}
#endif

CUO_FN void silk_decode_pulses(RangeDec *d, int16_t *pulses, int signalType, int quantOffsetType, int frame_length)
{
#if CUO_LANE_PATHS
    silk_decode_pulses_lanes(d, pulses, signalType, quantOffsetType, frame_length);
    return;
#endif
    CUO_PROF_SCOPE(kProfSilkPulses);
    int32_t sum_pulses[20], nLshifts[20];
    int RateLevelIndex = ec_icdf(d, tab::kRateLevelsIcdf + 9 * (signalType >> 1), 8);
    int iter = frame_length >> 4;
    if (iter * 16 < frame_length) iter++;
    const uint8_t *cdf = tab::kPulsesPerBlockIcdf + 18 * RateLevelIndex;
    for (int i = 0; i < iter; i++) {
        nLshifts[i] = 0;
        sum_pulses[i] = ec_icdf(d, cdf, 8);
        while (sum_pulses[i] == 16 + 1) {
            nLshifts[i]++;
            sum_pulses[i] = ec_icdf(d, tab::kPulsesPerBlockIcdf + 18 * 9 + (nLshifts[i] == 10), 8);
        }
    }
    for (int i = 0; i < iter; i++) {
        if (sum_pulses[i] > 0) silk_shell_decoder(&pulses[i * 16], d, sum_pulses[i]);
        else for (int k = 0; k < 16; k++) pulses[i * 16 + k] = 0;
    }
    for (int i = 0; i < iter; i++) {
        if (nLshifts[i] > 0) {
            int nLS = nLshifts[i];
            int16_t *pp = &pulses[i * 16];
            for (int k = 0; k < 16; k++) {
                int abs_q = pp[k];
                for (int j = 0; j < nLS; j++) {
                    abs_q = silk_lshift(abs_q, 1);
                    abs_q += ec_icdf(d, tab::kLsbIcdf, 8);
                }
                pp[k] = (int16_t)abs_q;
            }
            sum_pulses[i] |= nLS << 5;
        }
    }
    // signs (code_signs.c:75-115)
    // each sign is one symbol of the two-entry icdf {icdf0, 0}
    int16_t *q = pulses;
    const uint8_t *icdf_ptr = &tab::kSignIcdf[silk_smulbb(7, quantOffsetType + (signalType << 1))];
    int length = (frame_length + 16 / 2) >> 4;
    for (int i = 0; i < length; i++) {
        int p = sum_pulses[i];
        if (p > 0) {
            uint32_t icdf0 = ld_u8(icdf_ptr, imin(p & 0x1F, 6));
            for (int j = 0; j < 16; j++)
                if (q[j] > 0) q[j] = (int16_t)(q[j] * (silk_lshift(ec_icdf2(d, icdf0, 8), 1) - 1));
        }
        q += 16;
    }
}

#if CUO_LANE_PATHS
// ---- parameter decoding on the wavefront.  Coefficient vectors (NLSFs,
// polynomial and predictor coefficients) live one element per lane; the
// serial recurrences of NLSF2A.c, LPC_fit.c, LPC_inv_pred_gain.c and
// NLSF_stabilize.c become one lane-parallel step per stage.  Each stage
// computes every element from the previous stage's values exactly as the
// serial loop does, so the results are identical.

// silk_bwexpander_32 (bwexpander_32.c:38-50): the chirp powers are a scalar
// recurrence; element i is scaled by power i.
__device__ inline int32_t bwexpander32_lanes(int32_t a, int d, int32_t chirp_Q16)
{
    const int32_t cm1 = chirp_Q16 - 65536;
    int32_t cv = 0;
    for (int i = 0; i < d; i++) {
        cv = lane_set(cv, i, chirp_Q16);
        chirp_Q16 += silk_rshift_round(silk_mul(chirp_Q16, cm1), 16);
    }
    return lane_id() < d ? silk_smulww(cv, a) : a;
}

// LPC_inv_pred_gain.c:45-121 on Q12 coefficients a (lane k < d); 0 if unstable
__device__ inline int32_t inv_pred_gain_lanes(int32_t aQ12, int d)
{
    const int lane = lane_id();
    const int32_t A_LIMIT = silk_fix(0.99975, 24);
    const int32_t MIN_INV = silk_fix(1.0f / 1e4f, 30);
    if (wave_sum(lane < d ? aQ12 : 0) >= 4096) return 0;
    int32_t A = lane < d ? silk_lshift(aQ12, 24 - 12) : 0;
    int32_t invGain = 1 << 30;
    for (int k = d - 1; k > 0; k--) {
        const int32_t Ak = lane_get(A, k);
        if (Ak > A_LIMIT || Ak < -A_LIMIT) return 0;
        const int32_t rc_Q31 = -silk_lshift(Ak, 31 - 24);
        const int32_t rc_mult1_Q30 = (1 << 30) - silk_smmul(rc_Q31, rc_Q31);
        invGain = silk_lshift(silk_smmul(invGain, rc_mult1_Q30), 2);
        if (invGain < MIN_INV) return 0;
        const int mult2Q = 32 - silk_clz32(rc_mult1_Q30 < 0 ? -rc_mult1_Q30 : rc_mult1_Q30);
        const int32_t rc_mult2 = silk_inverse32_varq(rc_mult1_Q30, mult2Q + 30);
        const int32_t mirror = __shfl(A, k - 1 - lane, 64);
        int32_t nv = A;
        int bad = 0;
        if (lane < k) {
            const int64_t t64 = silk_rshift_round64(
                silk_smull(silk_sub_sat32(A, (int32_t)silk_rshift_round64(silk_smull(mirror, rc_Q31), 31)), rc_mult2),
                mult2Q);
            bad = t64 > INT32_MAX || t64 < INT32_MIN;
            nv = (int32_t)t64;
        }
        if (wave_ballot(bad)) return 0;
        A = nv;
    }
    const int32_t A0 = lane_get(A, 0);
    if (A0 > A_LIMIT || A0 < -A_LIMIT) return 0;
    const int32_t rc_Q31 = -silk_lshift(A0, 31 - 24);
    const int32_t rc_mult1_Q30 = (1 << 30) - silk_smmul(rc_Q31, rc_Q31);
    invGain = silk_lshift(silk_smmul(invGain, rc_mult1_Q30), 2);
    if (invGain < MIN_INV) return 0;
    return invGain;
}

// NLSF_stabilize.c:47-142 on NLSFs v (lane i < L); fb is LDS scratch for the
// rare fallback after 20 rounds, which runs the serial code.
__device__ inline int32_t nlsf_stabilize_lanes(int32_t v, const int16_t *dmin, int L, int16_t *fb)
{
    const int lane = lane_id();
    const int32_t dm = lane <= L ? (int32_t)dmin[lane] : 0;
    const int32_t dmin0 = ld_s16(dmin, 0), dminL = ld_s16(dmin, L);
    for (int loops = 0; loops < 20; loops++) {
        const int32_t pv = __shfl_up(v, 1, 64);
        const int32_t base = lane < L ? v : (1 << 15);
        const int32_t diff = lane <= L ? base - ((lane == 0 ? 0 : pv) + dm) : INT32_MAX;
        const int32_t min_diff = wave_min(diff);
        if (min_diff >= 0) return v;
        const int I = __builtin_ctzll(lanes_s<kCmpEq>(diff, min_diff));
        if (I == 0) {
            v = lane == 0 ? dmin0 : v;
        } else if (I == L) {
            v = lane == L - 1 ? (int32_t)(int16_t)((1 << 15) - dminL) : v;
        } else {
            const int32_t dI = lane_get(dm, I);
            const int32_t min_center = wave_sum(lane < I ? dm : 0) + (dI >> 1);
            const int32_t max_center = (1 << 15) - wave_sum(lane > I ? dm : 0) - (dI >> 1);
            const int32_t center = (int16_t)silk_limit(silk_rshift_round(lane_get(v, I - 1) + lane_get(v, I), 1),
                                                       min_center, max_center);
            const int32_t n1 = (int16_t)(center - (dI >> 1));
            const int32_t n2 = (int16_t)(n1 + dI);
            v = lane == I - 1 ? n1 : (lane == I ? n2 : v);
        }
    }
    if (lane < L) fb[lane] = (int16_t)v;
    for (int i = 1; i < L; i++) {  // insertion sort, increasing
        int16_t x = fb[i];
        int j;
        for (j = i - 1; j >= 0 && x < fb[j]; j--) fb[j + 1] = fb[j];
        fb[j + 1] = x;
    }
    fb[0] = (int16_t)imax(fb[0], dmin0);
    for (int i = 1; i < L; i++) fb[i] = (int16_t)imax(fb[i], silk_sat16((int32_t)fb[i - 1] + ld_s16(dmin, i)));
    fb[L - 1] = (int16_t)imin(fb[L - 1], (1 << 15) - dminL);
    for (int i = L - 2; i >= 0; i--) fb[i] = (int16_t)imin(fb[i], fb[i + 1] - ld_s16(dmin, i + 1));
    return lane < L ? (int32_t)fb[lane] : 0;
}

// NLSF_decode.c:63-93: NLSFs one per lane
__device__ inline int32_t nlsf_decode_lanes(const int8_t *NLSFIndices, const NlsfCb &cb, int16_t *fb)
{
    CUO_PROF_SCOPE(kProfSilkNlsf);
    const int lane = lane_id();
    const int order = cb.order;
    const int cb1 = uniform((int)NLSFIndices[0]);
    const uint8_t *sel = &cb.ec_sel[cb1 * order / 2];
    int32_t out_Q10 = 0, res = 0;
    for (int i = order - 1; i >= 0; i--) {
        const int entry = (int)ld_u8(sel, i >> 1);
        const int pred = (int)ld_u8(cb.pred_Q8, i + ((entry >> ((i & 1) ? 4 : 0)) & 1) * (order - 1));
        const int32_t pred_Q10 = silk_smulbb(out_Q10, (int16_t)pred) >> 8;
        out_Q10 = silk_lshift(uniform((int)NLSFIndices[1 + i]), 10);
        if (out_Q10 > 0) out_Q10 = out_Q10 - silk_fix(0.1, 10);
        else if (out_Q10 < 0) out_Q10 = out_Q10 + silk_fix(0.1, 10);
        out_Q10 = silk_smlawb(pred_Q10, out_Q10, cb.quantStep_Q16);
        res = lane_set(res, i, (int16_t)out_Q10);
    }
    int32_t v = 0;
    if (lane < order) {
        const int32_t w = cb.CB1_Wght_Q9[cb1 * order + lane];
        const int32_t c = cb.CB1_Q8[cb1 * order + lane];
        v = silk_limit(silk_lshift(res, 14) / w + silk_lshift((int16_t)c, 7), 0, 32767);
    }
    return nlsf_stabilize_lanes(v, cb.deltaMin_Q15, order, fb);
}

// NLSF2A.c:66-141 for order D; a_Q12 written to LDS
template <int D>
__device__ inline void nlsf2a_lanes(int16_t *a_Q12, int32_t nlsf)
{
    CUO_PROF_SCOPE(kProfSilkNlsf2a);
    constexpr int dd = D / 2;
    const int lane = lane_id();
    // cos(NLSF) by table interpolation (NLSF2A.c:89-100), lane k
    int32_t cosv = 0;
    if (lane < D) {
        const int32_t f_int = nlsf >> (15 - 7);
        const int32_t f_frac = nlsf - silk_lshift(f_int, 15 - 7);
        const int32_t cos_val = tab::kLsfCos_Q12[f_int];
        const int32_t delta = tab::kLsfCos_Q12[f_int + 1] - cos_val;
        cosv = silk_rshift_round(silk_lshift(cos_val, 8) + silk_mul(delta, f_frac), 20 - 16);
    }
    // cos_LSF_QA[ordering[k]] = cos of NLSF k, so position j holds lane inv[j]
    const uint8_t *inv = D == 16 ? kNlsfInv16 : kNlsfInv10;
    // P (row 0, lanes 0..dd) from positions 0,2,4..; Q (row 1, lanes 16..16+dd) from 1,3,5..
    const int row = lane >> 4, n = lane & 15;
    int32_t poly = n == 0 ? (1 << 16) : (n == 1 ? -(row == 0 ? lane_get(cosv, inv[0]) : lane_get(cosv, inv[1])) : 0);
#pragma unroll
    for (int k = 1; k < dd; k++) {
        const int32_t fP = lane_get(cosv, inv[2 * k]), fQ = lane_get(cosv, inv[2 * k + 1]);
        const int32_t ftmp = row == 0 ? fP : fQ;
        const int32_t p1 = __builtin_amdgcn_update_dpp(0, poly, 0x111, 0xF, 0xF, false);  // out[n-1]
        const int32_t p2 = __builtin_amdgcn_update_dpp(0, poly, 0x112, 0xF, 0xF, false);  // out[n-2]
        const int32_t t = (int32_t)silk_rshift_round64(silk_smull(ftmp, p1), 16);
        if (n >= 1 && n <= k) poly = poly + p2 - t;
        else if (n == k + 1) poly = silk_lshift(p2, 1) - t;
    }
    // a32_QA1 (NLSF2A.c:114-120): lane i < dd takes -Qtmp - Ptmp of i, lane i >= dd
    // takes Qtmp - Ptmp of D - 1 - i
    const int m = lane < dd ? lane : D - 1 - lane;
    const int mi = m < 0 ? 0 : m;
    const int32_t P0 = __shfl(poly, mi, 64), P1 = __shfl(poly, mi + 1, 64);
    const int32_t Q0 = __shfl(poly, 16 + mi, 64), Q1 = __shfl(poly, 16 + mi + 1, 64);
    const int32_t Ptmp = P1 + P0, Qtmp = Q1 - Q0;
    int32_t a = lane < dd ? -Qtmp - Ptmp : (lane < D ? Qtmp - Ptmp : 0);
    // silk_LPC_fit (LPC_fit.c:36-81), QOUT 12, QIN 17
    int it;
    for (it = 0; it < 10; it++) {
        const int32_t absv = lane < D ? (a > 0 ? a : -a) : 0;
        int32_t maxabs = wave_max(absv);
        const int idx = __builtin_ctzll(lanes_s<kCmpSlt>(lane, D) & lanes_s<kCmpEq>(absv, maxabs));
        maxabs = silk_rshift_round(maxabs, 17 - 12);
        if (maxabs > 32767) {
            maxabs = min32(maxabs, 163838);
            const int32_t chirp_Q16 = silk_fix(0.999, 16) - silk_lshift(maxabs - 32767, 14) / ((maxabs * (idx + 1)) >> 2);
            a = bwexpander32_lanes(a, D, chirp_Q16);
        } else {
            break;
        }
    }
    int32_t q12;
    if (it == 10) {
        q12 = silk_sat16(silk_rshift_round(a, 17 - 12));
        a = silk_lshift(q12, 17 - 12);
    } else {
        q12 = (int16_t)silk_rshift_round(a, 17 - 12);
    }
    for (int i = 0; inv_pred_gain_lanes(q12, D) == 0 && i < 16; i++) {
        a = bwexpander32_lanes(a, D, 65536 - silk_lshift(2, i));
        q12 = (int16_t)silk_rshift_round(a, 17 - 12);
    }
    if (lane < D) a_Q12[lane] = (int16_t)q12;
}

// silk_decode_parameters (decode_parameters.c:35-115) on the wavefront
__device__ inline void silk_decode_parameters_lanes(SilkChannel *ch, SilkCtrl *ctrl, int condCoding, int16_t *fb)
{
    CUO_PROF_SCOPE(kProfSilkParams);
    const int lane = lane_id();
    silk_gains_dequant(ctrl->Gains_Q16, ch->indices.GainsIndices, &ch->LastGainIndex, condCoding == CODE_COND,
                       ch->nb_subfr);
    const NlsfCb cb = nlsf_cb(uniform(ch->fs_kHz) == 16);
    const int order = uniform(ch->LPC_order);
    const int32_t nlsf = nlsf_decode_lanes(ch->indices.NLSFIndices, cb, fb);
    if (order == 16) nlsf2a_lanes<16>(ctrl->PredCoef_Q12[1], nlsf);
    else nlsf2a_lanes<10>(ctrl->PredCoef_Q12[1], nlsf);
    if (uniform(ch->first_frame_after_reset) == 1) ch->indices.NLSFInterpCoef_Q2 = 4;
    const int icoef = uniform((int)ch->indices.NLSFInterpCoef_Q2);
    if (icoef < 4) {
        const int32_t prev = lane < order ? (int32_t)ch->prevNLSF_Q15[lane] : 0;
        const int32_t nlsf0 = (int16_t)(prev + (silk_mul(icoef, nlsf - prev) >> 2));
        if (order == 16) nlsf2a_lanes<16>(ctrl->PredCoef_Q12[0], nlsf0);
        else nlsf2a_lanes<10>(ctrl->PredCoef_Q12[0], nlsf0);
    } else {
        CUO_PFOR(i, 0, order) ctrl->PredCoef_Q12[0][i] = ctrl->PredCoef_Q12[1][i];
    }
    if (lane < order) ch->prevNLSF_Q15[lane] = (int16_t)nlsf;
    if (uniform(ch->lossCnt)) {
        silk_bwexpander(ctrl->PredCoef_Q12[0], order, 63570);
        silk_bwexpander(ctrl->PredCoef_Q12[1], order, 63570);
    }
    const int nb_subfr = uniform(ch->nb_subfr);
    if (uniform((int)ch->indices.signalType) == SIG_VOICED) {
        silk_decode_pitch(ch->indices.lagIndex, ch->indices.contourIndex, ctrl->pitchL, ch->fs_kHz, nb_subfr);
        const int8_t *cbk = ltp_vq(uniform((int)ch->indices.PERIndex));
        CUO_PFOR(j, 0, nb_subfr * kLtpOrder) {
            const int k = j / kLtpOrder, i = j - k * kLtpOrder;
            ctrl->LTPCoef_Q14[j] = (int16_t)silk_lshift(cbk[ch->indices.LTPIndex[k] * kLtpOrder + i], 7);
        }
        ctrl->LTP_scale_Q14 = tab::kLtpScales_Q14[ch->indices.LTP_scaleIndex];
    } else {
        CUO_PFOR(k, 0, nb_subfr) ctrl->pitchL[k] = 0;
        CUO_PFOR(k, 0, kLtpOrder * nb_subfr) ctrl->LTPCoef_Q14[k] = 0;
        ch->indices.PERIndex = 0;
        ctrl->LTP_scale_Q14 = 0;
    }
}
#endif

CUO_FN void silk_decode_parameters(SilkChannel *ch, SilkCtrl *ctrl, int condCoding)
{
    CUO_PROF_SCOPE(kProfSilkParams);
    int16_t pNLSF_Q15[kSilkMaxOrder], pNLSF0_Q15[kSilkMaxOrder];
    silk_gains_dequant(ctrl->Gains_Q16, ch->indices.GainsIndices, &ch->LastGainIndex, condCoding == CODE_COND,
                       ch->nb_subfr);
    NlsfCb cb = nlsf_cb(ch->fs_kHz == 16);
    silk_nlsf_decode(pNLSF_Q15, ch->indices.NLSFIndices, cb);
    silk_nlsf2a(ctrl->PredCoef_Q12[1], pNLSF_Q15, ch->LPC_order);
    if (ch->first_frame_after_reset == 1) ch->indices.NLSFInterpCoef_Q2 = 4;
    if (ch->indices.NLSFInterpCoef_Q2 < 4) {
        for (int i = 0; i < ch->LPC_order; i++)
            pNLSF0_Q15[i] = (int16_t)(ch->prevNLSF_Q15[i] +
                (silk_mul(ch->indices.NLSFInterpCoef_Q2, pNLSF_Q15[i] - ch->prevNLSF_Q15[i]) >> 2));
        silk_nlsf2a(ctrl->PredCoef_Q12[0], pNLSF0_Q15, ch->LPC_order);
    } else {
        for (int i = 0; i < ch->LPC_order; i++) ctrl->PredCoef_Q12[0][i] = ctrl->PredCoef_Q12[1][i];
    }
    for (int i = 0; i < ch->LPC_order; i++) ch->prevNLSF_Q15[i] = pNLSF_Q15[i];
    if (ch->lossCnt) {
        silk_bwexpander(ctrl->PredCoef_Q12[0], ch->LPC_order, 63570);
        silk_bwexpander(ctrl->PredCoef_Q12[1], ch->LPC_order, 63570);
    }
    if (ch->indices.signalType == SIG_VOICED) {
        silk_decode_pitch(ch->indices.lagIndex, ch->indices.contourIndex, ctrl->pitchL, ch->fs_kHz, ch->nb_subfr);
        const int8_t *cbk = ltp_vq(ch->indices.PERIndex);
        for (int k = 0; k < ch->nb_subfr; k++) {
            int Ix = ch->indices.LTPIndex[k];
            for (int i = 0; i < kLtpOrder; i++)
                ctrl->LTPCoef_Q14[k * kLtpOrder + i] = (int16_t)silk_lshift(cbk[Ix * kLtpOrder + i], 7);
        }
        ctrl->LTP_scale_Q14 = tab::kLtpScales_Q14[ch->indices.LTP_scaleIndex];
    } else {
        for (int k = 0; k < ch->nb_subfr; k++) ctrl->pitchL[k] = 0;
        for (int k = 0; k < kLtpOrder * ch->nb_subfr; k++) ctrl->LTPCoef_Q14[k] = 0;
        ch->indices.PERIndex = 0;
        ctrl->LTP_scale_Q14 = 0;
    }
}

// ------------------------------------------------------------------ synthesis
CUO_FN int32_t lpc_pred_q10(const int32_t *s, const int16_t *A, int order)
{
    // s points at the current sample position; taps s[-1] .. s[-order]
    int32_t p = order >> 1;
    for (int j = 0; j < order; j++) p = silk_smlawb(p, s[-1 - j], A[j]);
    return p;
}

#if CUO_LANE_PATHS
// Excitation (decode_core.c:73-88) across the lanes.  The seed recurrence
// s' = rand(s) + pulse = a*s + c + pulse (mod 2^32) is an affine map per
// sample: each lane composes the maps of its block of consecutive samples, an
// exclusive scan of the compositions over the lanes gives each block's
// starting seed, and each lane then replays its block.  Same seeds, same
// excitation as the serial loop.
__device__ inline void silk_excitation_lanes(int32_t *exc, const int16_t *pulses, int L, int32_t seed,
                                             int32_t offset_Q10)
{
    const uint32_t a = 196314165u, c = 907633515u;
    const int lane = lane_id();
    const int B = (L + kLanes - 1) / kLanes;
    const int i0 = lane * B, i1 = min32(i0 + B, L);
    uint32_t A = 1, C = 0;
    for (int i = i0; i < i1; i++) {
        A = a * A;
        C = a * C + c + (uint32_t)(int32_t)pulses[i];
    }
    for (int o = 1; o < kLanes; o <<= 1) {
        uint32_t Ap = (uint32_t)__shfl_up((int)A, o, 64), Cp = (uint32_t)__shfl_up((int)C, o, 64);
        if (lane >= o) {
            C = A * Cp + C;
            A = A * Ap;
        }
    }
    uint32_t Ae = (uint32_t)__shfl_up((int)A, 1, 64), Ce = (uint32_t)__shfl_up((int)C, 1, 64);
    if (lane == 0) {
        Ae = 1;
        Ce = 0;
    }
    uint32_t sd = Ae * (uint32_t)seed + Ce;
    for (int i = i0; i < i1; i++) {
        sd = a * sd + c;
        int32_t e = silk_lshift((int32_t)pulses[i], 14);
        if (e > 0) e -= 80 << 4;
        else if (e < 0) e += 80 << 4;
        e += offset_Q10 << 4;
        if ((int32_t)sd < 0) e = -e;
        exc[i] = e;
        sd += (uint32_t)(int32_t)pulses[i];
    }
}

// silk_decode_core (decode_core.c:38-237) on the wavefront.  The LPC
// synthesis recursion stays serial per sample, but its order-16 prediction is
// one multiply per lane (lane j holds output n-1-j and coefficient j) and a
// 16-lane DPP sum; the synthesis history lives in that register across
// subframes.  The LTP synthesis reads its own output only lag-2 samples back,
// so blocks of lag-2 outputs are computed one per lane.  The rewhitening
// filter and state rescaling are lane-parallel loops.  Integer sums wrap like
// the serial code's, so the order of summation does not change any value.
__device__ inline void silk_decode_core_lanes(SilkChannel *ch, SilkCtrl *ctrl, int16_t *xq, const int16_t *pulses,
                                              SilkScratch *sc)
{
    CUO_PROF_SCOPE(kProfSilkCore);
    int16_t *sLTP = sc->sLTP;
    int32_t *sLTP_Q15 = sc->sLTP_Q15;
    int32_t *res_Q14 = sc->res_Q14;
    const int lane = lane_id();
    const int frame_length = uniform(ch->frame_length);
    const int subfr = uniform(ch->subfr_length);
    const int nb_subfr = uniform(ch->nb_subfr);
    const int order = uniform(ch->LPC_order);
    const int ltp_mem = uniform(ch->ltp_mem_length);
    const int sig0 = uniform((int)ch->indices.signalType);
    const int32_t offset_Q10 = tab::kQuantOffsets_Q10[(sig0 >> 1) * 2 + uniform((int)ch->indices.quantOffsetType)];
    const int NLSF_interpolation_flag = uniform((int)ch->indices.NLSFInterpCoef_Q2) < 1 << 2;
    silk_excitation_lanes(ch->exc_Q14, pulses, frame_length, uniform((int32_t)ch->indices.Seed), offset_Q10);
    int32_t win = lane < kSilkMaxOrder ? ch->sLPC_Q14_buf[kSilkMaxOrder - 1 - lane] : 0;
    const int32_t *pexc = ch->exc_Q14;
    int16_t *pxq = xq;
    int sLTP_buf_idx = ltp_mem;
    int lag = 0;
    for (int k = 0; k < nb_subfr; k++) {
        const int16_t *A_Q12 = ctrl->PredCoef_Q12[k >> 1];
        const int32_t a16 = lane < order ? (int32_t)((uint32_t)(uint16_t)A_Q12[lane] << 16) : 0;
        int16_t *B_Q14 = &ctrl->LTPCoef_Q14[k * kLtpOrder];
        int signalType = sig0;
        const int32_t gain = uniform(ctrl->Gains_Q16[k]);
        const int32_t prev_gain = uniform(ch->prev_gain_Q16);
        const int32_t Gain_Q10 = gain >> 6;
        int32_t inv_gain_Q31 = silk_inverse32_varq(gain, 47);
        int32_t gain_adj_Q16 = 1 << 16;
        if (gain != prev_gain) {
            gain_adj_Q16 = silk_div32_varq(prev_gain, gain, 16);
            win = silk_smulww(gain_adj_Q16, win);
        }
        ch->prev_gain_Q16 = gain;
        if (uniform(ch->lossCnt) && uniform(ch->prevSignalType) == SIG_VOICED && sig0 != SIG_VOICED && k < 2) {
            CUO_PFOR(i, 0, kLtpOrder) B_Q14[i] = i == kLtpOrder / 2 ? (int16_t)silk_fix(0.25, 14) : 0;
            signalType = SIG_VOICED;
            ctrl->pitchL[k] = ch->lagPrev;
        }
        const int32_t *pres = pexc;
        if (signalType == SIG_VOICED) {
            lag = uniform(ctrl->pitchL[k]);
            if (k == 0 || (k == 2 && NLSF_interpolation_flag)) {
                const int start_idx = ltp_mem - lag - order - kLtpOrder / 2;
                if (k == 2) CUO_PFOR(i, 0, 2 * subfr) ch->outBuf[ltp_mem + i] = xq[i];
                silk_lpc_analysis_filter(&sLTP[start_idx], &ch->outBuf[start_idx + k * subfr], A_Q12,
                                         ltp_mem - start_idx, order);
                if (k == 0) inv_gain_Q31 = silk_lshift(silk_smulwb(inv_gain_Q31, uniform(ctrl->LTP_scale_Q14)), 2);
                CUO_PFOR(i, 0, lag + kLtpOrder / 2)
                    sLTP_Q15[sLTP_buf_idx - i - 1] = silk_smulwb(inv_gain_Q31, sLTP[ltp_mem - i - 1]);
            } else if (gain_adj_Q16 != 1 << 16) {
                CUO_PFOR(i, 0, lag + kLtpOrder / 2)
                    sLTP_Q15[sLTP_buf_idx - i - 1] = silk_smulww(gain_adj_Q16, sLTP_Q15[sLTP_buf_idx - i - 1]);
            }
            const int32_t b0 = uniform((int32_t)B_Q14[0]), b1 = uniform((int32_t)B_Q14[1]),
                          b2 = uniform((int32_t)B_Q14[2]), b3 = uniform((int32_t)B_Q14[3]),
                          b4 = uniform((int32_t)B_Q14[4]);
            const int blk = max32(1, min32(lag - 2, kLanes));
            for (int i0 = 0; i0 < subfr; i0 += blk) {
                const int i = i0 + lane;
                if (lane < blk && i < subfr) {
                    const int32_t *pl = &sLTP_Q15[sLTP_buf_idx + i - lag + kLtpOrder / 2];
                    int32_t p = 2;
                    p = silk_smlawb(p, pl[0], b0);
                    p = silk_smlawb(p, pl[-1], b1);
                    p = silk_smlawb(p, pl[-2], b2);
                    p = silk_smlawb(p, pl[-3], b3);
                    p = silk_smlawb(p, pl[-4], b4);
                    const int32_t r = pexc[i] + silk_lshift(p, 1);
                    res_Q14[i] = r;
                    sLTP_Q15[sLTP_buf_idx + i] = silk_lshift(r, 1);
                }
            }
            sLTP_buf_idx += subfr;
            pres = res_Q14;
        }
        const int32_t rnd = order >> 1;
        for (int n0 = 0; n0 < subfr; n0 += kLanes) {
            const int cnt = min32(kLanes, subfr - n0);
            const int32_t rv = lane < cnt ? pres[n0 + lane] : 0;
            int32_t outv = 0;
            for (int i = 0; i < cnt; i++) {
                const int32_t t = row16_prefix_sum(__mulhi(win, a16));
                const int32_t p = (int32_t)((uint32_t)__builtin_amdgcn_readlane(t, 15) + (uint32_t)rnd);
                const int32_t y = silk_add_sat32(__builtin_amdgcn_readlane(rv, i), silk_lshift_sat32(p, 4));
                outv = lane == i ? silk_sat16(silk_rshift_round(silk_smulww(y, Gain_Q10), 8)) : outv;
                win = row16_shift_in(win, y);
            }
            if (lane < cnt) pxq[n0 + lane] = (int16_t)outv;
        }
        pexc += subfr;
        pxq += subfr;
    }
    if (lane < kSilkMaxOrder) ch->sLPC_Q14_buf[kSilkMaxOrder - 1 - lane] = win;
}
#endif

// silk_decode_core (decode_core.c:38-237)
CUO_FN void silk_decode_core(SilkChannel *ch, SilkCtrl *ctrl, int16_t *xq, const int16_t *pulses, SilkScratch *sc)
{
#if CUO_LANE_PATHS
    silk_decode_core_lanes(ch, ctrl, xq, pulses, sc);
    return;
#endif
    CUO_PROF_SCOPE(kProfSilkCore);
    int16_t *sLTP = sc->sLTP;
    int32_t *sLTP_Q15 = sc->sLTP_Q15;
    int32_t *res_Q14 = sc->res_Q14;
    int32_t *sLPC_Q14 = sc->sLPC_Q14;
    int lag = 0;
    int16_t A_Q12_tmp[kSilkMaxOrder];
    int32_t offset_Q10 = tab::kQuantOffsets_Q10[(ch->indices.signalType >> 1) * 2 + ch->indices.quantOffsetType];
    int NLSF_interpolation_flag = ch->indices.NLSFInterpCoef_Q2 < 1 << 2;
    int32_t rand_seed = ch->indices.Seed;
    for (int i = 0; i < ch->frame_length; i++) {
        rand_seed = silk_rand(rand_seed);
        int32_t e = silk_lshift((int32_t)pulses[i], 14);
        if (e > 0) e -= 80 << 4;
        else if (e < 0) e += 80 << 4;
        e += offset_Q10 << 4;
        if (rand_seed < 0) e = -e;
        ch->exc_Q14[i] = e;
        rand_seed = add32_ovf(rand_seed, pulses[i]);
    }
    for (int i = 0; i < kSilkMaxOrder; i++) sLPC_Q14[i] = ch->sLPC_Q14_buf[i];
    const int32_t *pexc = ch->exc_Q14;
    int16_t *pxq = xq;
    int sLTP_buf_idx = ch->ltp_mem_length;
    for (int k = 0; k < ch->nb_subfr; k++) {
        int32_t *pres = res_Q14;
        const int16_t *A_Q12 = ctrl->PredCoef_Q12[k >> 1];
        for (int i = 0; i < ch->LPC_order; i++) A_Q12_tmp[i] = A_Q12[i];
        int16_t *B_Q14 = &ctrl->LTPCoef_Q14[k * kLtpOrder];
        int signalType = ch->indices.signalType;
        int32_t Gain_Q10 = ctrl->Gains_Q16[k] >> 6;
        int32_t inv_gain_Q31 = silk_inverse32_varq(ctrl->Gains_Q16[k], 47);
        int32_t gain_adj_Q16;
        if (ctrl->Gains_Q16[k] != ch->prev_gain_Q16) {
            gain_adj_Q16 = silk_div32_varq(ch->prev_gain_Q16, ctrl->Gains_Q16[k], 16);
            for (int i = 0; i < kSilkMaxOrder; i++) sLPC_Q14[i] = silk_smulww(gain_adj_Q16, sLPC_Q14[i]);
        } else {
            gain_adj_Q16 = 1 << 16;
        }
        ch->prev_gain_Q16 = ctrl->Gains_Q16[k];
        if (ch->lossCnt && ch->prevSignalType == SIG_VOICED && ch->indices.signalType != SIG_VOICED && k < 2) {
            for (int i = 0; i < kLtpOrder; i++) B_Q14[i] = 0;
            B_Q14[kLtpOrder / 2] = (int16_t)silk_fix(0.25, 14);
            signalType = SIG_VOICED;
            ctrl->pitchL[k] = ch->lagPrev;
        }
        if (signalType == SIG_VOICED) {
            lag = ctrl->pitchL[k];
            if (k == 0 || (k == 2 && NLSF_interpolation_flag)) {
                int start_idx = ch->ltp_mem_length - lag - ch->LPC_order - kLtpOrder / 2;
                if (k == 2)
                    for (int i = 0; i < 2 * ch->subfr_length; i++) ch->outBuf[ch->ltp_mem_length + i] = xq[i];
                silk_lpc_analysis_filter(&sLTP[start_idx], &ch->outBuf[start_idx + k * ch->subfr_length], A_Q12,
                                         ch->ltp_mem_length - start_idx, ch->LPC_order);
                if (k == 0) inv_gain_Q31 = silk_lshift(silk_smulwb(inv_gain_Q31, ctrl->LTP_scale_Q14), 2);
                for (int i = 0; i < lag + kLtpOrder / 2; i++)
                    sLTP_Q15[sLTP_buf_idx - i - 1] = silk_smulwb(inv_gain_Q31, sLTP[ch->ltp_mem_length - i - 1]);
            } else if (gain_adj_Q16 != 1 << 16) {
                for (int i = 0; i < lag + kLtpOrder / 2; i++)
                    sLTP_Q15[sLTP_buf_idx - i - 1] = silk_smulww(gain_adj_Q16, sLTP_Q15[sLTP_buf_idx - i - 1]);
            }
        }
        if (signalType == SIG_VOICED) {
            const int32_t *pl = &sLTP_Q15[sLTP_buf_idx - lag + kLtpOrder / 2];
            for (int i = 0; i < ch->subfr_length; i++) {
                int32_t p = 2;
                p = silk_smlawb(p, pl[0], B_Q14[0]);
                p = silk_smlawb(p, pl[-1], B_Q14[1]);
                p = silk_smlawb(p, pl[-2], B_Q14[2]);
                p = silk_smlawb(p, pl[-3], B_Q14[3]);
                p = silk_smlawb(p, pl[-4], B_Q14[4]);
                pl++;
                pres[i] = pexc[i] + silk_lshift(p, 1);
                sLTP_Q15[sLTP_buf_idx] = silk_lshift(pres[i], 1);
                sLTP_buf_idx++;
            }
        } else {
            pres = (int32_t *)pexc;
        }
        for (int i = 0; i < ch->subfr_length; i++) {
            int32_t p = lpc_pred_q10(&sLPC_Q14[kSilkMaxOrder + i], A_Q12_tmp, ch->LPC_order);
            sLPC_Q14[kSilkMaxOrder + i] = silk_add_sat32(pres[i], silk_lshift_sat32(p, 4));
            pxq[i] = (int16_t)silk_sat16(silk_rshift_round(silk_smulww(sLPC_Q14[kSilkMaxOrder + i], Gain_Q10), 8));
        }
        for (int i = 0; i < kSilkMaxOrder; i++) sLPC_Q14[i] = sLPC_Q14[ch->subfr_length + i];
        pexc += ch->subfr_length;
        pxq += ch->subfr_length;
    }
    for (int i = 0; i < kSilkMaxOrder; i++) ch->sLPC_Q14_buf[i] = sLPC_Q14[i];
}

// ------------------------------------------------------------------ PLC (PLC.c)
static constexpr int16_t kHarmAtt_Q15[2] = {32440, 31130};
static constexpr int16_t kPlcRandAttV_Q15[2] = {31130, 26214};
static constexpr int16_t kPlcRandAttUV_Q15[2] = {32440, 29491};

CUO_FN void silk_plc_update(SilkChannel *ch, const SilkCtrl *ctrl)
{
    SilkPlcState *plc = &ch->plc;
    ch->prevSignalType = ch->indices.signalType;
    int32_t LTP_Gain_Q14 = 0;
    if (ch->indices.signalType == SIG_VOICED) {
        for (int j = 0; j * ch->subfr_length < ctrl->pitchL[ch->nb_subfr - 1]; j++) {
            if (j == ch->nb_subfr) break;
            int32_t t = 0;
            for (int i = 0; i < kLtpOrder; i++) t += ctrl->LTPCoef_Q14[(ch->nb_subfr - 1 - j) * kLtpOrder + i];
            if (t > LTP_Gain_Q14) {
                LTP_Gain_Q14 = t;
                for (int i = 0; i < kLtpOrder; i++)
                    plc->LTPCoef_Q14[i] = ctrl->LTPCoef_Q14[silk_smulbb(ch->nb_subfr - 1 - j, kLtpOrder) + i];
                plc->pitchL_Q8 = silk_lshift(ctrl->pitchL[ch->nb_subfr - 1 - j], 8);
            }
        }
        for (int i = 0; i < kLtpOrder; i++) plc->LTPCoef_Q14[i] = 0;
        plc->LTPCoef_Q14[kLtpOrder / 2] = (int16_t)LTP_Gain_Q14;
        if (LTP_Gain_Q14 < 11469) {
            int32_t scale_Q10 = silk_lshift(11469, 10) / imax(LTP_Gain_Q14, 1);
            for (int i = 0; i < kLtpOrder; i++)
                plc->LTPCoef_Q14[i] = (int16_t)(silk_smulbb(plc->LTPCoef_Q14[i], scale_Q10) >> 10);
        } else if (LTP_Gain_Q14 > 15565) {
            int32_t scale_Q14 = silk_lshift(15565, 14) / imax(LTP_Gain_Q14, 1);
            for (int i = 0; i < kLtpOrder; i++)
                plc->LTPCoef_Q14[i] = (int16_t)(silk_smulbb(plc->LTPCoef_Q14[i], scale_Q14) >> 14);
        }
    } else {
        plc->pitchL_Q8 = silk_lshift(silk_smulbb(ch->fs_kHz, 18), 8);
        for (int i = 0; i < kLtpOrder; i++) plc->LTPCoef_Q14[i] = 0;
    }
    for (int i = 0; i < ch->LPC_order; i++) plc->prevLPC_Q12[i] = ctrl->PredCoef_Q12[1][i];
    plc->prevLTP_scale_Q14 = (int16_t)ctrl->LTP_scale_Q14;
    plc->prevGain_Q16[0] = ctrl->Gains_Q16[ch->nb_subfr - 2];
    plc->prevGain_Q16[1] = ctrl->Gains_Q16[ch->nb_subfr - 1];
    plc->subfr_length = ch->subfr_length;
    plc->nb_subfr = ch->nb_subfr;
}

CUO_FN void silk_plc_conceal(SilkChannel *ch, SilkCtrl *ctrl, int16_t *frame, SilkScratch *sc)
{
    SilkPlcState *plc = &ch->plc;
    int32_t *sLTP_Q14 = sc->sLTP_Q15;
    int16_t *sLTP = sc->sLTP;
    int16_t A_Q12[kSilkMaxOrder];
    int32_t prevGain_Q10[2];
    prevGain_Q10[0] = plc->prevGain_Q16[0] >> 6;
    prevGain_Q10[1] = plc->prevGain_Q16[1] >> 6;
    if (ch->first_frame_after_reset)
        for (int i = 0; i < kSilkMaxOrder; i++) plc->prevLPC_Q12[i] = 0;
    // silk_PLC_energy
    int32_t energy1, energy2, shift1, shift2;
    {
        int16_t *exc_buf = sc->exc_buf;  // 2 * subfr_length <= 160 samples
        for (int k = 0; k < 2; k++)
            for (int i = 0; i < ch->subfr_length; i++)
                exc_buf[k * ch->subfr_length + i] = (int16_t)silk_sat16(
                    silk_smulww(ch->exc_Q14[i + (k + ch->nb_subfr - 2) * ch->subfr_length], prevGain_Q10[k]) >> 8);
        silk_sum_sqr_shift(&energy1, &shift1, exc_buf, ch->subfr_length);
        silk_sum_sqr_shift(&energy2, &shift2, &exc_buf[ch->subfr_length], ch->subfr_length);
    }
    const int32_t *rand_ptr;
    if ((energy1 >> shift2) < (energy2 >> shift1))
        rand_ptr = &ch->exc_Q14[imax(0, (plc->nb_subfr - 1) * plc->subfr_length - 128)];
    else
        rand_ptr = &ch->exc_Q14[imax(0, plc->nb_subfr * plc->subfr_length - 128)];
    int16_t *B_Q14 = plc->LTPCoef_Q14;
    int16_t rand_scale_Q14 = plc->randScale_Q14;
    int32_t harm_Gain_Q15 = kHarmAtt_Q15[imin(1, ch->lossCnt)];
    int32_t rand_Gain_Q15 = ch->prevSignalType == SIG_VOICED ? kPlcRandAttV_Q15[imin(1, ch->lossCnt)]
                                                             : kPlcRandAttUV_Q15[imin(1, ch->lossCnt)];
    silk_bwexpander(plc->prevLPC_Q12, ch->LPC_order, silk_fix(0.99, 16));
    for (int i = 0; i < ch->LPC_order; i++) A_Q12[i] = plc->prevLPC_Q12[i];
    if (ch->lossCnt == 0) {
        rand_scale_Q14 = 1 << 14;
        if (ch->prevSignalType == SIG_VOICED) {
            for (int i = 0; i < kLtpOrder; i++) rand_scale_Q14 = (int16_t)(rand_scale_Q14 - B_Q14[i]);
            rand_scale_Q14 = max16(3277, rand_scale_Q14);
            rand_scale_Q14 = (int16_t)(silk_smulbb(rand_scale_Q14, plc->prevLTP_scale_Q14) >> 14);
        } else {
            int32_t invGain_Q30 = silk_lpc_inverse_pred_gain(plc->prevLPC_Q12, ch->LPC_order);
            int32_t down = min32((1 << 30) >> 3, invGain_Q30);
            down = max32((1 << 30) >> 8, down);
            down = silk_lshift(down, 3);
            rand_Gain_Q15 = silk_smulwb(down, rand_Gain_Q15) >> 14;
        }
    }
    int32_t rand_seed = plc->rand_seed;
    int lag = silk_rshift_round(plc->pitchL_Q8, 8);
    int sLTP_buf_idx = ch->ltp_mem_length;
    int idx = ch->ltp_mem_length - lag - ch->LPC_order - kLtpOrder / 2;
    silk_lpc_analysis_filter(&sLTP[idx], &ch->outBuf[idx], A_Q12, ch->ltp_mem_length - idx, ch->LPC_order);
    int32_t inv_gain_Q30 = silk_inverse32_varq(plc->prevGain_Q16[1], 46);
    inv_gain_Q30 = min32(inv_gain_Q30, INT32_MAX >> 1);
    for (int i = idx + ch->LPC_order; i < ch->ltp_mem_length; i++) sLTP_Q14[i] = silk_smulwb(inv_gain_Q30, sLTP[i]);
    for (int k = 0; k < ch->nb_subfr; k++) {
        const int32_t *pl = &sLTP_Q14[sLTP_buf_idx - lag + kLtpOrder / 2];
        for (int i = 0; i < ch->subfr_length; i++) {
            int32_t p = 2;
            p = silk_smlawb(p, pl[0], B_Q14[0]);
            p = silk_smlawb(p, pl[-1], B_Q14[1]);
            p = silk_smlawb(p, pl[-2], B_Q14[2]);
            p = silk_smlawb(p, pl[-3], B_Q14[3]);
            p = silk_smlawb(p, pl[-4], B_Q14[4]);
            pl++;
            rand_seed = silk_rand(rand_seed);
            int ri = (rand_seed >> 25) & 127;
            sLTP_Q14[sLTP_buf_idx] = silk_lshift(silk_smlawb(p, rand_ptr[ri], rand_scale_Q14), 2);
            sLTP_buf_idx++;
        }
        for (int j = 0; j < kLtpOrder; j++) B_Q14[j] = (int16_t)(silk_smulbb(harm_Gain_Q15, B_Q14[j]) >> 15);
        rand_scale_Q14 = (int16_t)(silk_smulbb(rand_scale_Q14, rand_Gain_Q15) >> 15);
        plc->pitchL_Q8 = silk_smlawb(plc->pitchL_Q8, plc->pitchL_Q8, 655);
        plc->pitchL_Q8 = min32(plc->pitchL_Q8, silk_lshift(silk_smulbb(18, ch->fs_kHz), 8));
        lag = silk_rshift_round(plc->pitchL_Q8, 8);
    }
    int32_t *sLPC = &sLTP_Q14[ch->ltp_mem_length - kSilkMaxOrder];
    for (int i = 0; i < kSilkMaxOrder; i++) sLPC[i] = ch->sLPC_Q14_buf[i];
    for (int i = 0; i < ch->frame_length; i++) {
        int32_t p = lpc_pred_q10(&sLPC[kSilkMaxOrder + i], A_Q12, ch->LPC_order);
        sLPC[kSilkMaxOrder + i] = silk_add_sat32(sLPC[kSilkMaxOrder + i], silk_lshift_sat32(p, 4));
        frame[i] = (int16_t)silk_sat16(silk_sat16(silk_rshift_round(silk_smulww(sLPC[kSilkMaxOrder + i], prevGain_Q10[1]), 8)));
    }
    for (int i = 0; i < kSilkMaxOrder; i++) ch->sLPC_Q14_buf[i] = sLPC[ch->frame_length + i];
    plc->rand_seed = rand_seed;
    plc->randScale_Q14 = rand_scale_Q14;
    for (int i = 0; i < 4; i++) ctrl->pitchL[i] = lag;
}

CUO_FN void silk_plc_glue_frames(SilkChannel *ch, int16_t *frame, int length)
{
    SilkPlcState *plc = &ch->plc;
    if (ch->lossCnt) {
        silk_sum_sqr_shift(&plc->conc_energy, &plc->conc_energy_shift, frame, length);
        plc->last_frame_lost = 1;
    } else {
        if (plc->last_frame_lost) {
            int32_t energy, energy_shift;
            silk_sum_sqr_shift(&energy, &energy_shift, frame, length);
            if (energy_shift > plc->conc_energy_shift)
                plc->conc_energy = plc->conc_energy >> (energy_shift - plc->conc_energy_shift);
            else if (energy_shift < plc->conc_energy_shift)
                energy = energy >> (plc->conc_energy_shift - energy_shift);
            if (energy > plc->conc_energy) {
                int32_t LZ = silk_clz32(plc->conc_energy) - 1;
                plc->conc_energy = silk_lshift(plc->conc_energy, LZ);
                energy = energy >> max32(24 - LZ, 0);
                int32_t frac_Q24 = plc->conc_energy / imax(energy, 1);
                int32_t gain_Q16 = silk_lshift(silk_sqrt_approx(frac_Q24), 4);
                int32_t slope_Q16 = ((1 << 16) - gain_Q16) / length;
                slope_Q16 = silk_lshift(slope_Q16, 2);
                for (int i = 0; i < length; i++) {
                    frame[i] = (int16_t)silk_smulwb(gain_Q16, frame[i]);
                    gain_Q16 += slope_Q16;
                    if (gain_Q16 > 1 << 16) break;
                }
            }
        }
        plc->last_frame_lost = 0;
    }
}

CUO_FN void silk_plc(SilkChannel *ch, SilkCtrl *ctrl, int16_t *frame, int lost, SilkScratch *sc)
{
    CUO_PROF_SCOPE(kProfSilkPlc);
    if (ch->fs_kHz != ch->plc.fs_kHz) {
        silk_plc_reset(ch);
        ch->plc.fs_kHz = ch->fs_kHz;
    }
    if (lost) {
        silk_plc_conceal(ch, ctrl, frame, sc);
        ch->lossCnt++;
    } else {
        silk_plc_update(ch, ctrl);
    }
}

// ------------------------------------------------------------------ CNG (CNG.c)
CUO_FN void silk_cng(SilkChannel *ch, const SilkCtrl *ctrl, int16_t *frame, int length, SilkScratch *sc)
{
    CUO_PROF_SCOPE(kProfSilkCng);
    SilkCngState *cng = &ch->cng;
    if (ch->fs_kHz != cng->fs_kHz) {
        silk_cng_reset(ch);
        cng->fs_kHz = ch->fs_kHz;
    }
    if (ch->lossCnt == 0 && ch->prevSignalType == SIG_NO_VOICE) {
        for (int i = 0; i < ch->LPC_order; i++)
            cng->smth_NLSF_Q15[i] = (int16_t)(cng->smth_NLSF_Q15[i] +
                silk_smulwb((int32_t)ch->prevNLSF_Q15[i] - (int32_t)cng->smth_NLSF_Q15[i], 16348));
        int32_t max_Gain_Q16 = 0;
        int subfr = 0;
        for (int i = 0; i < ch->nb_subfr; i++)
            if (ctrl->Gains_Q16[i] > max_Gain_Q16) { max_Gain_Q16 = ctrl->Gains_Q16[i]; subfr = i; }
        // exc_buf[L + i] = exc_buf[i] for i from (nb_subfr - 1) L - 1 down to
        // 0 (a move up by one subframe), then the loudest subframe to the
        // front.  Blocks of kLanes from the top: each block is read whole
        // before it is written, and what is left below it lies under the
        // block's destination.
        int32_t *exc_buf = sc->cng_exc;
        const int L = ch->subfr_length;
        for (int top = (ch->nb_subfr - 1) * L; top > 0; top -= kLanes) {
            const int lo = imax(0, top - kLanes);
            const int i = lo + lane_id();
            const int32_t v = i < top ? exc_buf[i] : 0;
            if (i < top) exc_buf[L + i] = v;
        }
        CUO_PFOR(i, 0, L) exc_buf[i] = ch->exc_Q14[subfr * L + i];
        for (int i = 0; i < ch->nb_subfr; i++) {
            cng->smth_Gain_Q16 += silk_smulwb(ctrl->Gains_Q16[i] - cng->smth_Gain_Q16, 4634);
            if (silk_smulww(cng->smth_Gain_Q16, 46396) > ctrl->Gains_Q16[i]) cng->smth_Gain_Q16 = ctrl->Gains_Q16[i];
        }
    }
    if (ch->lossCnt) {
        int32_t *sig = sc->cng_sig;
        int16_t A_Q12[kSilkMaxOrder];
        int32_t gain_Q16 = silk_smulww(ch->plc.randScale_Q14, ch->plc.prevGain_Q16[1]);
        if (gain_Q16 >= (1 << 21) || cng->smth_Gain_Q16 > (1 << 23)) {
            gain_Q16 = (gain_Q16 >> 16) * (gain_Q16 >> 16);
            gain_Q16 = ((cng->smth_Gain_Q16 >> 16) * (cng->smth_Gain_Q16 >> 16)) - silk_lshift(gain_Q16, 5);
            gain_Q16 = silk_lshift(silk_sqrt_approx(gain_Q16), 16);
        } else {
            gain_Q16 = silk_smulww(gain_Q16, gain_Q16);
            gain_Q16 = silk_smulww(cng->smth_Gain_Q16, cng->smth_Gain_Q16) - silk_lshift(gain_Q16, 5);
            gain_Q16 = silk_lshift(silk_sqrt_approx(gain_Q16), 8);
        }
        int32_t gain_Q10 = gain_Q16 >> 6;
        // silk_CNG_exc
        int exc_mask = 255;
        while (exc_mask > length) exc_mask >>= 1;
        int32_t seed = cng->rand_seed;
        const int32_t *exc_buf = sc->cng_exc;
        for (int i0 = 0; i0 < length; i0 += kLanes) {
            // the serial seed sequence picks kLanes indices, then one gather
            const int nb = imin(kLanes, length - i0);
            int32_t idxv = 0;
            serial_lanes(nb, [&](auto k) {
                seed = silk_rand(seed);
                idxv = lane_set_s(idxv, k, (seed >> 24) & exc_mask);
            });
            if (lane_id() < nb) sig[kSilkMaxOrder + i0 + lane_id()] = exc_buf[idxv];
        }
        cng->rand_seed = seed;
        silk_nlsf2a(A_Q12, cng->smth_NLSF_Q15, ch->LPC_order);
        for (int i = 0; i < kSilkMaxOrder; i++) sig[i] = cng->synth_state[i];
        for (int i = 0; i < length; i++) {
            int32_t p = lpc_pred_q10(&sig[kSilkMaxOrder + i], A_Q12, ch->LPC_order);
            sig[kSilkMaxOrder + i] = silk_add_sat32(sig[kSilkMaxOrder + i], silk_lshift_sat32(p, 4));
            frame[i] = (int16_t)silk_sat16(frame[i] + silk_sat16(silk_rshift_round(silk_smulww(sig[kSilkMaxOrder + i], gain_Q10), 8)));
        }
        for (int i = 0; i < kSilkMaxOrder; i++) cng->synth_state[i] = sig[length + i];
    } else {
        for (int i = 0; i < ch->LPC_order; i++) cng->synth_state[i] = 0;
    }
}

// ------------------------------------------------------------------ frame + API
// silk_decode_frame (decode_frame.c:43-169)
CUO_FN void silk_decode_frame(SilkChannel *ch, RangeDec *d, int16_t *pOut, int32_t *pN, int lostFlag, int condCoding,
                              SilkScratch *sc)
{
    SilkCtrl *ctrl = &sc->ctrl;
    int L = ch->frame_length;
    ctrl->LTP_scale_Q14 = 0;
    if (lostFlag == FLAG_NORMAL || (lostFlag == FLAG_LBRR && ch->LBRR_flags[ch->nFramesDecoded] == 1)) {
        int16_t *pulses = sc->pulses;
        silk_decode_indices(ch, d, ch->nFramesDecoded, lostFlag, condCoding);
        silk_decode_pulses(d, pulses, ch->indices.signalType, ch->indices.quantOffsetType, ch->frame_length);
#if CUO_LANE_PATHS
        silk_decode_parameters_lanes(ch, ctrl, condCoding, sc->nlsf_tmp);
#else
        silk_decode_parameters(ch, ctrl, condCoding);
#endif
        silk_decode_core(ch, ctrl, pOut, pulses, sc);
        int mv_len = ch->ltp_mem_length - ch->frame_length;
        CUO_PFOR(i, 0, mv_len) ch->outBuf[i] = ch->outBuf[ch->frame_length + i];
        CUO_PFOR(i, 0, ch->frame_length) ch->outBuf[mv_len + i] = pOut[i];
        silk_plc(ch, ctrl, pOut, 0, sc);
        ch->lossCnt = 0;
        ch->prevSignalType = ch->indices.signalType;
        ch->first_frame_after_reset = 0;
    } else {
        silk_plc(ch, ctrl, pOut, 1, sc);
        int mv_len = ch->ltp_mem_length - ch->frame_length;
        CUO_PFOR(i, 0, mv_len) ch->outBuf[i] = ch->outBuf[ch->frame_length + i];
        CUO_PFOR(i, 0, ch->frame_length) ch->outBuf[mv_len + i] = pOut[i];
    }
    silk_cng(ch, ctrl, pOut, L, sc);
    silk_plc_glue_frames(ch, pOut, L);
    ch->lagPrev = ctrl->pitchL[ch->nb_subfr - 1];
    *pN = L;
}

struct SilkControl {
    int32_t internalSampleRate;
    int32_t payloadSize_ms;
    int32_t nChannelsInternal;  // coded channels of the packet (1, or 2 in the stereo build)
    int32_t nChannelsAPI;       // output channels
};

// silk_stereo_decode_pred (stereo_decode_pred.c:34-67): mid/side predictors
CUO_FN void silk_stereo_decode_pred(RangeDec *d, int32_t *pred0, int32_t *pred1)
{
    int n = ec_icdf(d, tab::kStereoPredJointIcdf, 8);
    int ix02 = n / 5, ix12 = n - 5 * ix02;
    int ix00 = ec_icdf(d, tab::kUniform3Icdf, 8);
    int ix01 = ec_icdf(d, tab::kUniform5Icdf, 8);
    int ix10 = ec_icdf(d, tab::kUniform3Icdf, 8);
    int ix11 = ec_icdf(d, tab::kUniform5Icdf, 8);
    int32_t p[2];
    const int ix[2][3] = {{ix00, ix01, ix02}, {ix10, ix11, ix12}};
    for (int k = 0; k < 2; k++) {
        const int i0 = ix[k][0] + 3 * ix[k][2];
        const int32_t low_Q13 = ld_s16(tab::kStereoPredQuant_Q13, i0);
        const int32_t step_Q13 = silk_smulwb(ld_s16(tab::kStereoPredQuant_Q13, i0 + 1) - low_Q13,
                                             silk_fix(0.5 / 5, 16));  // 0.5 / STEREO_QUANT_SUB_STEPS
        p[k] = silk_smlabb(low_Q13, step_Q13, 2 * ix[k][1] + 1);
    }
    *pred0 = p[0] - p[1];
    *pred1 = p[1];
}

// silk_stereo_MS_to_LR (stereo_MS_to_LR.c:34-85): x1 / x2 hold two samples
// of history then the frame; the side gets the interpolated prediction from
// the mid, then both become left / right.  Lane-parallel: every output
// depends only on inputs of the same sample and its two neighbours.
CUO_FN void silk_stereo_ms_to_lr(SilkDecoder *dec, int16_t *x1, int16_t *x2, int32_t pred0_Q13, int32_t pred1_Q13,
                                 int fs_kHz, int frame_length)
{
    x1[0] = dec->sMid[0]; x1[1] = dec->sMid[1];
    x2[0] = dec->sSide[0]; x2[1] = dec->sSide[1];
    const int16_t m0 = x1[frame_length], m1 = x1[frame_length + 1];
    const int16_t s0 = x2[frame_length], s1 = x2[frame_length + 1];
    dec->sMid[0] = m0; dec->sMid[1] = m1;
    dec->sSide[0] = s0; dec->sSide[1] = s1;
    const int32_t p0 = dec->pred_prev_Q13[0], p1 = dec->pred_prev_Q13[1];
    const int interp = 8 * fs_kHz;  // STEREO_INTERP_LEN_MS
    const int32_t denom_Q16 = (1 << 16) / interp;
    const int32_t delta0_Q13 = silk_rshift_round(silk_smulbb(pred0_Q13 - p0, denom_Q16), 16);
    const int32_t delta1_Q13 = silk_rshift_round(silk_smulbb(pred1_Q13 - p1, denom_Q16), 16);
    // side[n + 1] from mid[n .. n + 2] and side[n + 1]: computed into
    // registers first, stored after, so no lane reads an updated side sample
    for (int n0 = 0; n0 < frame_length; n0 += kLanes) {
        const int n = n0 + lane_id();
        int16_t out = 0;
        if (n < frame_length) {
            int32_t q0, q1;
            if (n < interp) { q0 = p0 + (n + 1) * delta0_Q13; q1 = p1 + (n + 1) * delta1_Q13; }
            else { q0 = pred0_Q13; q1 = pred1_Q13; }
            int32_t sum = silk_lshift(silk_add_lshift32(x1[n] + (int32_t)x1[n + 2], x1[n + 1], 1), 9);
            sum = silk_smlawb(silk_lshift((int32_t)x2[n + 1], 8), sum, q0);
            sum = silk_smlawb(sum, silk_lshift((int32_t)x1[n + 1], 11), q1);
            out = (int16_t)silk_sat16(silk_rshift_round(sum, 8));
        }
        if (n < frame_length) x2[n + 1] = out;
    }
    dec->pred_prev_Q13[0] = pred0_Q13;
    dec->pred_prev_Q13[1] = pred1_Q13;
    CUO_PFOR(n, 0, frame_length) {
        const int32_t sum = x1[n + 1] + (int32_t)x2[n + 1];
        const int32_t diff = x1[n + 1] - (int32_t)x2[n + 1];
        x1[n + 1] = (int16_t)silk_sat16(sum);
        x2[n + 1] = (int16_t)silk_sat16(diff);
    }
}

// silk_Decode (dec_API.c:132-452) at 48 kHz: one or two coded channels (mid /
// side) to one or two output channels (interleaved).  A stereo stream to one
// output channel renders the mid channel.  Returns 0 or a negative SILK
// error.
CUO_FN int silk_decode(SilkDecoder *dec, const SilkControl *ctl, int lostFlag, int newPacketFlag, RangeDec *d,
                       int16_t *samplesOut, int32_t *nSamplesOut, SilkScratch *sc, int16_t *rs_tmp)
{
    CUO_PROF_SCOPE(kProfSilk);
    dec = lds_uniform(dec);
    ctl = lds_uniform(ctl);
    d = lds_uniform(d);
    sc = lds_uniform(sc);
    const int nInt = kMaxCh == 1 ? 1 : uniform(ctl->nChannelsInternal);
    const int nAPI = kMaxCh == 1 ? 1 : uniform(ctl->nChannelsAPI);
    SilkChannel *ch0 = &dec->ch[0];
    int ret = 0;
    sc->cng_sig = dec->cng_sig;
    if (newPacketFlag)
        for (int n = 0; n < nInt; n++) dec->ch[n].nFramesDecoded = 0;
    if (kMaxCh == 2 && nInt > dec->nChannelsInternal) silk_reset_channel_of(dec, kMaxCh - 1);
    const int stereo_to_mono = kMaxCh == 2 && nInt == 1 && dec->nChannelsInternal == 2 &&
                               ctl->internalSampleRate == 1000 * ch0->fs_kHz;
    if (ch0->nFramesDecoded == 0) {
        for (int n = 0; n < nInt; n++) {
            SilkChannel *ch = &dec->ch[n];
            switch (ctl->payloadSize_ms) {
            case 0: case 10: ch->nFramesPerPacket = 1; ch->nb_subfr = 2; break;
            case 20: ch->nFramesPerPacket = 1; ch->nb_subfr = 4; break;
            case 40: ch->nFramesPerPacket = 2; ch->nb_subfr = 4; break;
            case 60: ch->nFramesPerPacket = 3; ch->nb_subfr = 4; break;
            default: return -203;  // SILK_DEC_INVALID_FRAME_SIZE
            }
            int fs_kHz_dec = (ctl->internalSampleRate >> 10) + 1;
            if (fs_kHz_dec != 8 && fs_kHz_dec != 12 && fs_kHz_dec != 16) return -200;
            silk_set_fs(ch, fs_kHz_dec, 48000);
        }
    }
    if (kMaxCh == 2 && nAPI == 2 && nInt == 2 && (dec->nChannelsAPI == 1 || dec->nChannelsInternal == 1)) {
        dec->pred_prev_Q13[0] = dec->pred_prev_Q13[1] = 0;
        dec->sSide[0] = dec->sSide[1] = 0;
        dec->ch[kMaxCh - 1].rs = ch0->rs;
    }
    dec->nChannelsAPI = nAPI;
    dec->nChannelsInternal = nInt;
    int decode_only_middle = 0;
    if (lostFlag != FLAG_LOST && ch0->nFramesDecoded == 0) {
        for (int n = 0; n < nInt; n++) {
            SilkChannel *ch = &dec->ch[n];
            for (int i = 0; i < ch->nFramesPerPacket; i++) ch->VAD_flags[i] = ec_bit_logp(d, 1);
            ch->LBRR_flag = ec_bit_logp(d, 1);
        }
        for (int n = 0; n < nInt; n++) {
            SilkChannel *ch = &dec->ch[n];
            for (int i = 0; i < 3; i++) ch->LBRR_flags[i] = 0;
            if (ch->LBRR_flag) {
                if (ch->nFramesPerPacket == 1) {
                    ch->LBRR_flags[0] = 1;
                } else {
                    const uint8_t *icdf = ch->nFramesPerPacket == 2 ? tab::kLbrrFlags2Icdf : tab::kLbrrFlags3Icdf;
                    int sym = ec_icdf(d, icdf, 8) + 1;
                    for (int i = 0; i < ch->nFramesPerPacket; i++) ch->LBRR_flags[i] = (sym >> i) & 1;
                }
            }
        }
        if (lostFlag == FLAG_NORMAL) {
            // skip the LBRR payload of every frame in this packet
            for (int i = 0; i < ch0->nFramesPerPacket; i++) {
                for (int n = 0; n < nInt; n++) {
                    SilkChannel *ch = &dec->ch[n];
                    if (ch->LBRR_flags[i]) {
                        if (nInt == 2 && n == 0) {
                            int32_t q0, q1;
                            silk_stereo_decode_pred(d, &q0, &q1);
                            if (dec->ch[kMaxCh - 1].LBRR_flags[i] == 0)
                                decode_only_middle = ec_icdf(d, tab::kStereoOnlyCodeMidIcdf, 8);
                        }
                        int cc = (i > 0 && ch->LBRR_flags[i - 1]) ? CODE_COND : CODE_INDEP;
                        silk_decode_indices(ch, d, i, 1, cc);
                        silk_decode_pulses(d, sc->pulses, ch->indices.signalType, ch->indices.quantOffsetType,
                                           ch->frame_length);
                    }
                }
            }
        }
    }
    int32_t pred0 = 0, pred1 = 0;
    if (nInt == 2) {
        SilkChannel *ch1 = &dec->ch[kMaxCh - 1];
        const int fi = ch0->nFramesDecoded;
        if (lostFlag == FLAG_NORMAL || (lostFlag == FLAG_LBRR && ch0->LBRR_flags[fi] == 1)) {
            silk_stereo_decode_pred(d, &pred0, &pred1);
            if ((lostFlag == FLAG_NORMAL && ch1->VAD_flags[fi] == 0) ||
                (lostFlag == FLAG_LBRR && ch1->LBRR_flags[fi] == 0))
                decode_only_middle = ec_icdf(d, tab::kStereoOnlyCodeMidIcdf, 8);
            else
                decode_only_middle = 0;
        } else {
            pred0 = dec->pred_prev_Q13[0];
            pred1 = dec->pred_prev_Q13[1];
        }
        if (decode_only_middle == 0 && dec->prev_decode_only_middle == 1) {
            // the side resumes after mid-only frames: from a clean history
            for (int i = 0; i < kSilkMaxFrame + 2 * 80; i++) ch1->outBuf[i] = 0;
            for (int i = 0; i < kSilkMaxOrder; i++) ch1->sLPC_Q14_buf[i] = 0;
            ch1->lagPrev = 100;
            ch1->LastGainIndex = 10;
            ch1->prevSignalType = SIG_NO_VOICE;
            ch1->first_frame_after_reset = 1;
        }
    }
    int has_side;
    if (lostFlag == FLAG_NORMAL)
        has_side = !decode_only_middle;
    else
        has_side = !dec->prev_decode_only_middle ||
                   (nInt == 2 && lostFlag == FLAG_LBRR &&
                    dec->ch[kMaxCh - 1].LBRR_flags[dec->ch[kMaxCh - 1].nFramesDecoded] == 1);
    int32_t nSamplesOutDec = 0;
    for (int n = 0; n < nInt; n++) {
        SilkChannel *ch = &dec->ch[n];
        int16_t *fo = sc->frame_out[n];   // [0..1] previous samples, [2..] this frame
        if (n == 0 || has_side) {
            const int FrameIndex = ch0->nFramesDecoded - n;
            int condCoding;
            if (FrameIndex <= 0) condCoding = CODE_INDEP;
            else if (lostFlag == FLAG_LBRR) condCoding = ch->LBRR_flags[FrameIndex - 1] ? CODE_COND : CODE_INDEP;
            else if (n > 0 && dec->prev_decode_only_middle) condCoding = CODE_INDEP_NO_LTP_SCALING;
            else condCoding = CODE_COND;
            sc->cng_exc = dec->cng_exc + kSilkMaxFrame * n;
            silk_decode_frame(ch, d, &fo[2], &nSamplesOutDec, lostFlag, condCoding, sc);
        } else {
            CUO_PFOR(i, 0, nSamplesOutDec) fo[2 + i] = 0;
        }
        ch->nFramesDecoded++;
    }
    int16_t *tmp0 = sc->frame_out[0];
    if (kMaxCh == 2 && nAPI == 2 && nInt == 2) {
        silk_stereo_ms_to_lr(dec, tmp0, sc->frame_out[kMaxCh - 1], pred0, pred1, ch0->fs_kHz, nSamplesOutDec);
    } else {
        tmp0[0] = dec->sMid[0];
        tmp0[1] = dec->sMid[1];
        dec->sMid[0] = tmp0[nSamplesOutDec];
        dec->sMid[1] = tmp0[nSamplesOutDec + 1];
    }
    const int32_t nOut = nSamplesOutDec * 48000 / silk_smulbb(ch0->fs_kHz, 1000);
    *nSamplesOut = nOut;
    if (kMaxCh == 1 || nAPI == 1) {
        silk_resample(&ch0->rs, samplesOut, &tmp0[1], nSamplesOutDec, sc->rs_buf);
    } else {
        // two output channels: each resampled on its own, then interleaved
        for (int n = 0; n < 2; n++) {
            const bool own = n < nInt;
            if (!own && !stereo_to_mono) {
                CUO_PFOR(i, 0, nOut) samplesOut[2 * i + 1] = samplesOut[2 * i];
                break;
            }
            const int16_t *in = own ? &sc->frame_out[n][1] : &tmp0[1];
            silk_resample(&dec->ch[n < kMaxCh ? n : 0].rs, rs_tmp, in, nSamplesOutDec, sc->rs_buf);
            CUO_PFOR(i, 0, nOut) samplesOut[2 * i + n] = rs_tmp[i];
        }
    }
    if (lostFlag == FLAG_LOST) {
        for (int n = 0; n < nInt; n++) dec->ch[n].LastGainIndex = 10;
    } else {
        dec->prev_decode_only_middle = decode_only_middle;
    }
    return ret;
}

}  // namespace CUO_NS
