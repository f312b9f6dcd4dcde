/* gen_tables.c -- emit the constant tables the HIP decoder needs.
 *
 * Bit-exact decoding requires the codec's constant tables to be numerically
 * identical to libopus 1.5.2's (RFC 6716 normative tables, CELT 48 kHz/960
 * static mode, SILK codebooks).  This generator compiles against the libopus
 * sources that ship with the reference (vendor/opus-1.5.2, FIXED_POINT) and
 * prints the numbers as C++ constexpr arrays in this project's own layout
 * (concatenated where the decoder wants one array plus offsets).  Only
 * numbers are emitted; no code.
 *
 * Build + run (needs the reference tree; the output header is committed):
 *   make -C cuopus_amd/csrc/gen
 */
#include <stdio.h>
#include <stdlib.h>

#include "modes.h"
#include "modes.c" /* defines eband5ms/band_allocation and includes the static mode */
#include "celt.h"
#include "tables.h" /* silk */
#include "pitch_est_defines.h"
#include "resampler_rom.h"

/* Pull the static data tables of these translation units into this one. */
#include "cwrs.c"
#include "quant_bands.c"
#include "rate.c"
#include "bands.c"

static FILE *o;

#define EMIT(ctype, name, arr, n)                                            \
    do {                                                                     \
        fprintf(o, "alignas(4) static constexpr %s %s[%d] = {", ctype, name, (int)(n)); \
        for (int i_ = 0; i_ < (int)(n); i_++)                                \
            fprintf(o, "%s%s%ld", i_ ? "," : "", (i_ % 16) ? "" : "\n  ",    \
                    (long)(arr)[i_]);                                        \
        fprintf(o, "};\n");                                                  \
    } while (0)

#define EMITU(ctype, name, arr, n)                                           \
    do {                                                                     \
        fprintf(o, "alignas(4) static constexpr %s %s[%d] = {", ctype, name, (int)(n)); \
        for (int i_ = 0; i_ < (int)(n); i_++)                                \
            fprintf(o, "%s%s%luu", i_ ? "," : "", (i_ % 12) ? "" : "\n  ",   \
                    (unsigned long)(arr)[i_]);                               \
        fprintf(o, "};\n");                                                  \
    } while (0)

static void emit_nlsf(const char *tag, const silk_NLSF_CB_struct *cb)
{
    char nm[128];
    int nv = cb->nVectors, ord = cb->order;
    fprintf(o, "static constexpr int kNlsf%s_nVectors = %d;\n", tag, nv);
    fprintf(o, "static constexpr int kNlsf%s_order = %d;\n", tag, ord);
    fprintf(o, "static constexpr int kNlsf%s_quantStep_Q16 = %d;\n", tag, cb->quantStepSize_Q16);
    fprintf(o, "static constexpr int kNlsf%s_invQuantStep_Q6 = %d;\n", tag, cb->invQuantStepSize_Q6);
    snprintf(nm, sizeof nm, "kNlsf%s_CB1_Q8", tag);     EMIT("uint8_t", nm, cb->CB1_NLSF_Q8, nv * ord);
    snprintf(nm, sizeof nm, "kNlsf%s_CB1_Wght_Q9", tag); EMIT("int16_t", nm, cb->CB1_Wght_Q9, nv * ord);
    snprintf(nm, sizeof nm, "kNlsf%s_CB1_iCDF", tag);    EMIT("uint8_t", nm, cb->CB1_iCDF, 2 * nv);
    snprintf(nm, sizeof nm, "kNlsf%s_pred_Q8", tag);     EMIT("uint8_t", nm, cb->pred_Q8, 2 * (ord - 1));
    snprintf(nm, sizeof nm, "kNlsf%s_ec_sel", tag);      EMIT("uint8_t", nm, cb->ec_sel, nv * ord / 2);
    snprintf(nm, sizeof nm, "kNlsf%s_ec_iCDF", tag);     EMIT("uint8_t", nm, cb->ec_iCDF, 72);
    snprintf(nm, sizeof nm, "kNlsf%s_deltaMin_Q15", tag); EMIT("int16_t", nm, cb->deltaMin_Q15, ord + 1);
}

int main(int argc, char **argv)
{
    o = fopen(argc > 1 ? argv[1] : "tables_gen.h", "w");
    if (!o) { perror("open"); return 1; }
    const CELTMode *m = &mode48000_960_120;
    fprintf(o, "// Generated by cuopus_amd/csrc/gen/gen_tables.c -- numeric constant tables of\n"
               "// the Opus decoder (RFC 6716; libopus 1.5.2 fixed-point values).  Do not edit.\n"
               "#pragma once\n#include <stdint.h>\nnamespace CUO_NS { namespace tab {\n");

    /* ---------------- CELT mode 48000/960/120 ---------------- */
    fprintf(o, "static constexpr int kNbEBands = %d;\n", m->nbEBands);
    fprintf(o, "static constexpr int kEffEBands = %d;\n", m->effEBands);
    fprintf(o, "static constexpr int kOverlap = %d;\n", m->overlap);
    fprintf(o, "static constexpr int kMaxLM = %d;\n", m->maxLM);
    fprintf(o, "static constexpr int kNbShortMdcts = %d;\n", m->nbShortMdcts);
    fprintf(o, "static constexpr int kShortMdctSize = %d;\n", m->shortMdctSize);
    fprintf(o, "static constexpr int kNbAllocVectors = %d;\n", m->nbAllocVectors);
    EMIT("int16_t", "kPreemph", m->preemph, 4);
    EMIT("int16_t", "kEBands", m->eBands, m->nbEBands + 1);
    EMIT("uint8_t", "kAllocVectors", m->allocVectors, m->nbAllocVectors * m->nbEBands);
    EMIT("int16_t", "kWindow", m->window, m->overlap);
    EMIT("int16_t", "kLogN", m->logN, m->nbEBands);
    fprintf(o, "static constexpr int kCacheSize = %d;\n", m->cache.size);
    EMIT("int16_t", "kCacheIndex", m->cache.index, (m->maxLM + 2) * m->nbEBands);
    EMIT("uint8_t", "kCacheBits", m->cache.bits, m->cache.size);
    EMIT("uint8_t", "kCacheCaps", m->cache.caps, (m->maxLM + 1) * 2 * m->nbEBands);

    /* FFTs: 4 sizes sharing one twiddle table; bitrev tables concatenated */
    {
        const kiss_fft_state *st[4];
        for (int k = 0; k < 4; k++) st[k] = m->mdct.kfft[k];
        int tw[960];
        for (int i = 0; i < 480; i++) { tw[2 * i] = st[0]->twiddles[i].r; tw[2 * i + 1] = st[0]->twiddles[i].i; }
        EMIT("int16_t", "kFftTwiddles", tw, 960);
        int br[900], boff[4], pos = 0, fac[64], ss[4], sh[4], sc[4], nf[4];
        for (int k = 0; k < 4; k++) {
            boff[k] = pos;
            for (int i = 0; i < st[k]->nfft; i++) br[pos++] = st[k]->bitrev[i];
            for (int i = 0; i < 16; i++) fac[16 * k + i] = st[k]->factors[i];
            ss[k] = st[k]->scale_shift; sh[k] = st[k]->shift; sc[k] = st[k]->scale; nf[k] = st[k]->nfft;
            if (st[k]->twiddles != st[0]->twiddles) { fprintf(stderr, "twiddle table differs\n"); return 1; }
        }
        EMIT("int16_t", "kFftBitrev", br, pos);
        EMIT("int16_t", "kFftBitrevOff", boff, 4);
        EMIT("int16_t", "kFftFactors", fac, 64);
        EMIT("int16_t", "kFftNfft", nf, 4);
        EMIT("int16_t", "kFftScale", sc, 4);
        EMIT("int16_t", "kFftScaleShift", ss, 4);
        EMIT("int16_t", "kFftShift", sh, 4);
    }
    fprintf(o, "static constexpr int kMdctN = %d;\n", m->mdct.n);
    EMIT("int16_t", "kMdctTrig", m->mdct.trig, 1800);

    /* PVQ codeword counts U(N,K), rows as offsets into one array */
    {
        int roff[15];
        for (int r = 0; r < 15; r++) roff[r] = (int)(CELT_PVQ_U_ROW[r] - CELT_PVQ_U_DATA);
        EMITU("uint32_t", "kPvqU", CELT_PVQ_U_DATA, (int)(sizeof(CELT_PVQ_U_DATA) / sizeof(CELT_PVQ_U_DATA[0])));
        EMIT("int16_t", "kPvqURow", roff, 15);
    }
    /* energy quantisation */
    EMIT("uint8_t", "kEProbModel", &e_prob_model[0][0][0], 4 * 2 * 42);
    EMIT("int16_t", "kPredCoef", pred_coef, 4);
    EMIT("int16_t", "kBetaCoef", beta_coef, 4);
    fprintf(o, "static constexpr int kBetaIntra = %d;\n", (int)beta_intra);
    EMIT("uint8_t", "kSmallEnergyIcdf", small_energy_icdf, 3);
    EMIT("int8_t", "kEMeans", eMeans, 25);
    EMIT("uint8_t", "kLog2FracTable", LOG2_FRAC_TABLE, 24);
    EMIT("int8_t", "kTfSelect", &tf_select_table[0][0], 32);
    EMIT("uint8_t", "kTrimIcdf", trim_icdf, 11);
    EMIT("uint8_t", "kSpreadIcdf", spread_icdf, 4);
    EMIT("uint8_t", "kTapsetIcdf", tapset_icdf, 3);
    EMIT("int16_t", "kOrderyTable", ordery_table, 30);

    /* ---------------- SILK ---------------- */
    EMIT("uint8_t", "kGainIcdf", &silk_gain_iCDF[0][0], 24);
    EMIT("uint8_t", "kDeltaGainIcdf", silk_delta_gain_iCDF, 41);
    EMIT("uint8_t", "kPitchLagIcdf", silk_pitch_lag_iCDF, 32);
    EMIT("uint8_t", "kPitchDeltaIcdf", silk_pitch_delta_iCDF, 21);
    EMIT("uint8_t", "kPitchContourIcdf", silk_pitch_contour_iCDF, 34);
    EMIT("uint8_t", "kPitchContourNBIcdf", silk_pitch_contour_NB_iCDF, 11);
    EMIT("uint8_t", "kPitchContour10msIcdf", silk_pitch_contour_10_ms_iCDF, 12);
    EMIT("uint8_t", "kPitchContour10msNBIcdf", silk_pitch_contour_10_ms_NB_iCDF, 3);
    EMIT("uint8_t", "kPulsesPerBlockIcdf", &silk_pulses_per_block_iCDF[0][0], 10 * 18);
    EMIT("uint8_t", "kRateLevelsIcdf", &silk_rate_levels_iCDF[0][0], 18);
    EMIT("uint8_t", "kMaxPulsesTable", silk_max_pulses_table, 4);
    EMIT("uint8_t", "kShellTable0", silk_shell_code_table0, 152);
    EMIT("uint8_t", "kShellTable1", silk_shell_code_table1, 152);
    EMIT("uint8_t", "kShellTable2", silk_shell_code_table2, 152);
    EMIT("uint8_t", "kShellTable3", silk_shell_code_table3, 152);
    EMIT("uint8_t", "kShellTableOffsets", silk_shell_code_table_offsets, 17);
    EMIT("uint8_t", "kLsbIcdf", silk_lsb_iCDF, 2);
    EMIT("uint8_t", "kSignIcdf", silk_sign_iCDF, 42);
    EMIT("uint8_t", "kUniform3Icdf", silk_uniform3_iCDF, 3);
    EMIT("uint8_t", "kUniform4Icdf", silk_uniform4_iCDF, 4);
    EMIT("uint8_t", "kUniform5Icdf", silk_uniform5_iCDF, 5);
    EMIT("uint8_t", "kUniform6Icdf", silk_uniform6_iCDF, 6);
    EMIT("uint8_t", "kUniform8Icdf", silk_uniform8_iCDF, 8);
    EMIT("uint8_t", "kNlsfExtIcdf", silk_NLSF_EXT_iCDF, 7);
    EMIT("uint8_t", "kLtpPerIndexIcdf", silk_LTP_per_index_iCDF, 3);
    EMIT("uint8_t", "kLtpGainIcdf0", silk_LTP_gain_iCDF_ptrs[0], 8);
    EMIT("uint8_t", "kLtpGainIcdf1", silk_LTP_gain_iCDF_ptrs[1], 16);
    EMIT("uint8_t", "kLtpGainIcdf2", silk_LTP_gain_iCDF_ptrs[2], 32);
    EMIT("int8_t", "kLtpVq0_Q7", silk_LTP_vq_ptrs_Q7[0], 8 * 5);
    EMIT("int8_t", "kLtpVq1_Q7", silk_LTP_vq_ptrs_Q7[1], 16 * 5);
    EMIT("int8_t", "kLtpVq2_Q7", silk_LTP_vq_ptrs_Q7[2], 32 * 5);
    EMIT("int8_t", "kLtpVqSizes", silk_LTP_vq_sizes, 3);
    EMIT("uint8_t", "kLtpScaleIcdf", silk_LTPscale_iCDF, 3);
    EMIT("int16_t", "kLtpScales_Q14", silk_LTPScales_table_Q14, 3);
    EMIT("uint8_t", "kTypeOffsetVadIcdf", silk_type_offset_VAD_iCDF, 4);
    EMIT("uint8_t", "kTypeOffsetNoVadIcdf", silk_type_offset_no_VAD_iCDF, 2);
    EMIT("int16_t", "kStereoPredQuant_Q13", silk_stereo_pred_quant_Q13, 16);
    EMIT("uint8_t", "kStereoPredJointIcdf", silk_stereo_pred_joint_iCDF, 25);
    EMIT("uint8_t", "kStereoOnlyCodeMidIcdf", silk_stereo_only_code_mid_iCDF, 2);
    EMIT("uint8_t", "kLbrrFlags2Icdf", silk_LBRR_flags_iCDF_ptr[0], 3);
    EMIT("uint8_t", "kLbrrFlags3Icdf", silk_LBRR_flags_iCDF_ptr[1], 7);
    EMIT("uint8_t", "kNlsfInterpFactorIcdf", silk_NLSF_interpolation_factor_iCDF, 5);
    EMIT("int16_t", "kQuantOffsets_Q10", &silk_Quantization_Offsets_Q10[0][0], 4);
    EMIT("int16_t", "kLsfCos_Q12", silk_LSFCosTab_FIX_Q12, 129);
    emit_nlsf("NBMB", &silk_NLSF_CB_NB_MB);
    emit_nlsf("WB", &silk_NLSF_CB_WB);
    EMIT("int8_t", "kCbLagsStage2", &silk_CB_lags_stage2[0][0], 4 * 11);
    EMIT("int8_t", "kCbLagsStage3", &silk_CB_lags_stage3[0][0], 4 * 34);
    EMIT("int8_t", "kCbLagsStage2_10ms", &silk_CB_lags_stage2_10_ms[0][0], 4 * 3);
    EMIT("int8_t", "kCbLagsStage3_10ms", &silk_CB_lags_stage3_10_ms[0][0], 4 * 12);
    /* resampler ROM */
    EMIT("int16_t", "kResFracFir12", &silk_resampler_frac_FIR_12[0][0], 12 * 4);
    EMIT("int16_t", "kResUp2HQ0", silk_resampler_up2_hq_0, 3);
    EMIT("int16_t", "kResUp2HQ1", silk_resampler_up2_hq_1, 3);

    fprintf(o, "}}  // namespace CUO_NS::tab\n");
    fclose(o);
    return 0;
}
