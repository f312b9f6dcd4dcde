// logmel.h -- log-mel front-end over the resident f32 16 kHz output.
//
// The reason the reference keeps decoded PCM on the device is to feed a
// speech model's feature front-end without a host round trip
// (ROADMAP.md:149-156).  This kernel is that front-end: frame t covers input
// samples [160 t, 160 t + 400) (25 ms every 10 ms, no padding at the edges),
// multiplied by a periodic Hann window of length 400, zero-padded to 512; the
// power spectrum |X_k|^2, k = 0..256, is weighted by n_mels triangular mel
// filters (Slaney mel scale and Slaney area normalisation, 0-8000 Hz) and the
// output is ln(max(mel, 1e-10)), row-major (frames, n_mels).
//
// One wavefront per frame, four per workgroup, each wavefront taking
// kMelFramesPerWave consecutive frames in turn.  Each lane keeps the window
// values, butterfly twiddles and split twiddles it uses in registers across
// those frames, and the workgroup first copies the mel filter tables into
// LDS, so the frame loop reads no table from global memory.  The 512-point
// real transform runs as a 256-point complex FFT of z_m = y_2m + i y_2m+1
// (Stockham radix-4, four stages, one butterfly per lane per stage, ping-pong
// in LDS, natural-order output) followed by the real-input split
// X_k = (Z_k + conj Z_256-k) / 2 - i W512^k (Z_k - conj Z_256-k) / 2.
// Window, twiddles and filter weights are computed on the host in double,
// rounded to float and kept in device tables.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cuo {

constexpr int kMelFft = 512;
constexpr int kMelWin = 400;
constexpr int kMelHop = 160;
constexpr int kMelBins = kMelFft / 2 + 1;  // 257
constexpr int kMelMaxMels = 128;
constexpr int kMelMaxWeights = 2 * kMelBins + 2 * kMelMaxMels;  // each FFT bin lies in <= 2 filters
constexpr int kMelWaves = 4;           // wavefronts per workgroup
constexpr int kMelFramesPerWave = 8;   // frames each wavefront takes in turn
constexpr int kMelFramesPerBlock = kMelWaves * kMelFramesPerWave;

struct MelFilters {
    int32_t first[kMelMaxMels];   // first FFT bin of filter m
    int32_t count[kMelMaxMels];   // its bin count
    int32_t woff[kMelMaxMels];    // offset of its weights
    float weight[kMelMaxWeights];
};

struct MelTables {
    float window[kMelWin];
    float tw256[2 * 256];         // W256^e = (cos, -sin)(2 pi e / 256), e = 0..255
    float tw512[2 * kMelBins];    // W512^k, k = 0..256
    MelFilters filt;
};

__device__ __forceinline__ float2 cmul(float2 a, float2 b)
{
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

__global__ __launch_bounds__(64 * kMelWaves) __attribute__((amdgpu_waves_per_eu(6))) void k_log_mel(const float *__restrict__ x, int64_t n_frames,
                                                           const MelTables *__restrict__ gtab, int n_mels,
                                                           float *__restrict__ out)
{
    __shared__ MelFilters filt;
    __shared__ float2 buf[kMelWaves][2][256];
    __shared__ float pw[kMelWaves][kMelBins + 3];
    {
        // the filter tables, once per workgroup (their size is a multiple of 4 bytes)
        static_assert(sizeof(MelFilters) % 4 == 0, "MelFilters size");
        const uint32_t *src = (const uint32_t *)&gtab->filt;
        uint32_t *dst = (uint32_t *)&filt;
        for (int i = (int)threadIdx.x; i < (int)(sizeof(MelFilters) / 4); i += 64 * kMelWaves) dst[i] = src[i];
    }
    __syncthreads();
    const MelTables *tab = gtab;
    const int w = (int)threadIdx.x >> 6;
    const int lane = (int)threadIdx.x & 63;
    float2 *b0 = buf[w][0], *b1 = buf[w][1];
    float *p = pw[w];
    // what each lane reads of the tables is the same for every frame: the
    // window at its samples, its butterflies' twiddles, its bins' split
    // twiddles -- kept in registers across the wavefront's frames
    float win[8];
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const int j = 2 * (lane + 64 * i);
        win[2 * i] = j < kMelWin ? tab->window[j] : 0.f;
        win[2 * i + 1] = j + 1 < kMelWin ? tab->window[j + 1] : 0.f;
    }
    float2 tw[3][3];  // stages Ns = 4, 16, 64; r = 1..3 (stage Ns = 1 has only W^0)
#pragma unroll
    for (int st = 0; st < 3; st++) {
        const int ns = 4 << (2 * st);
        const int k = lane & (ns - 1);
#pragma unroll
        for (int r = 1; r < 4; r++) {
            const int e = (r * k * (64 / ns)) & 255;  // W_{4 Ns}^{r k} = W256^{r k 64 / Ns}
            tw[st][r - 1] = make_float2(tab->tw256[2 * e], tab->tw256[2 * e + 1]);
        }
    }
    float2 tws[5];  // W512^k for this lane's bins k = lane + 64 i
#pragma unroll
    for (int i = 0; i < 5; i++) {
        const int kk = lane + 64 * i;
        tws[i] = kk < kMelBins ? make_float2(tab->tw512[2 * kk], tab->tw512[2 * kk + 1]) : make_float2(0.f, 0.f);
    }
    const int64_t t0 = (int64_t)blockIdx.x * kMelFramesPerBlock + (int64_t)w * kMelFramesPerWave;
    for (int fi = 0; fi < kMelFramesPerWave; fi++) {
        const int64_t t = t0 + fi;
        if (t >= n_frames) break;  // wave-uniform; no block barrier below this point
        const float *xf = x + t * kMelHop;
        // windowed frame as 256 complex values (pairs of real samples), zero past 400
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int m = lane + 64 * i;
            const int j = 2 * m;
            const float a = j < kMelWin ? xf[j] * win[2 * i] : 0.f;
            const float c = j + 1 < kMelWin ? xf[j + 1] * win[2 * i + 1] : 0.f;
            b0[m] = make_float2(a, c);
        }
        __builtin_amdgcn_wave_barrier();
        // 256-point complex FFT, Stockham radix-4: stage Ns = 1, 4, 16, 64
        float2 *src = b0, *dst = b1;
#pragma unroll
        for (int ns = 1, st = -1; ns < 256; ns *= 4, st++) {
            const int j = lane;            // 64 butterflies per stage
            const int k = j & (ns - 1);    // j mod Ns
            float2 v[4];
#pragma unroll
            for (int r = 0; r < 4; r++) {
                v[r] = src[j + 64 * r];
                if (r && st >= 0) v[r] = cmul(v[r], tw[st][r - 1]);
            }
            // 4-point DFT (W4 = -i)
            const float2 s02 = make_float2(v[0].x + v[2].x, v[0].y + v[2].y);
            const float2 d02 = make_float2(v[0].x - v[2].x, v[0].y - v[2].y);
            const float2 s13 = make_float2(v[1].x + v[3].x, v[1].y + v[3].y);
            const float2 d13 = make_float2(v[1].x - v[3].x, v[1].y - v[3].y);
            const int base = (j / ns) * ns * 4 + k;
            __builtin_amdgcn_wave_barrier();
            dst[base] = make_float2(s02.x + s13.x, s02.y + s13.y);
            dst[base + ns] = make_float2(d02.x + d13.y, d02.y - d13.x);      // d02 - i d13
            dst[base + 2 * ns] = make_float2(s02.x - s13.x, s02.y - s13.y);
            dst[base + 3 * ns] = make_float2(d02.x - d13.y, d02.y + d13.x);  // d02 + i d13
            __builtin_amdgcn_wave_barrier();
            float2 *tmp = src;
            src = dst;
            dst = tmp;
        }
        // real-input split and power spectrum, k = 0..256
#pragma unroll
        for (int i = 0; i < 5; i++) {
            const int kk = lane + 64 * i;
            if (kk >= kMelBins) break;
            const float2 zk = src[kk & 255];
            const float2 zn = src[(256 - kk) & 255];
            // E = (Z_k + conj Z_n) / 2, O = (Z_k - conj Z_n) / 2i
            const float2 e = make_float2(0.5f * (zk.x + zn.x), 0.5f * (zk.y - zn.y));
            const float2 o = make_float2(0.5f * (zk.y + zn.y), -0.5f * (zk.x - zn.x));
            const float2 xo = cmul(tws[i], o);
            const float re = e.x + xo.x, im = e.y + xo.y;
            p[kk] = re * re + im * im;
        }
        __builtin_amdgcn_wave_barrier();
        float *o = out + t * n_mels;
        for (int m = lane; m < n_mels; m += 64) {
            const int f0 = filt.first[m], nc = filt.count[m];
            const float *wt = filt.weight + filt.woff[m];
            float acc = 0.f;
            for (int i = 0; i < nc; i++) acc = __builtin_fmaf(wt[i], p[f0 + i], acc);
            o[m] = logf(acc > 1e-10f ? acc : 1e-10f);
        }
        // the next frame rewrites this wavefront's buffers
        __builtin_amdgcn_wave_barrier();
    }
}

}  // namespace cuo
