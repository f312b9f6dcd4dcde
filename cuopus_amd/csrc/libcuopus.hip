// libcuopus.hip -- host runtime + C ABI of the MI355X batch Opus decoder.
//
// Drop-in for the reference's libcuopus.cu (same entry points and contract,
// see include/cuopus.h).  Per call: demux every file on the host
// (host/demux.cpp), merge all chains into one table, upload packets + tables,
// run the decode kernel (one wavefront per chain, keep-window compaction fused
// into its epilogue), run one batched resampler launch for every file, copy
// the per-chain results back, synchronise, and validate every chain (stop
// code 0 and decoded samples == the chain's TOC capacity) before reporting
// success (libcuopus.cu:436-452).
#include <hip/hip_runtime.h>
#include <dlfcn.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <string>
#include <vector>
#include <array>
#include <atomic>
#include <deque>
#include <memory>

#include "../../include/cuopus.h"
#include "host/demux.h"
#include "kernels.h"
#include "logmel.h"
#include "stereo_launch.h"

using namespace cuo;
using cuo_host::Chain;
using cuo_host::Demuxed;

namespace {

struct Err {
    char *buf;
    size_t cap;
    char local[8];
    Err(char *b, size_t c) : buf(b), cap(c)
    {
        if (!buf || !cap) { buf = local; cap = sizeof(local); }
        buf[0] = 0;
    }
    int set(int rc, const char *fmt, ...) __attribute__((format(printf, 3, 4)))
    {
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, cap, fmt, ap);
        va_end(ap);
        return rc;
    }
};

#define HIPCHK(e, call)                                                                                 \
    do {                                                                                                \
        hipError_t _s = (call);                                                                         \
        if (_s != hipSuccess)                                                                           \
            return (e).set(CUOPUS_RC_GPU, "HIP error %s at %s:%d: %s", #call, __FILE__, __LINE__,       \
                           hipGetErrorString(_s));                                                      \
    } while (0)

struct DevBuf {
    void *p = nullptr;
    size_t cap = 0;
};

long long g_dev_bytes = 0;

int grow(DevBuf &b, size_t need, Err &e)
{
    if (need == 0) need = 1;
    if (need <= b.cap) return 0;
    if (b.p) {
        hipFree(b.p);
        g_dev_bytes -= (long long)b.cap;
        b.p = nullptr;
        b.cap = 0;
    }
    hipError_t s = hipMalloc(&b.p, need);
    if (s != hipSuccess) {
        b.p = nullptr;
        return e.set(CUOPUS_RC_GPU, "hipMalloc %zu bytes: %s", need, hipGetErrorString(s));
    }
    b.cap = need;
    g_dev_bytes += (long long)need;
    return 0;
}

struct Workspaces {
    DevBuf data, off, len, lost, chains, order, cat, res, nsamp, rng, files, queue, heap;
} W;

double env_f(const char *name, double dflt)
{
    const char *v = getenv(name);
    return (v && *v) ? atof(v) : dflt;
}
unsigned env_u(const char *name, unsigned dflt)
{
    const char *v = getenv(name);
    return (v && atoi(v) > 0) ? (unsigned)atoi(v) : dflt;
}

// Launch wrappers (64-lane workgroups, one wavefront per chain).
void launch_chains(uint32_t grid, hipStream_t st, const DecodeArgs &a)
{
    hipLaunchKernelGGL(k_decode_chains, dim3(grid), dim3(64), 0, st, a);
}

void launch_persistent(uint32_t grid, hipStream_t st, const DecodeArgs &a)
{
    hipLaunchKernelGGL(k_decode_persistent, dim3(grid), dim3(64), 0, st, a);
}

// The decode launch of either kernel build (stereo: the stereo-capable one).
void launch_decode(bool stereo, bool persistent, uint32_t grid, hipStream_t st, const DecodeArgs &a)
{
    if (stereo) cuo_stereo_launch(persistent ? 1 : 0, grid, st, &a);
    else if (persistent) launch_persistent(grid, st, a);
    else launch_chains(grid, st, a);
}

// HBM heap bytes per wavefront of the kernel build that decodes a launch.
size_t heap_bytes(bool stereo) { return stereo ? cuo_stereo_heap_bytes() : sizeof(DecHeap); }

// Kaiser I0 by its power series, in double.
double bessel_i0(double x)
{
    double s = 1.0, t = 1.0, q = x * x / 4.0;
    for (int k = 1; k < 64; k++) {
        t *= q / ((double)k * k);
        s += t;
        if (t < 1e-18 * s) break;
    }
    return s;
}

// 49 taps: sinc at cutoff 1/3 x Kaiser(8.6), unity DC, double -> float
// (cuopus_kernels.cuh:481-494, gpu_resample.py:23-31).
RsTaps make_taps()
{
    RsTaps t;
    double h[kRsTaps], sum = 0.0;
    const double fc = 1.0 / 3.0;
    for (int i = 0; i < kRsTaps; i++) {
        double n = i - (kRsTaps - 1) / 2.0;
        double s = (n == 0.0) ? fc : sin(M_PI * fc * n) / (M_PI * n);
        double r = 2.0 * i / (kRsTaps - 1) - 1.0;
        double w = bessel_i0(8.6 * sqrt(1.0 - r * r)) / bessel_i0(8.6);
        h[i] = s * w;
        sum += h[i];
    }
    for (int i = 0; i < kRsTaps; i++) t.t[i] = (float)(h[i] / sum);
    return t;
}

// ---------------------------------------------------------------- tuned defaults
// Per-device launch defaults written by the autotuner (cuopus_amd/autotune.py)
// to cuopus_tuned.txt beside libcuopus.so, or the file CUOPUS_TUNED names.
// One line per device: "<gcnArchName>/<CU count> slots=<n> rs=<variant> order=<o>".
// Read once per process and device; absent file or device: built-in defaults.
struct TunedDefaults {
    uint32_t slots = 0;  // persistent wavefronts (0: kDefaultWaves)
    int rs = -1;         // resampler variant (-1: kRsDefaultVariant)
    int order = -1;      // work order (-1: the cost model)
};

std::string lib_dir()
{
    Dl_info info;
    if (dladdr((void *)&lib_dir, &info) && info.dli_fname) {
        std::string s = info.dli_fname;
        size_t k = s.rfind('/');
        return k == std::string::npos ? std::string(".") : s.substr(0, k);
    }
    return ".";
}

const TunedDefaults &tuned_defaults()
{
    static TunedDefaults cache[64];
    static bool loaded[64] = {false};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) dev = 0;
    if (loaded[dev]) return cache[dev];
    loaded[dev] = true;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, dev) != hipSuccess) return cache[dev];
    char key[160];
    std::string arch = prop.gcnArchName;
    arch = arch.substr(0, arch.find(':'));
    snprintf(key, sizeof(key), "%s/%d", arch.c_str(), prop.multiProcessorCount);
    const char *env = getenv("CUOPUS_TUNED");
    std::string path = env && *env ? std::string(env) : lib_dir() + "/cuopus_tuned.txt";
    FILE *f = fopen(path.c_str(), "r");
    if (!f) return cache[dev];
    char line[512];
    while (fgets(line, sizeof(line), f)) {
        char k[160];
        if (sscanf(line, "%159s", k) != 1 || strcmp(k, key) != 0) continue;
        const char *s = strstr(line, "slots=");
        if (s) cache[dev].slots = (uint32_t)strtoul(s + 6, nullptr, 10);
        const char *r = strstr(line, "rs=");
        if (r) cache[dev].rs = atoi(r + 3);
        const char *o = strstr(line, "order=");
        if (o) cache[dev].order = atoi(o + 6);
    }
    fclose(f);
    return cache[dev];
}

// Per-frame decode costs of the work-order model, by TOC, in microseconds of
// one wavefront on a loaded MI355X: fitted (non-negative least squares of
// per-chain device times on the chains' frame counts, cuopus_amd/autotune.py
// calibrate) on the 2 h CELT workload and the 50-file mixed set.  CELT by
// frame size 2.5/5/10/20 ms (a 2.5 ms frame costs a quarter of a 20 ms one,
// so twice as much per second of audio); hybrid 20 ms and SILK 20 ms fitted,
// hybrid 10 ms and SILK 10/40/60 ms scaled from them; a lost packet
// (concealment) about a frame.
constexpr uint32_t kCostCelt[4] = {134, 227, 315, 539};
constexpr uint32_t kCostHybrid[2] = {280, 461};
constexpr uint32_t kCostSilk[4] = {130, 236, 470, 700};
constexpr uint32_t kCostLost = 200;

// Resampler launch: variant v picks outputs per thread (rs_variants table).
int resample_variant()
{
    unsigned v = env_u("CUOPUS_RS", 0);
    if (v) return (int)v - 1;
    const int t = tuned_defaults().rs;
    return t >= 0 && t < kRsVariants ? t : kRsDefaultVariant;
}

uint64_t rs_tiles(uint64_t out_len, int v) { return (out_len + rs_tile(v) - 1) / rs_tile(v); }

void launch_resample(uint64_t tiles, hipStream_t st, const int16_t *in, const RsFile *files, int nfiles,
                     const RsTaps &taps, int v)
{
    switch (v) {
#define CUO_RS(V)                                                                                              \
    case V:                                                                                                    \
        hipLaunchKernelGGL((k_resample_48to16<kRsShape[V][0], kRsShape[V][1]>), dim3((unsigned)tiles),          \
                           dim3(kRsShape[V][0]), 0, st, in, files, nfiles, taps);                              \
        break;
        CUO_RS(1)
        CUO_RS(2)
        CUO_RS(3)
        CUO_RS(4)
        CUO_RS(0)
#undef CUO_RS
    default: break;
    }
}

// ---------------------------------------------------------------- planning
struct FilePlan {
    std::shared_ptr<const Demuxed> d;  // shared by every plan that decodes this file (never copied)
    std::vector<Chain> chains;
    uint64_t chain0 = 0;     // first merged chain index
    uint64_t keep = 0;       // playable samples (per channel)
    uint64_t cat_off = 0;    // offset of the file in the compacted output (int16 elements)
    int64_t out16 = 0;
    int out_channels = 1;    // decoder output channels: 1 (mono render, the f32 path) or the stream's
};

// The decoder reads packet bytes as aligned 32-bit words and never depends
// on a byte past a packet; the packet buffer still ends in this many zero
// bytes so the last word read is inside it and holds nothing stale.
constexpr size_t kPktPad = 4;

struct Plan {
    std::vector<FilePlan> files;
    std::vector<uint8_t> data;   // every file's packets back to back + kPktPad zero bytes
    std::vector<uint32_t> off, len;
    std::vector<uint8_t> lost;
    std::vector<ChainDesc> chains;
    std::vector<uint64_t> cap;   // per chain: sum of TOC durations
    std::vector<uint32_t> cost;  // per chain: estimated decode cost (work order)
    uint64_t total_keep = 0;  // playable samples of all files
    uint64_t cat_len = 0;     // int16 elements of the compacted PCM buffer (file starts padded to 8)
    bool stereo = false;      // some stream is stereo: the stereo-capable kernel build decodes the launch
    // Merged with borrow = true: data and len stay empty and each file's own
    // packet bytes and lengths are uploaded in place (then kPktPad zeros);
    // off is built only when there are several files (shifted offsets).
    bool borrowed = false;
    uint64_t npk_total = 0, nbytes_total = 0;
    uint64_t n_packets() const { return borrowed ? npk_total : off.size(); }
    uint64_t data_bytes() const { return borrowed ? nbytes_total + kPktPad : data.size(); }
};

// OpusHead output gain (RFC 7845 5.1, Q7.8 dB) is applied unless
// CUOPUS_OUTPUT_GAIN=0 (the reference's decode path ignores it).
bool apply_output_gain()
{
    const char *v = getenv("CUOPUS_OUTPUT_GAIN");
    return !(v && v[0] == '0');
}

// Demuxes running at once in this process: each takes an equal share of the
// host threads (files of a multi-file call are demuxed side by side).
std::atomic<int> g_demuxing{0};

int demux_file(const uint8_t *ogg, size_t n, FilePlan *fp, bool with_chains, Err &e)
{
    std::string msg;
    if (!ogg || !n) return e.set(CUOPUS_RC_ARG, "empty input");
    auto dm = std::make_shared<Demuxed>();
    cuo_host::DemuxOpts dopts = cuo_host::demux_opts_from_env();
    const int active = ++g_demuxing;
    dopts.threads = std::max(1, cuo_host::host_threads(dopts) / active);
    const bool ok = cuo_host::demux(ogg, n, dm.get(), &msg, dopts);
    --g_demuxing;
    if (!ok) return e.set(CUOPUS_RC_SOFT, "%s", msg.c_str());
    fp->d = dm;
    if (fp->d->channels != 1 && fp->d->channels != 2)
        return e.set(CUOPUS_RC_SOFT, "only mono and stereo streams are supported (channels=%d)", fp->d->channels);
    if (with_chains) {
        double seg_s = env_f("CUOPUS_SEG_S", 1.0);
        double preroll_ms = env_f("CUOPUS_PREROLL_MS", 250.0);
        if (!cuo_host::build_chains(*fp->d, seg_s, preroll_ms, &fp->chains, &msg))
            return e.set(CUOPUS_RC_SOFT, "%s", msg.c_str());
    }
    fp->keep = fp->d->pcm_samples_48k > 0 ? (uint64_t)fp->d->pcm_samples_48k : 0;
    fp->out16 = cuo_host::out16_len(fp->d->pcm_samples_48k);
    return 0;
}

// Decode cost of one packet (microseconds of one wavefront) from its TOC:
// the per-frame cost of its mode and frame duration times its frame count.
uint32_t packet_cost(const uint8_t *p, uint32_t len)
{
    if (len == 0) return kCostLost;
    const uint8_t toc = p[0];
    int frames = 1;
    if ((toc & 3) == 1 || (toc & 3) == 2) frames = 2;
    else if ((toc & 3) == 3) frames = len > 1 ? (p[1] & 0x3F) : 1;
    uint32_t per;
    if (toc & 0x80) {
        per = kCostCelt[(toc >> 3) & 3];
    } else if ((toc & 0x60) == 0x60) {
        per = kCostHybrid[(toc >> 3) & 1];
    } else {
        per = kCostSilk[(toc >> 3) & 3];
    }
    return per * (uint32_t)(frames > 0 ? frames : 1);
}

// Merge the demuxed files into one packet table and one chain table.
int merge(Plan &pl, Err &e, bool borrow = false)
{
    uint64_t nbytes = 0, npk = 0, nch = 0;
    for (auto &f : pl.files) {
        nbytes += f.d->data.size();
        npk += f.d->off.size();
        nch += f.chains.size();
    }
    if (nbytes + kPktPad >= (1ull << 32)) return e.set(CUOPUS_RC_ARG, "merged packet data exceeds 4 GiB; split the call");
    pl.borrowed = borrow;
    pl.npk_total = npk;
    pl.nbytes_total = nbytes;
    if (!borrow) {
        pl.data.assign(nbytes + kPktPad, 0);
        pl.off.resize(npk);
        pl.len.resize(npk);
    } else if (pl.files.size() > 1) {
        pl.off.resize(npk);
    }
    pl.chains.resize(nch);
    pl.cap.resize(nch);
    pl.cost.resize(nch);
    const bool gain_on = apply_output_gain();
    const int threads = cuo_host::host_threads(cuo_host::demux_opts_from_env());
    uint64_t db = 0, pb = 0, cb = 0, cat = 0, keep_sum = 0;
    for (auto &f : pl.files) {
        if (!borrow) {
            if (!f.d->data.empty()) memcpy(pl.data.data() + db, f.d->data.data(), f.d->data.size());
            for (size_t k = 0; k < f.d->off.size(); k++) {
                pl.off[pb + k] = f.d->off[k] + (uint32_t)db;
                pl.len[pb + k] = f.d->len[k];
            }
        } else if (pl.files.size() > 1) {
            for (size_t k = 0; k < f.d->off.size(); k++) pl.off[pb + k] = f.d->off[k] + (uint32_t)db;
        }
        f.chain0 = cb;
        f.cat_off = cat;
        // each chain's TOC capacity and cost over its packets, in parallel
        // chain ranges (chains overlap by their preroll, so this reads more
        // packets than the file has)
        const size_t nc = f.chains.size();
        std::vector<uint64_t> cap_s(nc), cost_s(nc);
        {
            const int nr = (int)std::min<size_t>((size_t)threads, std::max<size_t>(1, nc / 256));
            cuo_host::parallel_for(nr, [&](int r) {
                for (size_t i = nc * (size_t)r / nr; i < nc * (size_t)(r + 1) / nr; i++) {
                    const Chain &c = f.chains[i];
                    uint64_t s = 0, cost = 0;
                    if ((uint64_t)c.first + c.npkts <= f.d->off.size())
                        for (uint32_t k = 0; k < c.npkts; k++) {
                            const uint32_t p = c.first + k;
                            const uint8_t *pk = f.d->data.data() + f.d->off[p];
                            s += (uint64_t)cuo_host::pkt_samples48(pk, f.d->len[p]);
                            cost += packet_cost(pk, f.d->len[p]);
                        }
                    cap_s[i] = s;
                    cost_s[i] = cost;
                }
            });
        }
        uint64_t fk = 0;
        for (size_t i = 0; i < f.chains.size(); i++) {
            const Chain &c = f.chains[i];
            if ((uint64_t)c.first + c.npkts > f.d->off.size())
                return e.set(CUOPUS_RC_SOFT, "chain %zu out of packet range", i);
            const uint64_t s = cap_s[i], cost = cost_s[i];
            if (s < (uint64_t)c.trim + c.keep)
                return e.set(CUOPUS_RC_SOFT, "chain %zu capacity %llu < trim %u + keep %u", i,
                             (unsigned long long)s, c.trim, c.keep);
            // chains never cross a link: the link holding the first packet gives the gain
            int gain = 0;
            for (const cuo_host::Link &lk : f.d->links)
                if (c.first >= lk.first && c.first < lk.first + lk.npkts) gain = lk.gain_q8;
            ChainDesc &cd = pl.chains[cb + i];
            cd.first = c.first + (uint32_t)pb;
            cd.npkts = c.npkts;
            cd.trim = c.trim;
            cd.keep = c.keep;
            cd.dst = cat / (uint64_t)f.out_channels + fk;  // in sample frames (cat is a multiple of 8)
            cd.rec_end = 0;  // no per-packet records outside the gate
            cd.gain = (int16_t)(gain_on ? gain : 0);
            cd.channels = (uint8_t)f.out_channels;
            cd.pad = 0;
            pl.cap[cb + i] = s;
            pl.cost[cb + i] = (uint32_t)std::min<uint64_t>(cost, 0xFFFFFFFFu);
            fk += c.keep;
        }
        if (fk != f.keep) return e.set(CUOPUS_RC_SOFT, "internal: keep mismatch");
        // each file's PCM starts 16-byte aligned (the resampler's wide loads)
        cat += (fk * (uint64_t)f.out_channels + 7) & ~7ull;
        pl.stereo = pl.stereo || cuo_host::needs_stereo_build(*f.d);
        keep_sum += fk;
        db += f.d->data.size();
        pb += f.d->off.size();
        cb += f.chains.size();
    }
    pl.total_keep = keep_sum;
    pl.cat_len = cat;
    return 0;
}

int refuse_capture(hipStream_t st, Err &e)
{
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    hipError_t s = hipStreamIsCapturing(st, &cs);
    if (s == hipSuccess && cs != hipStreamCaptureStatusNone)
        return e.set(CUOPUS_RC_ARG, "refusing to decode during HIP stream capture");
    (void)hipGetLastError();
    return 0;
}

// Work order: most expensive chains first (longest processing time first)
// by the per-packet cost model, so the persistent kernel's last chains are
// short ones and its tail stays short.  cost_model 0 keeps file order.
void work_order(const Plan &pl, int cost_model, std::vector<uint32_t> *order)
{
    const size_t n = pl.chains.size();
    order->resize(n);
    for (size_t i = 0; i < n; i++) (*order)[i] = (uint32_t)i;
    if (cost_model == 0) return;
    std::vector<uint64_t> key(n);
    for (size_t i = 0; i < n; i++)
        key[i] = cost_model == 2 ? pl.cap[i] + 960ull * pl.chains[i].npkts : pl.cost[i];
    std::stable_sort(order->begin(), order->end(), [&](uint32_t a, uint32_t b) { return key[a] > key[b]; });
}

// ---------------------------------------------------------------- device buffers of one call
struct CallBufs {
    uint8_t *data = nullptr;
    uint32_t *off = nullptr, *len = nullptr;
    uint8_t *lost = nullptr;
    ChainDesc *chains = nullptr;
    uint32_t *order = nullptr;
    ChainResult *res = nullptr;
    int16_t *cat = nullptr;
    RsFile *files = nullptr;
    uint32_t *queue = nullptr;
    DecHeap *heap = nullptr;
};

struct CallNeed {
    size_t data = 0, npk = 0, lost = 0, nch = 0, cat = 0, nfiles = 0, nheaps = 0;
    size_t heap = sizeof(DecHeap);  // bytes of one wavefront's heap (the kernel build's)
};

int bufs_from_workspace(const CallNeed &n, CallBufs *b, Err &e)
{
    int rc;
    if ((rc = grow(W.data, n.data, e)) || (rc = grow(W.off, 4 * n.npk, e)) || (rc = grow(W.len, 4 * n.npk, e)) ||
        (rc = grow(W.chains, sizeof(ChainDesc) * n.nch, e)) || (rc = grow(W.order, 4 * n.nch, e)) ||
        (rc = grow(W.res, sizeof(ChainResult) * n.nch, e)) || (rc = grow(W.queue, 256, e)) ||
        (rc = grow(W.heap, n.heap * n.nheaps, e)))
        return rc;
    if (n.lost && (rc = grow(W.lost, n.lost, e))) return rc;
    if (n.cat && (rc = grow(W.cat, 2 * n.cat, e))) return rc;
    if (n.nfiles && (rc = grow(W.files, sizeof(RsFile) * n.nfiles, e))) return rc;
    b->data = (uint8_t *)W.data.p;
    b->off = (uint32_t *)W.off.p;
    b->len = (uint32_t *)W.len.p;
    b->lost = n.lost ? (uint8_t *)W.lost.p : nullptr;
    b->chains = (ChainDesc *)W.chains.p;
    b->order = (uint32_t *)W.order.p;
    b->res = (ChainResult *)W.res.p;
    b->cat = n.cat ? (int16_t *)W.cat.p : nullptr;
    b->files = n.nfiles ? (RsFile *)W.files.p : nullptr;
    b->queue = (uint32_t *)W.queue.p;
    b->heap = (DecHeap *)W.heap.p;
    return 0;
}

// ---- the allocation-free entry point's device memory
// A file-scope device array, one copy per GPU, live until the process ends.
// Host code never passes `scratch` itself to a kernel or a copy: its name is
// a host-side stand-in, not a device address.  scratch_base() looks up each
// device's address once (hipGetSymbolAddress with that device current) and
// every use goes through that pointer.
constexpr size_t kScratchBytes = 128ull << 20;
static __device__ __align__(256) uint8_t scratch[kScratchBytes];
constexpr int kMaxDevices = 64;
void *g_scratch_ptr[kMaxDevices] = {nullptr};

int scratch_base(uint8_t **out, Err &e)
{
    int dev = -1;
    HIPCHK(e, hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return e.set(CUOPUS_RC_ARG, "device %d out of range", dev);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        hipError_t s = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (s != hipSuccess || !p)
            return e.set(CUOPUS_RC_GPU, "hipGetSymbolAddress(scratch) on device %d: %s", dev, hipGetErrorString(s));
        g_scratch_ptr[dev] = p;
    }
    *out = (uint8_t *)g_scratch_ptr[dev];
    return 0;
}

// Bytes a call needs in the static scratch (each region 256-byte aligned).
size_t scratch_bytes(const CallNeed &n)
{
    auto a = [](size_t v) { return (v + 255) & ~(size_t)255; };
    return a(n.data) + 2 * a(4 * n.npk) + a(n.lost) + a(sizeof(ChainDesc) * n.nch) + a(4 * n.nch) +
           a(sizeof(ChainResult) * n.nch) + a(2 * n.cat) + a(sizeof(RsFile) * n.nfiles) + a(256) +
           a(n.heap * n.nheaps);
}

int bufs_from_scratch(const CallNeed &n, CallBufs *b, Err &e)
{
    const size_t need = scratch_bytes(n);
    if (need > kScratchBytes)
        return e.set(CUOPUS_RC_ARG, "call needs %zu bytes of device scratch, the allocation-free path has %zu", need,
                     kScratchBytes);
    uint8_t *base;
    int rc = scratch_base(&base, e);
    if (rc) return rc;
    size_t o = 0;
    auto take = [&](size_t bytes) {
        uint8_t *p = base + o;
        o += (bytes + 255) & ~(size_t)255;
        return p;
    };
    b->data = take(n.data);
    b->off = (uint32_t *)take(4 * n.npk);
    b->len = (uint32_t *)take(4 * n.npk);
    b->lost = n.lost ? take(n.lost) : nullptr;
    b->chains = (ChainDesc *)take(sizeof(ChainDesc) * n.nch);
    b->order = (uint32_t *)take(4 * n.nch);
    b->res = (ChainResult *)take(sizeof(ChainResult) * n.nch);
    b->cat = n.cat ? (int16_t *)take(2 * n.cat) : nullptr;
    b->files = n.nfiles ? (RsFile *)take(sizeof(RsFile) * n.nfiles) : nullptr;
    b->queue = (uint32_t *)take(256);
    b->heap = (DecHeap *)take(n.heap * n.nheaps);
    return 0;
}

// Resident wavefronts of the persistent decoder when the caller does not
// choose: CUOPUS_SLOTS, else the tuned value for this device, else every
// resident slot of the chip.
// Work order when the caller does not choose: the tuned one, else the cost model.
int default_order()
{
    const int o = tuned_defaults().order;
    return o >= 0 && o <= 2 ? o : 1;
}

uint32_t default_slots()
{
    unsigned s = env_u("CUOPUS_SLOTS", 0);
    if (s) return s;
    const TunedDefaults &t = tuned_defaults();
    return t.slots ? t.slots : kDefaultWaves;
}

struct RunOpts {
    int16_t *s16_out = nullptr;          // decode straight into this (else workspace)
    std::vector<float *> f16_out;        // per file resample targets (empty: no resample)
    int32_t *pkt_nsamp = nullptr;        // device, per packet (gate)
    uint32_t *pkt_rng = nullptr;
    bool persistent = true;              // persistent work-queue kernel (else one workgroup per chain)
    uint32_t n_slots = 0;                // persistent wavefronts (0: default_slots())
    bool noalloc = false;                // device memory from the static scratch only
};

// Host-phase timing of a drop-in call, printed to stderr when CUOPUS_TRACE
// is set (demux, plan, upload, device work, validation).
struct HostTrace {
    bool on;
    std::chrono::steady_clock::time_point last;
    HostTrace() : on(getenv("CUOPUS_TRACE") != nullptr), last(std::chrono::steady_clock::now()) {}
    void mark(const char *phase)
    {
        if (!on) return;
        const auto t = std::chrono::steady_clock::now();
        fprintf(stderr, "[cuopus] %-16s %9.3f ms\n", phase,
                std::chrono::duration<double, std::milli>(t - last).count());
        last = t;
    }
};

// Kernel selection of the drop-in entry points: the persistent work-queue
// kernel with min(chains, slots) wavefronts (its heap memory is O(slots));
// CUOPUS_KERNEL=plain forces one workgroup per chain.
bool use_persistent()
{
    const char *v = getenv("CUOPUS_KERNEL");
    return !(v && strcmp(v, "plain") == 0);
}

// Upload, decode, resample, read back and validate.  results receives the
// per-chain results (host).
int run_plan(Plan &pl, hipStream_t st, RunOpts &o, std::vector<ChainResult> &results, Err &e)
{
    const uint32_t nch = (uint32_t)pl.chains.size();
    const uint64_t npk = pl.n_packets();
    int rc;
    std::vector<RsFile> rf;
    uint64_t tiles = 0;
    const int rsv = resample_variant();
    if (!o.f16_out.empty()) {
        for (size_t f = 0; f < pl.files.size(); f++) {
            if (pl.files[f].out16 <= 0) continue;
            RsFile r;
            r.in_off = pl.files[f].cat_off;
            r.in_len = pl.files[f].keep;
            r.out_len = (uint64_t)pl.files[f].out16;
            r.out = o.f16_out[f];
            r.tile0 = tiles;
            tiles += rs_tiles(r.out_len, rsv);
            rf.push_back(r);
        }
    }
    uint32_t n_waves = std::max<uint32_t>(1, std::min<uint32_t>(o.n_slots ? o.n_slots : default_slots(), nch));
    CallNeed need;
    need.data = pl.data_bytes();
    need.npk = npk;
    need.lost = pl.lost.empty() ? 0 : pl.lost.size();
    need.nch = nch;
    need.cat = o.s16_out ? 0 : (pl.cat_len ? pl.cat_len : 1);
    need.nfiles = rf.size();
    need.nheaps = o.persistent ? n_waves : nch;
    need.heap = heap_bytes(pl.stereo);
    CallBufs b;
    if (o.noalloc) {
        // fewer resident wavefronts when their heaps would not fit (>= 64)
        while (need.nheaps > 64 && scratch_bytes(need) > kScratchBytes) need.nheaps = std::max<size_t>(64, need.nheaps / 2);
        n_waves = (uint32_t)need.nheaps;
        rc = bufs_from_scratch(need, &b, e);
    } else {
        rc = bufs_from_workspace(need, &b, e);
    }
    if (rc) return rc;
    int16_t *out = o.s16_out ? o.s16_out : b.cat;
    HostTrace tr;

    std::vector<uint32_t> order;
    work_order(pl, default_order(), &order);
    tr.mark("order");
    if (pl.borrowed) {
        static const uint8_t kZeros[kPktPad] = {};
        uint64_t db = 0, pb = 0;
        for (const FilePlan &f : pl.files) {
            const Demuxed &dm = *f.d;
            const uint64_t np = dm.off.size();
            if (dm.data.size()) HIPCHK(e, hipMemcpyAsync(b.data + db, dm.data.data(), dm.data.size(), hipMemcpyHostToDevice, st));
            if (np) {
                HIPCHK(e, hipMemcpyAsync(b.len + pb, dm.len.data(), 4 * np, hipMemcpyHostToDevice, st));
                if (pl.files.size() == 1)
                    HIPCHK(e, hipMemcpyAsync(b.off, dm.off.data(), 4 * np, hipMemcpyHostToDevice, st));
            }
            db += dm.data.size();
            pb += np;
        }
        HIPCHK(e, hipMemcpyAsync(b.data + db, kZeros, kPktPad, hipMemcpyHostToDevice, st));
        if (pl.files.size() > 1 && npk) HIPCHK(e, hipMemcpyAsync(b.off, pl.off.data(), 4 * npk, hipMemcpyHostToDevice, st));
    } else {
        HIPCHK(e, hipMemcpyAsync(b.data, pl.data.data(), pl.data.size(), hipMemcpyHostToDevice, st));
        if (npk) {
            HIPCHK(e, hipMemcpyAsync(b.off, pl.off.data(), 4 * npk, hipMemcpyHostToDevice, st));
            HIPCHK(e, hipMemcpyAsync(b.len, pl.len.data(), 4 * npk, hipMemcpyHostToDevice, st));
        }
    }
    HIPCHK(e, hipMemcpyAsync(b.chains, pl.chains.data(), sizeof(ChainDesc) * nch, hipMemcpyHostToDevice, st));
    HIPCHK(e, hipMemcpyAsync(b.order, order.data(), 4ull * nch, hipMemcpyHostToDevice, st));
    if (b.lost) HIPCHK(e, hipMemcpyAsync(b.lost, pl.lost.data(), pl.lost.size(), hipMemcpyHostToDevice, st));
    if (!rf.empty()) HIPCHK(e, hipMemcpyAsync(b.files, rf.data(), sizeof(RsFile) * rf.size(), hipMemcpyHostToDevice, st));
    tr.mark("upload issued");

    DecodeArgs a;
    memset(&a, 0, sizeof(a));
    a.data = b.data;
    a.off = b.off;
    a.len = b.len;
    a.lost = b.lost;
    a.chains = b.chains;
    a.order = b.order;
    a.n_chains = nch;
    a.out = out;
    a.res = b.res;
    a.heap = b.heap;
    a.pkt_nsamp = o.pkt_nsamp;
    a.pkt_rng = o.pkt_rng;
    a.queue = b.queue;
    if (o.persistent) {
        // the queue counter is zeroed by a kernel (device buffers keep the last
        // call's contents; no memset node if a caller captures the sequence)
        hipLaunchKernelGGL(k_reset_counter, dim3(1), dim3(64), 0, st, b.queue);
        HIPCHK(e, hipGetLastError());
        launch_decode(pl.stereo, true, n_waves, st, a);
    } else {
        launch_decode(pl.stereo, false, nch, st, a);
    }
    HIPCHK(e, hipGetLastError());
    // one resampler launch over every file's tiles
    if (!rf.empty()) {
        static RsTaps taps = make_taps();
        launch_resample(tiles, st, out, b.files, (int)rf.size(), taps, rsv);
        HIPCHK(e, hipGetLastError());
    }
    results.resize(nch);
    HIPCHK(e, hipMemcpyAsync(results.data(), b.res, sizeof(ChainResult) * nch, hipMemcpyDeviceToHost, st));
    tr.mark("launches issued");
    HIPCHK(e, hipStreamSynchronize(st));
    tr.mark("device done");
    for (uint32_t c = 0; c < nch; c++) {
        if (results[c].stop != 0) return e.set(CUOPUS_RC_SOFT, "chain %u: device decode error %d", c, results[c].stop);
        if (results[c].samples != pl.cap[c])
            return e.set(CUOPUS_RC_SOFT, "chain %u: decoded %u samples, expected %llu", c, results[c].samples,
                         (unsigned long long)pl.cap[c]);
    }
    tr.mark("validated");
    return 0;
}

// Merge already-demuxed files (pl.files), attach the loss mask and run.
int decode_plan(Plan &pl, const uint8_t *lost, size_t mask_n, RunOpts &o, hipStream_t st, Err &e)
{
    int rc = refuse_capture(st, e);
    if (rc) return rc;
    HostTrace tr;
    if ((rc = merge(pl, e, true))) return rc;
    tr.mark("merge");
    if (lost) {
        if (mask_n != pl.n_packets())
            return e.set(CUOPUS_RC_ARG, "loss mask has %zu entries, stream has %llu packets", mask_n,
                         (unsigned long long)pl.n_packets());
        pl.lost.assign(lost, lost + mask_n);
    }
    if (pl.chains.empty()) return 0;
    std::vector<ChainResult> res;
    return run_plan(pl, st, o, res, e);
}

int last_ovf = 0;

}  // namespace

// =====================================================================================
// Demuxed files: demux once, size the output, then decode (the Python
// binding's path, so a file is parsed once per call).
// =====================================================================================
struct cuopus_file {
    FilePlan fp;
};

extern "C" int cuopus_file_open(const uint8_t *ogg, size_t n, cuopus_file **out, int64_t *info6, char *errbuf,
                                size_t errcap)
{
    Err e(errbuf, errcap);
    if (!out) return e.set(CUOPUS_RC_ARG, "out is NULL");
    *out = nullptr;
    HostTrace tr;
    cuopus_file *f = new cuopus_file();
    int rc = demux_file(ogg, n, &f->fp, true, e);
    if (rc) { delete f; return rc; }
    tr.mark("demux + chains");
    if (info6) {
        info6[0] = f->fp.out16;
        info6[1] = (int64_t)f->fp.keep;
        info6[2] = f->fp.d->channels;
        info6[3] = (int64_t)f->fp.d->off.size();
        info6[4] = (int64_t)f->fp.chains.size();
        info6[5] = f->fp.d->links.empty() ? 0 : f->fp.d->links[0].gain_q8;
    }
    *out = f;
    return 0;
}

extern "C" void cuopus_file_close(cuopus_file *f) { delete f; }

namespace {

int files_f16k(cuopus_file *const *fs, int nfiles, const uint8_t *lost, size_t mask_n, float *const *out_devs,
               const size_t *out_caps, int64_t *out_lens, void *stream, bool noalloc, Err &e)
{
    last_ovf = 0;
    if (out_lens)
        for (int f = 0; f < nfiles; f++) out_lens[f] = 0;
    if (nfiles <= 0) return 0;
    if (!fs || !out_devs || !out_caps) return e.set(CUOPUS_RC_ARG, "NULL argument array");
    HostTrace tr;
    Plan pl;
    RunOpts o;
    o.persistent = use_persistent();
    o.noalloc = noalloc;
    pl.files.resize((size_t)nfiles);
    for (int f = 0; f < nfiles; f++) {
        if (!fs[f]) return e.set(CUOPUS_RC_ARG, "file %d is NULL", f);
        const FilePlan &fp = fs[f]->fp;
        if ((uint64_t)fp.out16 > out_caps[f])
            return nfiles > 1 ? e.set(CUOPUS_RC_ARG, "file %d: output capacity %zu < required %lld", f, out_caps[f],
                                      (long long)fp.out16)
                              : e.set(CUOPUS_RC_ARG, "output capacity %zu < required %lld samples", out_caps[f],
                                      (long long)fp.out16);
        if (fp.out16 > 0 && !out_devs[f])
            return nfiles > 1 ? e.set(CUOPUS_RC_ARG, "file %d: out_dev is NULL", f)
                              : e.set(CUOPUS_RC_ARG, "out_dev is NULL");
        pl.files[f] = fp;
        o.f16_out.push_back(out_devs[f]);
    }
    tr.mark("plan files");
    int rc = decode_plan(pl, lost, mask_n, o, (hipStream_t)stream, e);
    if (rc) return rc;
    if (out_lens)
        for (int f = 0; f < nfiles; f++) out_lens[f] = pl.files[f].out16;
    return 0;
}

// Demux files [0, nfiles) into fps[f] on up to the host threads (files side
// by side); rc of the first failing file in file order, its message in e.
int demux_files(const uint8_t *const *oggs, const size_t *ns, int nfiles, FilePlan *const *fps, Err &e)
{
    std::vector<int> rcs((size_t)nfiles, 0);
    std::vector<std::array<char, 256>> msgs((size_t)nfiles);
    std::atomic<int> next{0};
    const int nt = std::min(nfiles, cuo_host::host_threads(cuo_host::demux_opts_from_env()));
    cuo_host::parallel_for(nt, [&](int) {
        for (int f; (f = next++) < nfiles;) {
            Err fe(msgs[f].data(), msgs[f].size());
            rcs[f] = demux_file(oggs[f], ns[f], fps[f], true, fe);
        }
    });
    for (int f = 0; f < nfiles; f++)
        if (rcs[f]) return nfiles > 1 ? e.set(rcs[f], "file %d: %s", f, msgs[f].data()) : e.set(rcs[f], "%s", msgs[f].data());
    return 0;
}

int files_f16k_from_ogg(const uint8_t *const *oggs, const size_t *ns, int nfiles, const uint8_t *lost,
                        size_t mask_n, float *const *out_devs, const size_t *out_caps, int64_t *out_lens,
                        void *stream, bool noalloc, Err &e)
{
    if (out_lens)
        for (int f = 0; f < nfiles; f++) out_lens[f] = 0;
    if (nfiles <= 0) return 0;
    if (!oggs || !ns) return e.set(CUOPUS_RC_ARG, "NULL argument array");
    std::vector<cuopus_file> files((size_t)nfiles);
    std::vector<cuopus_file *> fs((size_t)nfiles);
    std::vector<FilePlan *> fps((size_t)nfiles);
    for (int f = 0; f < nfiles; f++) {
        fs[f] = &files[f];
        fps[f] = &files[f].fp;
    }
    if (int rc = demux_files(oggs, ns, nfiles, fps.data(), e)) return rc;
    return files_f16k(fs.data(), nfiles, lost, mask_n, out_devs, out_caps, out_lens, stream, noalloc, e);
}

}  // namespace

extern "C" int cuopus_file_decode_f16k_on(cuopus_file *f, float *out_dev, size_t out_cap, int64_t *out_len,
                                          void *stream, char *errbuf, size_t errcap)
{
    Err e(errbuf, errcap);
    if (!f) return e.set(CUOPUS_RC_ARG, "file is NULL");
    return files_f16k(&f, 1, nullptr, 0, &out_dev, &out_cap, out_len, stream, false, e);
}

extern "C" int cuopus_files_decode_f16k_multi(cuopus_file *const *fs, int nfiles, float *const *out_devs,
                                              const size_t *out_caps, int64_t *out_lens, void *stream,
                                              char *errbuf, size_t errcap)
{
    Err e(errbuf, errcap);
    return files_f16k(fs, nfiles, nullptr, 0, out_devs, out_caps, out_lens, stream, false, e);
}

// =====================================================================================
extern "C" int cuopus_last_arena_overflow(void) { return last_ovf; }

extern "C" long long cuopus_workspace_bytes(void) { return g_dev_bytes; }

static void release_mel_tables();

extern "C" void cuopus_release_workspaces(void)
{
    release_mel_tables();
    DevBuf *all[] = {&W.data, &W.off, &W.len, &W.lost, &W.chains, &W.order, &W.heap,
                     &W.cat, &W.res, &W.nsamp, &W.rng, &W.files, &W.queue};
    for (DevBuf *b : all) {
        if (b->p) {
            hipFree(b->p);
            g_dev_bytes -= (long long)b->cap;
        }
        b->p = nullptr;
        b->cap = 0;
    }
}

extern "C" int cuopus_probe(const uint8_t *ogg, size_t n, int64_t *samples16k, char *errbuf, size_t errcap)
{
    Err e(errbuf, errcap);
    if (!samples16k) return e.set(CUOPUS_RC_ARG, "samples16k is NULL");
    FilePlan fp;
    int rc = demux_file(ogg, n, &fp, false, e);
    if (rc) return rc;
    *samples16k = fp.out16;
    return 0;
}

extern "C" int cuopus_probe_packets(const uint8_t *ogg, size_t n, int64_t *n_packets, int64_t *n_chains, char *errbuf,
                                    size_t errcap)
{
    Err e(errbuf, errcap);
    FilePlan fp;
    int rc = demux_file(ogg, n, &fp, true, e);
    if (rc) return rc;
    if (n_packets) *n_packets = (int64_t)fp.d->off.size();
    if (n_chains) *n_chains = (int64_t)fp.chains.size();
    return 0;
}

extern "C" int cuopus_dump_chains(const uint8_t *ogg, size_t n, uint32_t *out4, size_t cap_chains, int64_t *n_chains,
                                  char *errbuf, size_t errcap)
{
    Err e(errbuf, errcap);
    FilePlan fp;
    int rc = demux_file(ogg, n, &fp, true, e);
    if (rc) return rc;
    if (n_chains) *n_chains = (int64_t)fp.chains.size();
    if (!out4) return 0;
    if (cap_chains < fp.chains.size())
        return e.set(CUOPUS_RC_ARG, "chain capacity %zu < %zu", cap_chains, fp.chains.size());
    for (size_t i = 0; i < fp.chains.size(); i++) {
        out4[4 * i + 0] = fp.chains[i].first;
        out4[4 * i + 1] = fp.chains[i].npkts;
        out4[4 * i + 2] = fp.chains[i].trim;
        out4[4 * i + 3] = fp.chains[i].keep;
    }
    return 0;
}

extern "C" int cuopus_dump_packets(const uint8_t *ogg, size_t n, uint8_t *data, size_t data_cap, uint32_t *lens,
                                   size_t cap_packets, int64_t *n_packets, int64_t *n_bytes, char *errbuf,
                                   size_t errcap)
{
    Err e(errbuf, errcap);
    FilePlan fp;
    int rc = demux_file(ogg, n, &fp, false, e);
    if (rc) return rc;
    const size_t np = fp.d->off.size();
    const size_t nb = np ? (size_t)fp.d->off[np - 1] + fp.d->len[np - 1] : 0;
    if (n_packets) *n_packets = (int64_t)np;
    if (n_bytes) *n_bytes = (int64_t)nb;
    if (lens) {
        if (cap_packets < np) return e.set(CUOPUS_RC_ARG, "packet capacity %zu < %zu", cap_packets, np);
        memcpy(lens, fp.d->len.data(), np * sizeof(uint32_t));
    }
    if (data) {
        if (data_cap < nb) return e.set(CUOPUS_RC_ARG, "byte capacity %zu < %zu", data_cap, nb);
        memcpy(data, fp.d->data.data(), nb);
    }
    return 0;
}

extern "C" int cuopus_demux_stats(const uint8_t *ogg, size_t n, int64_t *s10, char *errbuf, size_t errcap)
{
    Err e(errbuf, errcap);
    if (!s10) return e.set(CUOPUS_RC_ARG, "null output");
    FilePlan fp;
    int rc = demux_file(ogg, n, &fp, false, e);
    if (rc) return rc;
    const cuo_host::DemuxStats &st = fp.d->stats;
    s10[0] = st.pages;
    s10[1] = (int64_t)fp.d->links.size();
    s10[2] = st.crc_bad;
    s10[3] = st.foreign_pages;
    s10[4] = st.dropped_packets;
    s10[5] = st.granule_warnings;
    s10[6] = fp.d->granule_checked;
    s10[7] = (int64_t)st.skipped_bytes;
    s10[8] = (int64_t)fp.d->off.size();
    s10[9] = fp.d->pcm_samples_48k;
    return 0;
}

extern "C" int cuopus_decode_f16k_on(const uint8_t *ogg, size_t n, float *out_dev, size_t out_cap, int64_t *out_len,
                                     void *stream, char *errbuf, size_t errcap)
{
    Err e(errbuf, errcap);
    return files_f16k_from_ogg(&ogg, &n, 1, nullptr, 0, &out_dev, &out_cap, out_len, stream, false, e);
}

extern "C" int cuopus_decode_f16k(const uint8_t *ogg, size_t n, float *out_dev, size_t out_cap, int64_t *out_len,
                                  char *errbuf, size_t errcap)
{
    return cuopus_decode_f16k_on(ogg, n, out_dev, out_cap, out_len, nullptr, errbuf, errcap);
}

extern "C" int cuopus_decode_f16k_noalloc_on(const uint8_t *ogg, size_t n, float *out_dev, size_t out_cap,
                                             int64_t *out_len, void *stream, char *errbuf, size_t errcap)
{
    Err e(errbuf, errcap);
    return files_f16k_from_ogg(&ogg, &n, 1, nullptr, 0, &out_dev, &out_cap, out_len, stream, true, e);
}

extern "C" int cuopus_decode_lossy_f16k_on(const uint8_t *ogg, size_t n, const uint8_t *lost, size_t mask_n,
                                           float *out_dev, size_t out_cap, int64_t *out_len, void *stream,
                                           char *errbuf, size_t errcap)
{
    Err e(errbuf, errcap);
    if (!lost) return e.set(CUOPUS_RC_ARG, "lost mask is NULL");
    return files_f16k_from_ogg(&ogg, &n, 1, lost, mask_n, &out_dev, &out_cap, out_len, stream, false, e);
}

extern "C" int cuopus_decode_s16_on(const uint8_t *ogg, size_t n, int16_t *out_dev, size_t out_cap, int64_t *out_len,
                                    void *stream, char *errbuf, size_t errcap)
{
    Err e(errbuf, errcap);
    last_ovf = 0;
    if (out_len) *out_len = 0;
    Plan pl;
    pl.files.resize(1);
    int rc = demux_file(ogg, n, &pl.files[0], true, e);
    if (rc) return rc;
    FilePlan &fp = pl.files[0];
    fp.out_channels = fp.d->channels;  // the stream's own channels, interleaved
    const uint64_t keep = fp.keep, elems = keep * (uint64_t)fp.out_channels;
    if (elems > out_cap)
        return e.set(CUOPUS_RC_ARG, "output capacity %zu < required %llu samples", out_cap, (unsigned long long)elems);
    if (elems > 0 && !out_dev) return e.set(CUOPUS_RC_ARG, "out_dev is NULL");
    RunOpts o;
    o.s16_out = out_dev;
    o.persistent = use_persistent();
    rc = decode_plan(pl, nullptr, 0, o, (hipStream_t)stream, e);
    if (rc) return rc;
    if (out_len) *out_len = (int64_t)keep;
    return 0;
}

extern "C" int cuopus_decode_f16k_multi(const uint8_t *const *oggs, const size_t *ns, int nfiles,
                                        float *const *out_devs, const size_t *out_caps, int64_t *out_lens,
                                        void *stream, char *errbuf, size_t errcap)
{
    Err e(errbuf, errcap);
    return files_f16k_from_ogg(oggs, ns, nfiles, nullptr, 0, out_devs, out_caps, out_lens, stream, false, e);
}

namespace {
int gate_chains(const uint8_t *data, size_t nbytes, const uint32_t *off, const uint32_t *len, uint32_t n_packets,
                const uint8_t *lost, const uint32_t *first, const uint32_t *npkts, const uint32_t *trim,
                const uint32_t *keep, uint32_t n_chains, int channels, bool stereo_build, int16_t *pcm_out,
                int32_t *pkt_nsamp, uint32_t *pkt_rng, int32_t *chain_stop, void *stream, Err &e)
{
    if (channels != 1 && channels != 2) return e.set(CUOPUS_RC_ARG, "channels must be 1 or 2");
    hipStream_t st = (hipStream_t)stream;
    int rc = refuse_capture(st, e);
    if (rc) return rc;
    if (!data || !off || !len || !first || !npkts || !trim || !keep || !pcm_out || !pkt_nsamp || !pkt_rng ||
        !chain_stop)
        return e.set(CUOPUS_RC_ARG, "NULL argument");
    if (nbytes + kPktPad >= (1ull << 32)) return e.set(CUOPUS_RC_ARG, "packet data exceeds 4 GiB");
    Plan pl;
    pl.data.assign(nbytes + kPktPad, 0);
    if (nbytes) memcpy(pl.data.data(), data, nbytes);
    pl.off.assign(off, off + n_packets);
    pl.len.assign(len, len + n_packets);
    for (uint32_t p = 0; p < n_packets; p++)
        if ((uint64_t)off[p] + len[p] > nbytes) return e.set(CUOPUS_RC_ARG, "packet %u out of range", p);
    if (lost) pl.lost.assign(lost, lost + n_packets);
    pl.chains.resize(n_chains);
    pl.cap.resize(n_chains);
    pl.cost.resize(n_chains);
    uint64_t cat = 0;
    for (uint32_t c = 0; c < n_chains; c++) {
        if ((uint64_t)first[c] + npkts[c] > n_packets) return e.set(CUOPUS_RC_ARG, "chain %u out of range", c);
        ChainDesc &cd = pl.chains[c];
        cd.first = first[c];
        cd.npkts = npkts[c];
        cd.trim = trim[c];
        cd.keep = keep[c];
        cd.dst = cat;
        cd.rec_end = c + 1 < n_chains ? first[c + 1] : 0xFFFFFFFFu;
        cd.gain = 0;
        cd.channels = (uint8_t)channels;
        cd.pad = 0;
        uint64_t s = 0, cost = 0;
        for (uint32_t k = 0; k < npkts[c]; k++) {
            const uint32_t p = first[c] + k;
            s += (uint64_t)cuo_host::pkt_samples48(data + off[p], len[p]);
            cost += packet_cost(data + off[p], len[p]);
        }
        pl.cap[c] = s;
        pl.cost[c] = (uint32_t)std::min<uint64_t>(cost, 0xFFFFFFFFu);
        cat += keep[c];
    }
    pl.total_keep = cat;
    pl.cat_len = cat * (uint64_t)channels;
    pl.stereo = stereo_build || channels == 2;
    if (n_chains == 0) return 0;
    if ((rc = grow(W.nsamp, 4ull * n_packets, e))) return rc;
    if ((rc = grow(W.rng, 4ull * n_packets, e))) return rc;
    HIPCHK(e, hipMemsetAsync(W.nsamp.p, 0, 4ull * n_packets, st));
    HIPCHK(e, hipMemsetAsync(W.rng.p, 0, 4ull * n_packets, st));
    RunOpts o;
    o.pkt_nsamp = (int32_t *)W.nsamp.p;
    o.pkt_rng = (uint32_t *)W.rng.p;
    o.persistent = use_persistent();
    const unsigned slots = env_u("CUOPUS_GATE_SLOTS", 0);
    if (slots) o.n_slots = slots;
    std::vector<ChainResult> res;
    rc = run_plan(pl, st, o, res, e);
    // results are copied back even when a chain failed validation
    if (rc && rc != CUOPUS_RC_SOFT) return rc;
    int rc2 = rc;
    for (uint32_t c = 0; c < n_chains && c < res.size(); c++) chain_stop[c] = res[c].stop;
    HIPCHK(e, hipMemcpyAsync(pkt_nsamp, W.nsamp.p, 4ull * n_packets, hipMemcpyDeviceToHost, st));
    HIPCHK(e, hipMemcpyAsync(pkt_rng, W.rng.p, 4ull * n_packets, hipMemcpyDeviceToHost, st));
    HIPCHK(e, hipMemcpyAsync(pcm_out, W.cat.p, 2 * pl.cat_len, hipMemcpyDeviceToHost, st));
    HIPCHK(e, hipStreamSynchronize(st));
    return rc2;
}
}  // namespace

extern "C" int cuopus_gate_chains(const uint8_t *data, size_t nbytes, const uint32_t *off, const uint32_t *len,
                                  uint32_t n_packets, const uint8_t *lost, const uint32_t *first,
                                  const uint32_t *npkts, const uint32_t *trim, const uint32_t *keep,
                                  uint32_t n_chains, int16_t *pcm_out, int32_t *pkt_nsamp, uint32_t *pkt_rng,
                                  int32_t *chain_stop, void *stream, char *errbuf, size_t errcap)
{
    Err e(errbuf, errcap);
    return gate_chains(data, nbytes, off, len, n_packets, lost, first, npkts, trim, keep, n_chains, 1, false,
                       pcm_out, pkt_nsamp, pkt_rng, chain_stop, stream, e);
}

extern "C" int cuopus_gate_chains_ch(const uint8_t *data, size_t nbytes, const uint32_t *off, const uint32_t *len,
                                     uint32_t n_packets, const uint8_t *lost, const uint32_t *first,
                                     const uint32_t *npkts, const uint32_t *trim, const uint32_t *keep,
                                     uint32_t n_chains, int channels, int stereo_build, int16_t *pcm_out,
                                     int32_t *pkt_nsamp, uint32_t *pkt_rng, int32_t *chain_stop, void *stream,
                                     char *errbuf, size_t errcap)
{
    Err e(errbuf, errcap);
    return gate_chains(data, nbytes, off, len, n_packets, lost, first, npkts, trim, keep, n_chains, channels,
                       stereo_build != 0, pcm_out, pkt_nsamp, pkt_rng, chain_stop, stream, e);
}

extern "C" int cuopus_resample_s16_to_f16k(const int16_t *in_dev, int64_t n48, float *out_dev, size_t out_cap,
                                           int64_t *out_len, void *stream, char *errbuf, size_t errcap)
{
    Err e(errbuf, errcap);
    hipStream_t st = (hipStream_t)stream;
    int64_t n16 = cuo_host::out16_len(n48);
    if (out_len) *out_len = 0;
    if ((uint64_t)n16 > out_cap) return e.set(CUOPUS_RC_ARG, "output capacity %zu < required %lld", out_cap, (long long)n16);
    if (n16 == 0) return 0;
    if (!in_dev || !out_dev) return e.set(CUOPUS_RC_ARG, "NULL pointer");
    RsFile r;
    r.in_off = 0;
    r.in_len = (uint64_t)n48;
    r.out_len = (uint64_t)n16;
    r.out = out_dev;
    r.tile0 = 0;
    int rc;
    if ((rc = grow(W.files, sizeof(RsFile), e))) return rc;
    HIPCHK(e, hipMemcpyAsync(W.files.p, &r, sizeof(RsFile), hipMemcpyHostToDevice, st));
    static RsTaps taps = make_taps();
    const int rsv = resample_variant();
    launch_resample(rs_tiles((uint64_t)n16, rsv), st, in_dev, (const RsFile *)W.files.p, 1, taps, rsv);
    HIPCHK(e, hipGetLastError());
    HIPCHK(e, hipStreamSynchronize(st));
    if (out_len) *out_len = n16;
    return 0;
}


// =====================================================================================
// Log-mel front-end over a resident f32 16 kHz buffer
// =====================================================================================
namespace {

// Slaney mel scale (linear below 1 kHz, logarithmic above; 15 mel at 1 kHz).
double hz_to_mel(double f)
{
    const double f_sp = 200.0 / 3.0, min_log_hz = 1000.0, min_log_mel = min_log_hz / f_sp;
    const double logstep = log(6.4) / 27.0;
    return f < min_log_hz ? f / f_sp : min_log_mel + log(f / min_log_hz) / logstep;
}
double mel_to_hz(double m)
{
    const double f_sp = 200.0 / 3.0, min_log_hz = 1000.0, min_log_mel = min_log_hz / f_sp;
    const double logstep = log(6.4) / 27.0;
    return m < min_log_mel ? f_sp * m : min_log_hz * exp(logstep * (m - min_log_mel));
}

// Window, twiddles and the n_mels triangular filters (0-8 kHz at 16 kHz,
// Slaney area normalisation 2 / (f_hi - f_lo)), in double, stored as float.
bool mel_tables(int n_mels, MelTables *t)
{
    memset(t, 0, sizeof(*t));
    for (int j = 0; j < kMelWin; j++) t->window[j] = (float)(0.5 - 0.5 * cos(2.0 * M_PI * j / kMelWin));
    for (int e = 0; e < 256; e++) {
        t->tw256[2 * e] = (float)cos(2.0 * M_PI * e / 256.0);
        t->tw256[2 * e + 1] = (float)-sin(2.0 * M_PI * e / 256.0);
    }
    for (int k = 0; k < kMelBins; k++) {
        t->tw512[2 * k] = (float)cos(2.0 * M_PI * k / 512.0);
        t->tw512[2 * k + 1] = (float)-sin(2.0 * M_PI * k / 512.0);
    }
    const double sr = 16000.0, fmax = sr / 2;
    std::vector<double> edge((size_t)n_mels + 2);
    const double m0 = hz_to_mel(0.0), m1 = hz_to_mel(fmax);
    for (int i = 0; i < n_mels + 2; i++) edge[i] = mel_to_hz(m0 + (m1 - m0) * i / (n_mels + 1));
    int nw = 0;
    for (int m = 0; m < n_mels; m++) {
        const double lo = edge[m], mid = edge[m + 1], hi = edge[m + 2];
        const double enorm = 2.0 / (hi - lo);
        int first = -1, count = 0;
        t->filt.woff[m] = nw;
        for (int k = 0; k < kMelBins; k++) {
            const double f = k * sr / kMelFft;
            const double w = std::max(0.0, std::min((f - lo) / (mid - lo), (hi - f) / (hi - mid)));
            if (w <= 0.0) {
                if (first >= 0) break;
                continue;
            }
            if (first < 0) first = k;
            if (nw >= kMelMaxWeights) return false;
            t->filt.weight[nw++] = (float)(w * enorm);
            count++;
        }
        t->filt.first[m] = first < 0 ? 0 : first;
        t->filt.count[m] = count;
    }
    return true;
}

struct MelCache {
    DevBuf tab;
    int n_mels = 0;
};
MelCache g_mel[kMaxDevices];

}  // namespace

static void release_mel_tables()
{
    for (MelCache &m : g_mel) {
        if (m.tab.p) {
            hipFree(m.tab.p);
            g_dev_bytes -= (long long)m.tab.cap;
        }
        m.tab.p = nullptr;
        m.tab.cap = 0;
        m.n_mels = 0;
    }
}

extern "C" int cuopus_mel_filters(int n_mels, float *out, size_t cap)
{
    if (n_mels < 1 || n_mels > kMelMaxMels || !out || cap < (size_t)n_mels * kMelBins) return CUOPUS_RC_ARG;
    MelTables t;
    if (!mel_tables(n_mels, &t)) return CUOPUS_RC_ARG;
    for (size_t i = 0; i < (size_t)n_mels * kMelBins; i++) out[i] = 0.f;
    for (int m = 0; m < n_mels; m++)
        for (int i = 0; i < t.filt.count[m]; i++) out[(size_t)m * kMelBins + t.filt.first[m] + i] = t.filt.weight[t.filt.woff[m] + i];
    return 0;
}

extern "C" int cuopus_tuned_defaults(int *slots, int *rs_variant, int *order)
{
    if (slots) *slots = (int)default_slots();
    if (rs_variant) *rs_variant = resample_variant();
    if (order) *order = default_order();
    return 0;
}

extern "C" int cuopus_log_mel(const float *in_dev, int64_t n16, float *out_dev, size_t out_cap, int64_t *frames,
                              int n_mels, void *stream, char *errbuf, size_t errcap)
{
    Err e(errbuf, errcap);
    if (frames) *frames = 0;
    if (n_mels < 1 || n_mels > kMelMaxMels) return e.set(CUOPUS_RC_ARG, "n_mels %d outside 1..%d", n_mels, kMelMaxMels);
    if (n16 < 0) return e.set(CUOPUS_RC_ARG, "negative input length");
    const int64_t nf = n16 >= kMelWin ? 1 + (n16 - kMelWin) / kMelHop : 0;
    if ((uint64_t)(nf * n_mels) > out_cap)
        return e.set(CUOPUS_RC_ARG, "output capacity %zu < %lld frames x %d mels", out_cap, (long long)nf, n_mels);
    if (nf == 0) return 0;
    if (!in_dev || !out_dev) return e.set(CUOPUS_RC_ARG, "NULL pointer");
    hipStream_t st = (hipStream_t)stream;
    int dev = 0;
    HIPCHK(e, hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return e.set(CUOPUS_RC_ARG, "device %d out of range", dev);
    MelCache &mc = g_mel[dev];
    if (mc.n_mels != n_mels) {
        MelTables t;
        if (!mel_tables(n_mels, &t)) return e.set(CUOPUS_RC_ARG, "mel filter table overflow");
        int rc = grow(mc.tab, sizeof(MelTables), e);
        if (rc) return rc;
        HIPCHK(e, hipMemcpy(mc.tab.p, &t, sizeof(MelTables), hipMemcpyHostToDevice));
        mc.n_mels = n_mels;
    }
    const uint64_t blocks = ((uint64_t)nf + kMelFramesPerBlock - 1) / kMelFramesPerBlock;
    hipLaunchKernelGGL(k_log_mel, dim3((unsigned)blocks), dim3(64 * kMelWaves), 0, st, in_dev, nf,
                       (const MelTables *)mc.tab.p, n_mels, out_dev);
    HIPCHK(e, hipGetLastError());
    HIPCHK(e, hipStreamSynchronize(st));
    if (frames) *frames = nf;
    return 0;
}

// =====================================================================================
// Resident batches
// =====================================================================================
struct cuopus_batch {
    Plan pl;
    cuopus_batch_opts opts;
    DevBuf data, off, len, lost, chains, order, cat, res, files, queue, heap;
    std::vector<RsFile> rsf;          // one entry per file with output
    std::vector<uint64_t> file_tiles; // per rsf entry: tile count
    std::vector<uint32_t> file_chain0, file_nchain;
    std::vector<int> file_rs;         // per file: its rsf entry or -1
    uint64_t tiles = 0;
    uint32_t n_slots = 0;
    uint64_t n_heaps = 0;  // HBM decoder heaps: one per workgroup of the largest launch
    int64_t chain_pkts = 0, chain_bytes = 0;
    RsTaps taps;
    hipGraphExec_t gexec = nullptr;
    hipGraph_t graph = nullptr;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    bool ran = false;
    int rs_variant = 0;
};

namespace {

int batch_grow(cuopus_batch *b, DevBuf &d, size_t need, Err &e) { return grow(d, need, e); }

void batch_free(cuopus_batch *b)
{
    DevBuf *all[] = {&b->data, &b->off, &b->len, &b->lost, &b->chains, &b->order, &b->heap,
                     &b->cat, &b->res, &b->files, &b->queue};
    for (DevBuf *d : all) {
        if (d->p) { hipFree(d->p); g_dev_bytes -= (long long)d->cap; }
        d->p = nullptr;
        d->cap = 0;
    }
    if (b->gexec) hipGraphExecDestroy(b->gexec);
    if (b->graph) hipGraphDestroy(b->graph);
    for (auto &ev : b->ev)
        if (ev) hipEventDestroy(ev);
}

// Enqueue the launch sequence of a batch (no host work, capturable).
int batch_enqueue(cuopus_batch *b, hipStream_t st, Err &e)
{
    Plan &pl = b->pl;
    DecodeArgs a;
    memset(&a, 0, sizeof(a));
    a.data = (const uint8_t *)b->data.p;
    a.off = (const uint32_t *)b->off.p;
    a.len = (const uint32_t *)b->len.p;
    a.lost = pl.lost.empty() ? nullptr : (const uint8_t *)b->lost.p;
    a.chains = (const ChainDesc *)b->chains.p;
    a.out = (int16_t *)b->cat.p;
    a.res = (ChainResult *)b->res.p;
    a.queue = (uint32_t *)b->queue.p;
    a.heap = (DecHeap *)b->heap.p;
    const bool prof = b->opts.profile && !b->opts.use_graph;
    const int mode = b->opts.mode;
    if (prof) HIPCHK(e, hipEventRecord(b->ev[0], st));
    if (mode == 0) {
        // plain per-file path: one decode launch and one resample launch per file
        for (size_t f = 0; f < pl.files.size(); f++) {
            uint32_t c0 = b->file_chain0[f], nc = b->file_nchain[f];
            if (nc) {
                a.order = (const uint32_t *)b->order.p + c0;
                a.n_chains = nc;
                launch_decode(pl.stereo, false, nc, st, a);
                HIPCHK(e, hipGetLastError());
            }
            const int r = b->file_rs[f];
            if (r >= 0) {
                launch_resample(b->file_tiles[r], st, (const int16_t *)b->cat.p, (const RsFile *)b->files.p + r, 1,
                                b->taps, b->rs_variant);
                HIPCHK(e, hipGetLastError());
            }
        }
        if (prof) {
            HIPCHK(e, hipEventRecord(b->ev[1], st));
            HIPCHK(e, hipEventRecord(b->ev[2], st));
            HIPCHK(e, hipEventRecord(b->ev[3], st));
        }
        return 0;
    }
    const uint32_t nch = (uint32_t)pl.chains.size();
    a.order = (const uint32_t *)b->order.p;
    if (mode == 2) {
        hipLaunchKernelGGL(k_reset_counter, dim3(1), dim3(64), 0, st, (uint32_t *)b->queue.p);
        HIPCHK(e, hipGetLastError());
        a.n_chains = nch;
        launch_decode(pl.stereo, true, b->n_slots, st, a);
        HIPCHK(e, hipGetLastError());
    } else {
        a.n_chains = nch;
        launch_decode(pl.stereo, false, nch, st, a);
        HIPCHK(e, hipGetLastError());
    }
    if (prof) HIPCHK(e, hipEventRecord(b->ev[1], st));
    if (b->tiles) {
        if (prof) HIPCHK(e, hipEventRecord(b->ev[2], st));
        launch_resample(b->tiles, st, (const int16_t *)b->cat.p, (const RsFile *)b->files.p, (int)b->rsf.size(),
                        b->taps, b->rs_variant);
        HIPCHK(e, hipGetLastError());
        if (prof) HIPCHK(e, hipEventRecord(b->ev[3], st));
    } else if (prof) {
        HIPCHK(e, hipEventRecord(b->ev[2], st));
        HIPCHK(e, hipEventRecord(b->ev[3], st));
    }
    return 0;
}

}  // namespace

extern "C" int cuopus_batch_create(const uint8_t *const *oggs, const size_t *ns, int nfiles,
                                   const uint8_t *const *lost_masks, const size_t *mask_ns, float *const *out_devs,
                                   const size_t *out_caps, const cuopus_batch_opts *opts, cuopus_batch **out,
                                   char *errbuf, size_t errcap)
{
    Err e(errbuf, errcap);
    if (!out) return e.set(CUOPUS_RC_ARG, "out is NULL");
    *out = nullptr;
    if (nfiles <= 0 || !oggs || !ns) return e.set(CUOPUS_RC_ARG, "no files");
    cuopus_batch *b = new cuopus_batch();
    if (opts) b->opts = *opts;
    else memset(&b->opts, 0, sizeof(b->opts));
    Plan &pl = b->pl;
    pl.files.resize((size_t)nfiles);
    std::vector<FilePlan *> fps((size_t)nfiles);
    for (int f = 0; f < nfiles; f++) fps[f] = &pl.files[f];
    int rc = demux_files(oggs, ns, nfiles, fps.data(), e);
    for (int f = 0; f < nfiles && !rc; f++) {
        if (out_devs && out_caps && (uint64_t)pl.files[f].out16 > out_caps[f])
            rc = e.set(CUOPUS_RC_ARG, "file %d: output capacity %zu < %lld", f, out_caps[f],
                       (long long)pl.files[f].out16);
        else if (out_devs && pl.files[f].out16 > 0 && !out_devs[f])
            rc = e.set(CUOPUS_RC_ARG, "file %d: out_dev is NULL", f);
    }
    if (!rc) rc = merge(pl, e);
    if (!rc && lost_masks) {
        pl.lost.assign(pl.off.size(), 0);
        uint64_t pb = 0;
        for (int f = 0; f < nfiles && !rc; f++) {
            size_t np = pl.files[f].d->off.size();
            if (lost_masks[f]) {
                if (!mask_ns || mask_ns[f] != np)
                    rc = e.set(CUOPUS_RC_ARG, "file %d: loss mask size != %zu packets", f, np);
                else memcpy(pl.lost.data() + pb, lost_masks[f], np);
            }
            pb += np;
        }
    }
    if (!rc && pl.chains.empty()) rc = e.set(CUOPUS_RC_SOFT, "no chains");
    if (rc) { delete b; return rc; }
    // per-file chain ranges and the work order
    std::vector<uint32_t> order;
    b->file_chain0.resize((size_t)nfiles);
    b->file_nchain.resize((size_t)nfiles);
    for (int f = 0; f < nfiles; f++) {
        b->file_chain0[f] = (uint32_t)pl.files[f].chain0;
        b->file_nchain[f] = (uint32_t)pl.files[f].chains.size();
    }
    work_order(pl, b->opts.mode == 0 ? 0 : (b->opts.lpt_order == 3 ? default_order() : b->opts.lpt_order), &order);
    const uint32_t nch = (uint32_t)pl.chains.size();
    // persistent grid: n_slots wavefronts (default: the chip's resident
    // capacity at the kernel's LDS footprint), never more than the chains
    b->n_slots = b->opts.n_slots > 0 ? (uint32_t)b->opts.n_slots : default_slots();
    if (b->n_slots > nch) b->n_slots = nch;
    if (b->opts.mode == 2) {
        b->n_heaps = b->n_slots;
    } else if (b->opts.mode == 1) {
        b->n_heaps = nch;
    } else {
        b->n_heaps = 1;
        for (int f = 0; f < nfiles; f++) b->n_heaps = std::max<uint64_t>(b->n_heaps, b->file_nchain[f]);
    }
    b->taps = make_taps();
    b->rs_variant = b->opts.rs_variant > 0 && b->opts.rs_variant <= kRsVariants ? b->opts.rs_variant - 1
                                                                               : resample_variant();
    b->file_rs.assign(pl.files.size(), -1);
    for (size_t f = 0; f < pl.files.size(); f++) {
        if (!out_devs || pl.files[f].out16 <= 0) continue;
        b->file_rs[f] = (int)b->rsf.size();
        RsFile r;
        r.in_off = pl.files[f].cat_off;
        r.in_len = pl.files[f].keep;
        r.out_len = (uint64_t)pl.files[f].out16;
        r.out = out_devs[f];
        r.tile0 = b->tiles;
        uint64_t t = rs_tiles(r.out_len, b->rs_variant);
        b->file_tiles.push_back(t);
        b->tiles += t;
        b->rsf.push_back(r);
    }
    // in mode 0 each file's resample launch indexes its own entry, so tile0 = 0
    std::vector<RsFile> rsf_upload = b->rsf;
    if (b->opts.mode == 0)
        for (auto &r : rsf_upload) r.tile0 = 0;
    const uint64_t npk = pl.off.size();
    for (const ChainDesc &c : pl.chains) {
        b->chain_pkts += c.npkts;
        for (uint32_t k = 0; k < c.npkts; k++) b->chain_bytes += pl.len[c.first + k];
    }
    if ((rc = grow(b->data, pl.data.size() + 4, e)) || (rc = grow(b->off, 4 * npk, e)) ||
        (rc = grow(b->len, 4 * npk, e)) || (rc = grow(b->chains, sizeof(ChainDesc) * nch, e)) ||
        (rc = grow(b->order, 4ull * nch, e)) || (rc = grow(b->res, sizeof(ChainResult) * nch, e)) ||
        (rc = grow(b->cat, 2 * (pl.cat_len ? pl.cat_len : 1), e)) || (rc = grow(b->queue, 256, e)) ||
        (rc = grow(b->heap, heap_bytes(pl.stereo) * (size_t)b->n_heaps, e)) ||
        (rc = grow(b->files, sizeof(RsFile) * (rsf_upload.size() ? rsf_upload.size() : 1), e)) ||
        (!pl.lost.empty() && (rc = grow(b->lost, pl.lost.size(), e)))) {
        batch_free(b);
        delete b;
        return rc;
    }
    hipError_t hs = hipSuccess;
    if (hs == hipSuccess) hs = hipMemcpy(b->data.p, pl.data.data(), pl.data.size(), hipMemcpyHostToDevice);
    if (hs == hipSuccess) hs = hipMemcpy(b->off.p, pl.off.data(), 4 * npk, hipMemcpyHostToDevice);
    if (hs == hipSuccess) hs = hipMemcpy(b->len.p, pl.len.data(), 4 * npk, hipMemcpyHostToDevice);
    if (hs == hipSuccess) hs = hipMemcpy(b->chains.p, pl.chains.data(), sizeof(ChainDesc) * nch, hipMemcpyHostToDevice);
    if (hs == hipSuccess) hs = hipMemcpy(b->order.p, order.data(), 4ull * nch, hipMemcpyHostToDevice);
    if (hs == hipSuccess && !rsf_upload.empty())
        hs = hipMemcpy(b->files.p, rsf_upload.data(), sizeof(RsFile) * rsf_upload.size(), hipMemcpyHostToDevice);
    if (hs == hipSuccess && !pl.lost.empty())
        hs = hipMemcpy(b->lost.p, pl.lost.data(), pl.lost.size(), hipMemcpyHostToDevice);
    for (int i = 0; i < 4 && hs == hipSuccess; i++) hs = hipEventCreate(&b->ev[i]);
    if (hs != hipSuccess) {
        batch_free(b);
        delete b;
        return e.set(CUOPUS_RC_GPU, "batch upload: %s", hipGetErrorString(hs));
    }
    // packet bytes are kept on the device; release the host copies
    std::vector<uint8_t>().swap(pl.data);
    *out = b;
    return 0;
}

extern "C" int cuopus_batch_run(cuopus_batch *b, void *stream, char *errbuf, size_t errcap)
{
    Err e(errbuf, errcap);
    if (!b) return e.set(CUOPUS_RC_ARG, "batch is NULL");
    hipStream_t st = (hipStream_t)stream;
    int rc;
    if (!b->opts.use_graph) {
        if ((rc = refuse_capture(st, e))) return rc;
        rc = batch_enqueue(b, st, e);
        b->ran = rc == 0;
        return rc;
    }
    if (!b->gexec) {
        // capture on a private stream so the caller's stream is never in
        // capture mode, then replay on the caller's stream
        hipStream_t cs;
        HIPCHK(e, hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
        hipError_t s1 = hipStreamBeginCapture(cs, hipStreamCaptureModeRelaxed);
        if (s1 != hipSuccess) { hipStreamDestroy(cs); return e.set(CUOPUS_RC_GPU, "begin capture: %s", hipGetErrorString(s1)); }
        rc = batch_enqueue(b, cs, e);
        hipError_t s2 = hipStreamEndCapture(cs, &b->graph);
        hipStreamDestroy(cs);
        if (rc) return rc;
        if (s2 != hipSuccess) return e.set(CUOPUS_RC_GPU, "end capture: %s", hipGetErrorString(s2));
        HIPCHK(e, hipGraphInstantiate(&b->gexec, b->graph, nullptr, nullptr, 0));
    }
    HIPCHK(e, hipGraphLaunch(b->gexec, st));
    b->ran = true;
    return 0;
}

extern "C" int cuopus_batch_check(cuopus_batch *b, void *stream, int64_t *out_lens, char *errbuf, size_t errcap)
{
    Err e(errbuf, errcap);
    if (!b) return e.set(CUOPUS_RC_ARG, "batch is NULL");
    if (!b->ran) return e.set(CUOPUS_RC_ARG, "batch has not run");
    hipStream_t st = (hipStream_t)stream;
    const size_t nch = b->pl.chains.size();
    std::vector<ChainResult> res(nch);
    HIPCHK(e, hipMemcpyAsync(res.data(), b->res.p, sizeof(ChainResult) * nch, hipMemcpyDeviceToHost, st));
    HIPCHK(e, hipStreamSynchronize(st));
    for (size_t c = 0; c < nch; c++) {
        if (res[c].stop != 0) return e.set(CUOPUS_RC_SOFT, "chain %zu: device decode error %d", c, res[c].stop);
        if (res[c].samples != b->pl.cap[c])
            return e.set(CUOPUS_RC_SOFT, "chain %zu: decoded %u samples, expected %llu", c, res[c].samples,
                         (unsigned long long)b->pl.cap[c]);
    }
    if (out_lens)
        for (size_t f = 0; f < b->pl.files.size(); f++) out_lens[f] = b->pl.files[f].out16;
    return 0;
}

extern "C" int cuopus_batch_info(const cuopus_batch *b, int64_t *c)
{
    if (!b || !c) return CUOPUS_RC_ARG;
    int64_t dec = 0, out16 = 0, bytes = 0;
    for (uint64_t v : b->pl.cap) dec += (int64_t)v;
    for (auto &f : b->pl.files) { out16 += f.out16; bytes += (int64_t)f.d->data.size(); }
    c[0] = (int64_t)b->pl.chains.size();
    c[1] = (int64_t)b->pl.off.size();
    c[2] = (int64_t)b->pl.total_keep;
    c[3] = dec;
    c[4] = out16;
    c[5] = bytes;
    c[6] = b->chain_pkts;
    c[7] = b->chain_bytes;
    return 0;
}

extern "C" int cuopus_batch_chain_times(cuopus_batch *b, void *stream, uint32_t *t0, uint32_t *t1, int64_t cap)
{
    if (!b || !b->ran || cap < (int64_t)b->pl.chains.size()) return CUOPUS_RC_ARG;
    const size_t nch = b->pl.chains.size();
    std::vector<ChainResult> res(nch);
    hipStream_t st = (hipStream_t)stream;
    if (hipMemcpyAsync(res.data(), b->res.p, sizeof(ChainResult) * nch, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
        return CUOPUS_RC_GPU;
    for (size_t c = 0; c < nch; c++) {
        if (t0) t0[c] = res[c].t0;
        if (t1) t1[c] = res[c].t1;
    }
    return 0;
}

extern "C" int cuopus_batch_kernel_ms(const cuopus_batch *b, float *decode_ms, float *resample_ms)
{
    if (!b || !b->opts.profile || b->opts.use_graph) return CUOPUS_RC_ARG;
    float d = 0, r = 0;
    if (hipEventElapsedTime(&d, b->ev[0], b->ev[1]) != hipSuccess) return CUOPUS_RC_GPU;
    if (hipEventElapsedTime(&r, b->ev[2], b->ev[3]) != hipSuccess) return CUOPUS_RC_GPU;
    if (decode_ms) *decode_ms = d;
    if (resample_ms) *resample_ms = r;
    return 0;
}

extern "C" void cuopus_batch_destroy(cuopus_batch *b)
{
    if (!b) return;
    batch_free(b);
    delete b;
}


// =====================================================================================
// Serving: continuous batching of decode requests
// =====================================================================================
// Requests are demuxed on submit and queue in arrival order.  poll() retires
// finished launches (event query, no blocking), then admits queued requests
// into a new launch on a free launch slot while the device memory the server
// holds stays within the budget.  Several launches run at once on their own
// streams, so requests that arrive while others decode start as soon as a
// slot and memory are free and their work overlaps on the device.  Each
// launch is the batch path: one persistent decode kernel over the admitted
// requests' chains (costliest first) and one resample launch over their files.
//
// Memory: each launch slot keeps its device buffers (packets, tables, int16
// PCM, decoder heaps, ...) for reuse; the budget bounds the sum of every
// slot's buffers, busy or idle.  Admission computes what a launch would add
// to its slot's buffers; when that does not fit, idle slots' buffers are
// released first.
//
// Ordering: submit records an event on the caller's stream; the launch waits
// for it before it writes the request's output, so work the caller queued on
// that memory earlier (the allocator reusing a block, a fill) finishes first.
struct SrvReq {
    uint64_t id = 0;
    FilePlan fp;
    std::vector<uint8_t> lost;
    float *out = nullptr;
    int64_t out16 = 0;
    hipEvent_t ready = nullptr;   // recorded on the caller's stream at submit
    int rc = 0;
    char err[160] = {0};
    ~SrvReq()
    {
        if (ready) hipEventDestroy(ready);
    }
};

// Device bytes of one launch, per buffer (what srv_launch grows).
struct SrvNeed {
    uint64_t data = kPktPad, npk = 0, nch = 0, cat = 0, files = 0, lost = 0;
    uint32_t waves = 0;
    bool stereo = false;
    void add(const SrvReq *r)
    {
        data += r->fp.d->data.size();
        npk += r->fp.d->off.size();
        nch += r->fp.chains.size();
        cat += (r->fp.keep + 7) & ~7ull;
        files += r->out16 > 0;
        lost = lost || !r->lost.empty();
        stereo = stereo || cuo_host::needs_stereo_build(*r->fp.d);
    }
};

struct SrvLaunch {
    hipStream_t st = nullptr;
    hipEvent_t done = nullptr;
    bool busy = false;
    std::vector<SrvReq *> reqs;
    std::vector<uint64_t> cap;         // per chain expected samples
    DevBuf data, off, len, lost, chains, order, cat, res, files, queue, heap;
    uint8_t *h_stage = nullptr;        // pinned staging for uploads
    size_t h_stage_cap = 0;
    ChainResult *h_res = nullptr;      // pinned results
    size_t h_res_cap = 0;
    // the device buffers in srv_extra's order
    std::array<DevBuf *, 11> bufs() { return {&data, &off, &len, &lost, &chains, &order, &cat, &res, &files, &queue, &heap}; }
    std::array<const DevBuf *, 11> bufs() const
    {
        return {&data, &off, &len, &lost, &chains, &order, &cat, &res, &files, &queue, &heap};
    }
    uint64_t held() const
    {
        uint64_t s = 0;
        for (const DevBuf *b : bufs()) s += b->cap;
        return s;
    }
    void release()
    {
        for (DevBuf *b : bufs()) {
            if (b->p) {
                hipFree(b->p);
                g_dev_bytes -= (long long)b->cap;
            }
            b->p = nullptr;
            b->cap = 0;
        }
    }
};

struct cuopus_server {
    cuopus_server_opts opts;
    std::deque<SrvReq *> pending;
    std::vector<SrvLaunch> launches;
    std::deque<SrvReq *> finished;
    uint64_t next_id = 1;
    RsTaps taps;
    uint32_t waves = 0;
    int rs_variant = 0;
    uint64_t held() const
    {
        uint64_t s = 0;
        for (const SrvLaunch &L : launches) s += L.held();
        return s;
    }
};

namespace {

int pinned_grow(uint8_t **p, size_t *cap, size_t need, Err &e)
{
    if (*cap >= need) return 0;
    if (*p) hipHostFree(*p);
    *p = nullptr;
    *cap = 0;
    size_t n = need + need / 4 + 4096;
    hipError_t s = hipHostMalloc((void **)p, n, hipHostMallocDefault);
    if (s != hipSuccess) return e.set(CUOPUS_RC_GPU, "hipHostMalloc(%zu): %s", n, hipGetErrorString(s));
    *cap = n;
    return 0;
}

// What a launch of `n` adds to slot L's device buffers (grow() replaces a
// buffer by one of exactly the needed size when it is too small).
uint64_t srv_extra(const cuopus_server *srv, const SrvLaunch &L, const SrvNeed &n)
{
    const uint64_t waves = std::max<uint64_t>(1, std::min<uint64_t>(srv->waves, n.nch));
    const uint64_t need[11] = {n.data, 4 * n.npk, 4 * n.npk, n.lost ? n.npk : 0, sizeof(ChainDesc) * n.nch, 4 * n.nch,
                               2 * std::max<uint64_t>(1, n.cat), sizeof(ChainResult) * n.nch,
                               sizeof(RsFile) * std::max<uint64_t>(1, n.files), 256, heap_bytes(n.stereo) * waves};
    const auto have = L.bufs();
    uint64_t extra = 0;
    for (int i = 0; i < 11; i++)
        if (need[i] > have[i]->cap) extra += need[i];
    return extra;
}

// Build and enqueue one launch for the requests in L.reqs (no host waits).
int srv_launch(cuopus_server *srv, SrvLaunch &L, Err &e)
{
    Plan pl;
    pl.files.reserve(L.reqs.size());
    bool any_lost = false;
    for (SrvReq *r : L.reqs) {
        pl.files.push_back(r->fp);  // copy: the request keeps its demux for error reports
        any_lost = any_lost || !r->lost.empty();
    }
    int rc = merge(pl, e);
    if (rc) return rc;
    const uint32_t nch = (uint32_t)pl.chains.size();
    const uint64_t npk = pl.off.size();
    if (any_lost) {
        pl.lost.assign(npk, 0);
        uint64_t pb = 0;
        for (size_t f = 0; f < L.reqs.size(); f++) {
            if (!L.reqs[f]->lost.empty()) memcpy(pl.lost.data() + pb, L.reqs[f]->lost.data(), L.reqs[f]->lost.size());
            pb += pl.files[f].d->off.size();
        }
    }
    std::vector<uint32_t> order;
    work_order(pl, default_order(), &order);
    std::vector<RsFile> rsf;
    uint64_t tiles = 0;
    for (size_t f = 0; f < pl.files.size(); f++) {
        if (L.reqs[f]->out16 <= 0) continue;
        RsFile r;
        r.in_off = pl.files[f].cat_off;
        r.in_len = pl.files[f].keep;
        r.out_len = (uint64_t)L.reqs[f]->out16;
        r.out = L.reqs[f]->out;
        r.tile0 = tiles;
        tiles += rs_tiles(r.out_len, srv->rs_variant);
        rsf.push_back(r);
    }
    const uint32_t waves = std::max<uint32_t>(1, std::min<uint32_t>(srv->waves, nch));
    // device buffers (kept per launch slot; sizes as srv_extra computes them)
    if ((rc = grow(L.data, pl.data.size(), e)) || (rc = grow(L.off, 4 * npk, e)) ||
        (rc = grow(L.len, 4 * npk, e)) || (rc = grow(L.chains, sizeof(ChainDesc) * nch, e)) ||
        (rc = grow(L.order, 4ull * nch, e)) || (rc = grow(L.res, sizeof(ChainResult) * nch, e)) ||
        (rc = grow(L.cat, 2 * (pl.cat_len ? pl.cat_len : 1), e)) || (rc = grow(L.queue, 256, e)) ||
        (rc = grow(L.files, sizeof(RsFile) * (rsf.empty() ? 1 : rsf.size()), e)) ||
        (rc = grow(L.heap, heap_bytes(pl.stereo) * (size_t)waves, e)) ||
        (any_lost && (rc = grow(L.lost, npk, e))))
        return rc;
    // stage everything in one pinned block, then async copies on the launch stream
    const size_t o_data = 0, o_off = o_data + pl.data.size(), o_len = o_off + 4 * npk, o_ch = o_len + 4 * npk,
                 o_ord = o_ch + sizeof(ChainDesc) * nch, o_rs = o_ord + 4ull * nch,
                 o_lost = o_rs + sizeof(RsFile) * rsf.size(), total = o_lost + (any_lost ? npk : 0);
    if ((rc = pinned_grow(&L.h_stage, &L.h_stage_cap, total, e))) return rc;
    memcpy(L.h_stage + o_data, pl.data.data(), pl.data.size());
    memcpy(L.h_stage + o_off, pl.off.data(), 4 * npk);
    memcpy(L.h_stage + o_len, pl.len.data(), 4 * npk);
    memcpy(L.h_stage + o_ch, pl.chains.data(), sizeof(ChainDesc) * nch);
    memcpy(L.h_stage + o_ord, order.data(), 4ull * nch);
    if (!rsf.empty()) memcpy(L.h_stage + o_rs, rsf.data(), sizeof(RsFile) * rsf.size());
    if (any_lost) memcpy(L.h_stage + o_lost, pl.lost.data(), npk);
    size_t rneed = sizeof(ChainResult) * nch;
    {
        uint8_t *hr = (uint8_t *)L.h_res;
        if ((rc = pinned_grow(&hr, &L.h_res_cap, rneed, e))) return rc;
        L.h_res = (ChainResult *)hr;
    }
    hipStream_t st = L.st;
    // the caller's stream reached each request's submit point
    for (SrvReq *r : L.reqs)
        if (r->ready) HIPCHK(e, hipStreamWaitEvent(st, r->ready, 0));
    HIPCHK(e, hipMemcpyAsync(L.data.p, L.h_stage + o_data, pl.data.size(), hipMemcpyHostToDevice, st));
    HIPCHK(e, hipMemcpyAsync(L.off.p, L.h_stage + o_off, 4 * npk, hipMemcpyHostToDevice, st));
    HIPCHK(e, hipMemcpyAsync(L.len.p, L.h_stage + o_len, 4 * npk, hipMemcpyHostToDevice, st));
    HIPCHK(e, hipMemcpyAsync(L.chains.p, L.h_stage + o_ch, sizeof(ChainDesc) * nch, hipMemcpyHostToDevice, st));
    HIPCHK(e, hipMemcpyAsync(L.order.p, L.h_stage + o_ord, 4ull * nch, hipMemcpyHostToDevice, st));
    if (!rsf.empty())
        HIPCHK(e, hipMemcpyAsync(L.files.p, L.h_stage + o_rs, sizeof(RsFile) * rsf.size(), hipMemcpyHostToDevice, st));
    if (any_lost) HIPCHK(e, hipMemcpyAsync(L.lost.p, L.h_stage + o_lost, npk, hipMemcpyHostToDevice, st));
    DecodeArgs a;
    memset(&a, 0, sizeof(a));
    a.data = (const uint8_t *)L.data.p;
    a.off = (const uint32_t *)L.off.p;
    a.len = (const uint32_t *)L.len.p;
    a.lost = any_lost ? (const uint8_t *)L.lost.p : nullptr;
    a.chains = (const ChainDesc *)L.chains.p;
    a.order = (const uint32_t *)L.order.p;
    a.n_chains = nch;
    a.heap = (DecHeap *)L.heap.p;
    a.out = (int16_t *)L.cat.p;
    a.res = (ChainResult *)L.res.p;
    a.queue = (uint32_t *)L.queue.p;
    hipLaunchKernelGGL(k_reset_counter, dim3(1), dim3(64), 0, st, a.queue);
    HIPCHK(e, hipGetLastError());
    launch_decode(pl.stereo, true, waves, st, a);
    HIPCHK(e, hipGetLastError());
    if (tiles) {
        launch_resample(tiles, st, (const int16_t *)L.cat.p, (const RsFile *)L.files.p, (int)rsf.size(), srv->taps,
                        srv->rs_variant);
        HIPCHK(e, hipGetLastError());
    }
    HIPCHK(e, hipMemcpyAsync(L.h_res, L.res.p, rneed, hipMemcpyDeviceToHost, st));
    HIPCHK(e, hipEventRecord(L.done, st));
    L.cap = pl.cap;
    L.busy = true;
    return 0;
}

// Retire a finished launch: validate each chain, report each request.
void srv_retire(cuopus_server *srv, SrvLaunch &L)
{
    size_t c0 = 0;
    for (SrvReq *r : L.reqs) {
        size_t nc = r->fp.chains.size();
        for (size_t c = c0; c < c0 + nc && !r->rc; c++) {
            if (L.h_res[c].stop != 0) {
                r->rc = CUOPUS_RC_SOFT;
                snprintf(r->err, sizeof(r->err), "chain %zu: device decode error %d", c - c0, L.h_res[c].stop);
            } else if (L.h_res[c].samples != L.cap[c]) {
                r->rc = CUOPUS_RC_SOFT;
                snprintf(r->err, sizeof(r->err), "chain %zu: decoded %u samples, expected %llu", c - c0,
                         L.h_res[c].samples, (unsigned long long)L.cap[c]);
            }
        }
        c0 += nc;
        srv->finished.push_back(r);
    }
    L.reqs.clear();
    L.busy = false;
}

}  // namespace

extern "C" int cuopus_server_create(const cuopus_server_opts *opts, cuopus_server **out, char *errbuf,
                                    size_t errcap)
{
    Err e(errbuf, errcap);
    if (!out) return e.set(CUOPUS_RC_ARG, "out is NULL");
    *out = nullptr;
    cuopus_server *s = new cuopus_server();
    memset(&s->opts, 0, sizeof(s->opts));
    if (opts) s->opts = *opts;
    if (s->opts.n_streams <= 0) s->opts.n_streams = 4;
    if (s->opts.mem_budget <= 0) s->opts.mem_budget = 8ll << 30;
    if (s->opts.max_chains_per_launch <= 0) s->opts.max_chains_per_launch = 1 << 20;
    s->taps = make_taps();
    s->waves = s->opts.n_waves > 0 ? (uint32_t)s->opts.n_waves : default_slots();
    s->rs_variant = resample_variant();
    s->launches.resize((size_t)s->opts.n_streams);
    for (auto &L : s->launches) {
        hipError_t a = hipStreamCreateWithFlags(&L.st, hipStreamNonBlocking);
        hipError_t b = a == hipSuccess ? hipEventCreateWithFlags(&L.done, hipEventDisableTiming) : a;
        if (b != hipSuccess) {
            cuopus_server_destroy(s);
            return e.set(CUOPUS_RC_GPU, "stream/event: %s", hipGetErrorString(b));
        }
    }
    *out = s;
    return 0;
}

extern "C" int cuopus_server_submit(cuopus_server *srv, const uint8_t *ogg, size_t n, const uint8_t *lost,
                                    size_t mask_n, float *out_dev, size_t out_cap, void *stream, uint64_t *req_id,
                                    char *errbuf, size_t errcap)
{
    Err e(errbuf, errcap);
    if (!srv || !req_id) return e.set(CUOPUS_RC_ARG, "NULL argument");
    SrvReq *r = new SrvReq();
    int rc = demux_file(ogg, n, &r->fp, true, e);
    if (!rc && lost) {
        if (mask_n != r->fp.d->off.size()) rc = e.set(CUOPUS_RC_ARG, "loss mask size != %zu packets", r->fp.d->off.size());
        else r->lost.assign(lost, lost + mask_n);
    }
    if (!rc && r->fp.out16 > 0 && (!out_dev || (uint64_t)r->fp.out16 > out_cap))
        rc = e.set(CUOPUS_RC_ARG, "output capacity %zu < %lld", out_cap, (long long)r->fp.out16);
    if (!rc && r->fp.chains.empty()) rc = e.set(CUOPUS_RC_SOFT, "no chains");
    if (rc) { delete r; return rc; }
    r->out = out_dev;
    r->out16 = r->fp.out16;
    SrvNeed one;
    one.add(r);
    SrvLaunch empty;
    const uint64_t alone = srv_extra(srv, empty, one);
    if (alone > (uint64_t)srv->opts.mem_budget) {
        delete r;
        return e.set(CUOPUS_RC_ARG, "request needs %llu device bytes, budget is %lld", (unsigned long long)alone,
                     (long long)srv->opts.mem_budget);
    }
    hipError_t s = hipEventCreateWithFlags(&r->ready, hipEventDisableTiming);
    if (s == hipSuccess) s = hipEventRecord(r->ready, (hipStream_t)stream);
    if (s != hipSuccess) {
        delete r;
        return e.set(CUOPUS_RC_GPU, "submit event: %s", hipGetErrorString(s));
    }
    r->id = srv->next_id++;
    *req_id = r->id;
    srv->pending.push_back(r);
    return 0;
}

extern "C" int cuopus_server_poll(cuopus_server *srv, uint64_t *ids, int *rcs, int cap, int *n_done, char *errbuf,
                                  size_t errcap)
{
    Err e(errbuf, errcap);
    if (!srv || !n_done) return e.set(CUOPUS_RC_ARG, "NULL argument");
    // 1. retire finished launches
    for (auto &L : srv->launches) {
        if (!L.busy) continue;
        hipError_t q = hipEventQuery(L.done);
        if (q == hipErrorNotReady) continue;
        if (q != hipSuccess) return e.set(CUOPUS_RC_GPU, "launch: %s", hipGetErrorString(q));
        srv_retire(srv, L);
    }
    // 2. admit queued requests onto free launch slots while the device bytes
    //    the server holds stay within the budget
    const uint64_t budget = (uint64_t)srv->opts.mem_budget;
    for (auto &L : srv->launches) {
        if (L.busy || srv->pending.empty()) continue;
        SrvNeed need;
        std::vector<SrvReq *> take;
        bool released = false;
        while (!srv->pending.empty()) {
            SrvReq *r = srv->pending[take.size()];
            SrvNeed next = need;
            next.add(r);
            if (!take.empty() && next.nch > (uint64_t)srv->opts.max_chains_per_launch) break;
            uint64_t extra = srv_extra(srv, L, next);
            if (srv->held() + extra > budget && !released) {
                // give back what idle slots keep, then try again
                for (auto &M : srv->launches)
                    if (!M.busy && &M != &L) M.release();
                released = true;
                extra = srv_extra(srv, L, next);
            }
            if (srv->held() + extra > budget && L.held() > 0 && take.empty()) {
                // this slot's own buffers are the wrong shape: drop them
                L.release();
                extra = srv_extra(srv, L, next);
            }
            if (srv->held() + extra > budget) break;
            need = next;
            take.push_back(r);
            if (take.size() == srv->pending.size()) break;
        }
        if (take.empty()) break;  // head request must wait for memory
        for (size_t i = 0; i < take.size(); i++) srv->pending.pop_front();
        L.reqs = take;
        int rc = srv_launch(srv, L, e);
        if (rc) {
            for (SrvReq *r : L.reqs) {
                r->rc = rc;
                snprintf(r->err, sizeof(r->err), "%s", errbuf ? errbuf : "launch failed");
                srv->finished.push_back(r);
            }
            L.reqs.clear();
            if (rc == CUOPUS_RC_GPU) return rc;
        }
    }
    // 3. report finished requests
    int k = 0;
    while (k < cap && !srv->finished.empty()) {
        SrvReq *r = srv->finished.front();
        srv->finished.pop_front();
        if (ids) ids[k] = r->id;
        if (rcs) rcs[k] = r->rc;
        if (r->rc) e.set(r->rc, "request %llu: %s", (unsigned long long)r->id, r->err);
        delete r;
        k++;
    }
    *n_done = k;
    return 0;
}

extern "C" int cuopus_server_stats(const cuopus_server *srv, int64_t *s4)
{
    if (!srv || !s4) return CUOPUS_RC_ARG;
    int64_t busy = 0;
    for (auto &L : srv->launches) busy += L.busy;
    s4[0] = (int64_t)srv->pending.size();
    s4[1] = busy;
    s4[2] = (int64_t)srv->held();
    s4[3] = (int64_t)srv->finished.size();
    return 0;
}

extern "C" void cuopus_server_destroy(cuopus_server *srv)
{
    if (!srv) return;
    for (auto &L : srv->launches) {
        if (L.st) hipStreamSynchronize(L.st);
        L.release();
        if (L.h_stage) hipHostFree(L.h_stage);
        if (L.h_res) hipHostFree(L.h_res);
        for (SrvReq *r : L.reqs) delete r;
        if (L.done) hipEventDestroy(L.done);
        if (L.st) hipStreamDestroy(L.st);
    }
    for (SrvReq *r : srv->pending) delete r;
    for (SrvReq *r : srv->finished) delete r;
    delete srv;
}

// Cycle accounting of a -DCUO_PROF build: copies and clears the per-scope
// cycle sums (returns the number of scopes, 0 in a normal build).
extern "C" int cuopus_prof_read(unsigned long long *out, int cap)
{
#if defined(CUO_PROF)
    unsigned long long h[kProfSlots];
    if (hipMemcpyFromSymbol(h, HIP_SYMBOL(g_prof), sizeof(h)) != hipSuccess) return -1;
    unsigned long long z[kProfSlots] = {0};
    if (hipMemcpyToSymbol(HIP_SYMBOL(g_prof), z, sizeof(z)) != hipSuccess) return -1;
    for (int i = 0; i < cap && i < kProfCount; i++) out[i] = h[i];
    return kProfCount;
#else
    (void)out;
    (void)cap;
    return 0;
#endif
}

extern "C" int cuopus_warmup(char *errbuf, size_t errcap)
{
    Err e(errbuf, errcap);
    int rc;
    if ((rc = grow(W.res, sizeof(ChainResult), e))) return rc;
    DecodeArgs a;
    memset(&a, 0, sizeof(a));
    a.n_chains = 0;  // every workgroup returns at once: only loads the code objects
    a.res = (ChainResult *)W.res.p;
    if ((rc = grow(W.heap, sizeof(DecHeap), e))) return rc;
    a.heap = (DecHeap *)W.heap.p;
    launch_chains(1, 0, a);
    HIPCHK(e, hipGetLastError());
    HIPCHK(e, hipDeviceSynchronize());
    // the allocation-free path's device address on this device
    uint8_t *base;
    return scratch_base(&base, e);
}
