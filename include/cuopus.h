/* cuopus.h -- C ABI of the MI355X batch Opus decoder (libcuopus.so, ROCm/HIP).
 *
 * Drop-in for the reference's libcuopus (libcuopus.cu:1-74 contract, :201-776
 * definitions): same symbol names, signatures, return codes and threading
 * contract.  `void *stream` is a hipStream_t (torch.cuda.current_stream().
 * cuda_stream on PyTorch-ROCm); device pointers are HIP device pointers.
 *
 * Return codes (libcuopus.cu:48-51):
 *   0 ok; 1 soft failure (bad stream / decode validation, GPU healthy);
 *   2 GPU (HIP) error; 3 caller error (capacity, NULL pointer, capture).
 * errbuf, when non-NULL, is always NUL-terminated.
 *
 * Threading / capture contract (libcuopus.cu:53-68): no context creation and
 * no hipSetDevice (runs on the caller's current device); no internal threads;
 * the single-file and multi-file calls issue all work on the caller's stream,
 * synchronise it before returning and refuse to run while that stream is
 * capturing.  Workspaces are library-global and grow-only: not thread-safe.
 */
#ifndef CUOPUS_H
#define CUOPUS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CUOPUS_RC_OK 0
#define CUOPUS_RC_SOFT 1
#define CUOPUS_RC_GPU 2
#define CUOPUS_RC_ARG 3

/* ---------------------------------------------------------------------------
 * The reference's eight entry points.
 * ------------------------------------------------------------------------- */

/* Demux + validate (mono, granule cross-check); *samples16k = exact f32 16 kHz
 * output length, max(0, (playable48 - 48) / 3).  Host only, no GPU work.
 * Replaces libcuopus.cu:223-236. */
int cuopus_probe(const uint8_t *ogg, size_t n, int64_t *samples16k, char *errbuf, size_t errcap);

/* Decode + resample into out_dev (capacity out_cap floats) on the default
 * stream.  Replaces libcuopus.cu:462-468. */
int cuopus_decode_f16k(const uint8_t *ogg, size_t n, float *out_dev, size_t out_cap, int64_t *out_len,
                       char *errbuf, size_t errcap);

/* Same on an explicit stream (hipStream_t).  Replaces libcuopus.cu:238-460. */
int cuopus_decode_f16k_on(const uint8_t *ogg, size_t n, float *out_dev, size_t out_cap, int64_t *out_len,
                          void *stream, char *errbuf, size_t errcap);

/* N files in one decode launch + one batched resample launch; per-file output
 * bit-identical to cuopus_decode_f16k_on.  Replaces libcuopus.cu:482-741
 * (without the reference's max-batch limit: the chain count is bounded only
 * by device memory). */
int cuopus_decode_f16k_multi(const uint8_t *const *oggs, const size_t *ns, int nfiles, float *const *out_devs,
                             const size_t *out_caps, int64_t *out_lens, void *stream, char *errbuf, size_t errcap);

/* Optional: load the kernels and size the workspaces once at start-up.
 * Replaces libcuopus.cu:746-776. */
int cuopus_warmup(char *errbuf, size_t errcap);

/* Development aid, no counterpart in the reference: in a library built with
 * -DCUO_PROF, copies the decoder's per-scope shader-clock cycle sums into
 * out (up to cap entries), clears them and returns the scope count; returns
 * 0 in a normal build. */
int cuopus_prof_read(unsigned long long *out, int cap);

/* 1 if the last decode failed for lack of per-chain decoder memory.  The
 * decoder state here is statically sized, so this is always 0 (kept for ABI
 * compatibility).  Replaces libcuopus.cu:201. */
int cuopus_last_arena_overflow(void);

/* Bytes held in grow-only device workspaces.  Replaces libcuopus.cu:203. */
long long cuopus_workspace_bytes(void);

/* Free every workspace.  Replaces libcuopus.cu:205-221. */
void cuopus_release_workspaces(void);

/* ---------------------------------------------------------------------------
 * Additions on the same path (no counterpart symbol in the reference).
 * ------------------------------------------------------------------------- */

/* Playable int16 48 kHz PCM (preroll and pre-skip removed) into out_dev, in
 * the stream's channels (stereo interleaved L R); out_cap in int16 elements,
 * *out_len = playable samples per channel.  The reference produces the mono
 * buffer internally (its cuopus_compact output, cuopus_kernels.cuh:420-433).
 * The f32 16 kHz entry points decode a stereo stream to one channel the way
 * libopus opened with one channel does (SILK mid channel, CELT bands summed
 * before synthesis). */
int cuopus_decode_s16_on(const uint8_t *ogg, size_t n, int16_t *out_dev, size_t out_cap, int64_t *out_len,
                         void *stream, char *errbuf, size_t errcap);

/* As cuopus_decode_f16k_on with a per-packet loss mask (lost[i] != 0: audio
 * packet i is treated as lost): it is decoded from the next packet's in-band
 * FEC (LBRR) when that packet is in the same chain, not lost and carries LBRR,
 * otherwise concealed for its TOC duration.  mask_n must equal the number of
 * audio packets (see cuopus_probe_packets). */
int cuopus_decode_lossy_f16k_on(const uint8_t *ogg, size_t n, const uint8_t *lost, size_t mask_n, float *out_dev,
                                size_t out_cap, int64_t *out_len, void *stream, char *errbuf, size_t errcap);

/* Number of audio packets and chains of a stream (host only). */
int cuopus_probe_packets(const uint8_t *ogg, size_t n, int64_t *n_packets, int64_t *n_chains, char *errbuf,
                         size_t errcap);

/* Chain table of a stream (host only): out4[4*i..4*i+3] = first packet,
 * packet count, trim, keep of chain i (the reference CLI's --dump-chains CUC2
 * records, decode_cli.cu:693-701 / cuopus_demux.h:396-411).  With out4 NULL
 * only *n_chains is returned. */
int cuopus_dump_chains(const uint8_t *ogg, size_t n, uint32_t *out4, size_t cap_chains, int64_t *n_chains,
                       char *errbuf, size_t errcap);

/* Audio packets of a stream (host only): the reference CLI's --dump-blob CUOP
 * payload (cuopus_demux.h:387-397 / ogg_index.py cmd_extract): lens[i] per
 * packet and the packet bytes back to back.  data or lens may be NULL to
 * query *n_packets / *n_bytes. */
int cuopus_dump_packets(const uint8_t *ogg, size_t n, uint8_t *data, size_t data_cap, uint32_t *lens,
                        size_t cap_packets, int64_t *n_packets, int64_t *n_bytes, char *errbuf, size_t errcap);

/* Demuxer report (host only; no counterpart in the reference, whose
 * ogg_index.py cmd_index prints page / CRC / granule figures): s10[0] pages
 * of the Opus stream(s) used, [1] chained links, [2] pages failing the CRC
 * (CUOPUS_CHECK_CRC=1), [3] pages of other logical streams, [4] packets
 * dropped as incomplete, [5] granule mismatches accepted (CUOPUS_GRANULE=warn),
 * [6] granules checked, [7] bytes outside accepted pages, [8] audio packets,
 * [9] playable 48 kHz samples. */
int cuopus_demux_stats(const uint8_t *ogg, size_t n, int64_t *s10, char *errbuf, size_t errcap);

/* Per-packet gate (the reference's harness2.cu:51-91 protocol, host
 * buffers): decode the chains of an already-demuxed packet table, each chain
 * from a fresh decoder.  Outputs per packet the decoded sample count and the
 * final range, per chain the stop code, and the compacted keep-window PCM
 * (pcm_out, sum(keep) int16).  lost may be NULL. */
int cuopus_gate_chains(const uint8_t *data, size_t nbytes, const uint32_t *off, const uint32_t *len,
                       uint32_t n_packets, const uint8_t *lost, const uint32_t *first, const uint32_t *npkts,
                       const uint32_t *trim, const uint32_t *keep, uint32_t n_chains, int16_t *pcm_out,
                       int32_t *pkt_nsamp, uint32_t *pkt_rng, int32_t *chain_stop, void *stream, char *errbuf,
                       size_t errcap);

/* cuopus_gate_chains with an output channel count (1 or 2; pcm_out holds
 * sum(keep) x channels interleaved samples) and the kernel build: stereo_build
 * != 0 (or channels == 2) decodes with the stereo-capable build, which also
 * decodes stereo packets for a one-channel output as libopus does. */
int cuopus_gate_chains_ch(const uint8_t *data, size_t nbytes, const uint32_t *off, const uint32_t *len,
                          uint32_t n_packets, const uint8_t *lost, const uint32_t *first, const uint32_t *npkts,
                          const uint32_t *trim, const uint32_t *keep, uint32_t n_chains, int channels,
                          int stereo_build, int16_t *pcm_out, int32_t *pkt_nsamp, uint32_t *pkt_rng,
                          int32_t *chain_stop, void *stream, char *errbuf, size_t errcap);

/* Parsed files: demux + chain table once (host only), then decode.  info6,
 * when non-NULL: [0] f32 16 kHz output length, [1] playable 48 kHz samples,
 * [2] channels, [3] audio packets, [4] chains, [5] OpusHead output gain
 * (Q7.8 dB).  The decode calls behave as cuopus_decode_f16k_on /
 * cuopus_decode_f16k_multi on the same bytes (same output, same return
 * codes) without parsing the stream again; the Python binding uses them so a
 * file is demuxed once per call.  A handle is freed by cuopus_file_close and
 * may be decoded any number of times. */
typedef struct cuopus_file cuopus_file;
int cuopus_file_open(const uint8_t *ogg, size_t n, cuopus_file **out, int64_t *info6, char *errbuf, size_t errcap);
int cuopus_file_decode_f16k_on(cuopus_file *f, float *out_dev, size_t out_cap, int64_t *out_len, void *stream,
                               char *errbuf, size_t errcap);
int cuopus_files_decode_f16k_multi(cuopus_file *const *fs, int nfiles, float *const *out_devs, const size_t *out_caps,
                                   int64_t *out_lens, void *stream, char *errbuf, size_t errcap);
void cuopus_file_close(cuopus_file *f);

/* Resample playable int16 48 kHz device PCM to f32 16 kHz (length
 * max(0,(n48-48)/3)) with the library's resampler kernel. */
int cuopus_resample_s16_to_f16k(const int16_t *in_dev, int64_t n48, float *out_dev, size_t out_cap,
                                int64_t *out_len, void *stream, char *errbuf, size_t errcap);

/* Log-mel features of a resident f32 16 kHz device buffer (the consumer the
 * reference keeps its PCM on the device for, ROADMAP.md:149-156; no
 * counterpart entry point).  Frame t covers in_dev[160 t, 160 t + 400): 25 ms
 * every 10 ms, no edge padding, *frames = n16 >= 400 ? 1 + (n16 - 400) / 160
 * : 0.  Periodic Hann window (400), zero-padded 512-point power spectrum,
 * n_mels (1..128) triangular filters on the Slaney mel scale from 0 to
 * 8000 Hz with Slaney area normalisation, out = ln(max(mel, 1e-10)) as
 * row-major (frames, n_mels) float32; out_cap in floats.  Synchronises the
 * stream before returning. */
int cuopus_log_mel(const float *in_dev, int64_t n16, float *out_dev, size_t out_cap, int64_t *frames, int n_mels,
                   void *stream, char *errbuf, size_t errcap);

/* The filter weights cuopus_log_mel applies (host only): out[m * 257 + k],
 * row-major (n_mels, 257) float32, cap in floats.  Returns 0 or
 * CUOPUS_RC_ARG. */
int cuopus_mel_filters(int n_mels, float *out, size_t cap);

/* ---------------------------------------------------------------------------
 * Resident batches: demux + upload once, then decode from HBM-resident
 * inputs any number of times.  A batch owns its device memory; decoded f32
 * 16 kHz output goes to caller-owned device buffers given at creation.
 * Runs are asynchronous on the given stream; cuopus_batch_check reads back
 * and validates the last run (synchronising the stream).
 * ------------------------------------------------------------------------- */
typedef struct cuopus_batch cuopus_batch;

typedef struct {
    int mode;        /* 0: one decode + one resample launch per file (the plain
                        per-file path); 1: every chain in one launch, one
                        workgroup per chain; 2: persistent kernel, n_slots
                        wavefronts pull chains from a device work queue */
    int n_slots;     /* mode 2: resident wavefronts (0 = library default) */
    int lpt_order;   /* work order (modes 1-2): 0 file order, 1 costliest
                        chains first by the per-packet cost model (TOC mode
                        and frame duration), 2 longest chains first by
                        samples + 960 x packets, 3 the library default
                        (tuned, else 1) */
    int use_graph;   /* 1: capture the launch sequence into a HIP graph on
                        the first run and replay it afterwards */
    int profile;     /* 1: time the decode and resample kernels with events */
    int rs_variant;  /* resampler launch shape: 0 = library default (tuned),
                        v + 1 = variant v (threads x outputs per thread:
                        0: 256x4, 1: 256x8, 2: 128x8, 3: 512x4, 4: 128x4) */
} cuopus_batch_opts;

/* lost_masks may be NULL (no loss) or hold one mask per file (NULL entries
 * allowed); mask_ns[i] must equal file i's audio packet count. */
int cuopus_batch_create(const uint8_t *const *oggs, const size_t *ns, int nfiles, const uint8_t *const *lost_masks,
                        const size_t *mask_ns, float *const *out_devs, const size_t *out_caps,
                        const cuopus_batch_opts *opts, cuopus_batch **out, char *errbuf, size_t errcap);
int cuopus_batch_run(cuopus_batch *b, void *stream, char *errbuf, size_t errcap);
int cuopus_batch_check(cuopus_batch *b, void *stream, int64_t *out_lens, char *errbuf, size_t errcap);
/* counts: [0] chains, [1] packets, [2] playable 48 kHz samples, [3] samples
 * decoded including preroll, [4] f32 16 kHz outputs, [5] packet bytes,
 * [6] packets read by the chains (preroll packets counted once per chain),
 * [7] packet bytes read by the chains */
int cuopus_batch_info(const cuopus_batch *b, int64_t *counts8);
/* The launch defaults in effect on the current device (no counterpart in
 * the reference): persistent wavefronts, resampler variant, work order --
 * CUOPUS_SLOTS / CUOPUS_RS, else this device's line of cuopus_tuned.txt
 * (beside the library, or the file CUOPUS_TUNED names), else built-in. */
int cuopus_tuned_defaults(int *slots, int *rs_variant, int *order);

/* Per-chain start / end times of the last run (constant 100 MHz device clock,
 * low 32 bits), indexed by chain in file order; cap >= chain count.  For
 * calibrating the work-order cost model and measuring the persistent
 * kernel's tail.  Synchronises the stream. */
int cuopus_batch_chain_times(cuopus_batch *b, void *stream, uint32_t *t0, uint32_t *t1, int64_t cap);
/* Kernel times of the last profiled run in milliseconds (decode, resample). */
int cuopus_batch_kernel_ms(const cuopus_batch *b, float *decode_ms, float *resample_ms);
void cuopus_batch_destroy(cuopus_batch *b);

/* ---------------------------------------------------------------------------
 * Serving: continuous batching (no counterpart in the reference).  Requests
 * are demuxed on submit and queue in arrival order; cuopus_server_poll never
 * blocks: it retires finished launches, admits queued requests into a new
 * launch on a free stream while their device bytes (packets, chain table,
 * int16 PCM, f32 output) fit the memory budget, and reports finished
 * requests.  Launches on different streams run concurrently.  Outputs are
 * written to the caller's device buffers given at submit; the caller keeps
 * them alive until the request is reported.  Single-threaded use only.
 * ------------------------------------------------------------------------- */
typedef struct cuopus_server cuopus_server;
typedef struct {
    int n_streams;              /* launches in flight (0: 4) */
    int n_waves;                /* persistent wavefronts per launch (0: default) */
    long long mem_budget;       /* device bytes the server may hold: every launch
                                   slot's buffers (packets, tables, int16 PCM,
                                   decoder heaps), busy or idle (0: 8 GiB) */
    int max_chains_per_launch;  /* 0: unlimited */
} cuopus_server_opts;

int cuopus_server_create(const cuopus_server_opts *opts, cuopus_server **out, char *errbuf, size_t errcap);
/* lost/mask_n: optional per-packet loss mask (FEC/PLC rule as the lossy
 * decode).  stream: the caller's stream (hipStream_t) that last used out_dev;
 * the request's launch waits for the point of this call on that stream
 * before it writes out_dev.  The output is complete when poll reports the
 * request. */
int cuopus_server_submit(cuopus_server *srv, const uint8_t *ogg, size_t n, const uint8_t *lost, size_t mask_n,
                         float *out_dev, size_t out_cap, void *stream, uint64_t *req_id, char *errbuf,
                         size_t errcap);
/* Up to cap finished requests: ids[i] and rcs[i] (0 ok, else the request's
 * error; errbuf then holds the last failing request's message). */
int cuopus_server_poll(cuopus_server *srv, uint64_t *ids, int *rcs, int cap, int *n_done, char *errbuf,
                       size_t errcap);
/* s4: [0] queued requests [1] launches in flight [2] device bytes the server
 * holds (never above mem_budget) [3] finished, not yet reported */
int cuopus_server_stats(const cuopus_server *srv, int64_t *s4);
void cuopus_server_destroy(cuopus_server *srv);

/* ---------------------------------------------------------------------------
 * Allocation-free decode.  No counterpart in the reference: for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  As cuopus_decode_f16k_on (same
 * kernels, same output), but every device buffer of the call (packets,
 * tables, decoder heaps, int16 PCM) lives in one static device array of
 * 128 MiB per GPU, reserved when the library loads and kept until the process
 * ends; the call makes no hipMalloc/hipFree and no allocating library call.
 * A stream whose buffers exceed it returns CUOPUS_RC_ARG with the byte count
 * (about 10 minutes of 64 kbps mono audio fit: 2 B per playable 48 kHz sample
 * and channel plus the packets, 64 B per chain and 38 KB per resident
 * wavefront, 60 KB for a stream that needs the stereo build, 64-4096 of
 * them).  Every call on a GPU shares that GPU's array: calls that overlap on
 * one GPU (two host threads, or two streams without ordering) race. */
int cuopus_decode_f16k_noalloc_on(const uint8_t *ogg, size_t n, float *out_dev, size_t out_cap, int64_t *out_len,
                                  void *stream, char *errbuf, size_t errcap);

#ifdef __cplusplus
}
#endif

#endif /* CUOPUS_H */
