// cuopus_decode -- command-line front-end over libcuopus.so.
//
// The reference CLI's --ogg mode (decode_cli.cu:680-734) with the same
// options:
//   cuopus_decode --ogg <in.ogg> [out.pcm|-] [--seg S] [--preroll MS]
//                 [--stride N] [--f16k] [--dump-blob F] [--dump-chains F]
// Output: the playable int16 48 kHz PCM (little-endian, interleaved when the
// stream is stereo), or with --f16k the float32 16 kHz samples; "-" writes to
// stdout.  --dump-blob writes the CUOP packet dump ('CUOP', u32 count, u32
// lengths, packet bytes) and --dump-chains the CUC2 chain table ('CUC2', u32
// count, 4 x u32 first/npkts/trim/keep per chain), both host-only.  --seg and
// --preroll set the chain length and decoder preroll (CUOPUS_SEG_S,
// CUOPUS_PREROLL_MS).  --stride is accepted for command-line compatibility:
// the reference's chains-per-warp knob; here every chain has a wavefront of
// its own.  The reference's legacy blob mode and --serve daemon are not part
// of this decode path.
//
// Uses only the C ABI (include/cuopus.h) plus hipMalloc/hipMemcpy for the
// output buffer.  Exit status 0 ok, 1 decode failure, 2 usage.
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <string>
#include <vector>

#include "cuopus.h"

static int usage(const char *argv0)
{
    fprintf(stderr,
            "usage: %s --ogg <in.ogg> [out.pcm|-] [--seg S] [--preroll MS] [--stride N] [--f16k]\n"
            "                 [--dump-blob F] [--dump-chains F]\n",
            argv0);
    return 2;
}

static bool read_file(const char *path, std::vector<uint8_t> *out)
{
    FILE *f = fopen(path, "rb");
    if (!f) return false;
    uint8_t buf[1 << 16];
    size_t n;
    while ((n = fread(buf, 1, sizeof(buf), f)) > 0) out->insert(out->end(), buf, buf + n);
    fclose(f);
    return true;
}

static bool put_u32(FILE *f, uint32_t v)
{
    uint8_t b[4] = {(uint8_t)v, (uint8_t)(v >> 8), (uint8_t)(v >> 16), (uint8_t)(v >> 24)};
    return fwrite(b, 1, 4, f) == 4;
}

static int dump_blob(const std::vector<uint8_t> &ogg, const char *path)
{
    char err[512];
    int64_t np = 0, nb = 0;
    int rc = cuopus_dump_packets(ogg.data(), ogg.size(), nullptr, 0, nullptr, 0, &np, &nb, err, sizeof(err));
    std::vector<uint8_t> data((size_t)(nb > 0 ? nb : 1));
    std::vector<uint32_t> lens((size_t)(np > 0 ? np : 1));
    if (!rc)
        rc = cuopus_dump_packets(ogg.data(), ogg.size(), data.data(), data.size(), lens.data(), lens.size(), &np, &nb,
                                 err, sizeof(err));
    if (rc) { fprintf(stderr, "cuopus_decode: %s\n", err); return 1; }
    FILE *f = fopen(path, "wb");
    if (!f) { perror(path); return 1; }
    bool ok = fwrite("CUOP", 1, 4, f) == 4 && put_u32(f, (uint32_t)np);
    for (int64_t i = 0; ok && i < np; i++) ok = put_u32(f, lens[(size_t)i]);
    ok = ok && (nb == 0 || fwrite(data.data(), 1, (size_t)nb, f) == (size_t)nb);
    ok = fclose(f) == 0 && ok;
    if (!ok) { fprintf(stderr, "cuopus_decode: write %s failed\n", path); return 1; }
    return 0;
}

static int dump_chains(const std::vector<uint8_t> &ogg, const char *path)
{
    char err[512];
    int64_t nc = 0;
    int rc = cuopus_dump_chains(ogg.data(), ogg.size(), nullptr, 0, &nc, err, sizeof(err));
    std::vector<uint32_t> t((size_t)(4 * (nc > 0 ? nc : 1)));
    if (!rc) rc = cuopus_dump_chains(ogg.data(), ogg.size(), t.data(), (size_t)nc, &nc, err, sizeof(err));
    if (rc) { fprintf(stderr, "cuopus_decode: %s\n", err); return 1; }
    FILE *f = fopen(path, "wb");
    if (!f) { perror(path); return 1; }
    bool ok = fwrite("CUC2", 1, 4, f) == 4 && put_u32(f, (uint32_t)nc);
    for (int64_t i = 0; ok && i < 4 * nc; i++) ok = put_u32(f, t[(size_t)i]);
    ok = fclose(f) == 0 && ok;
    if (!ok) { fprintf(stderr, "cuopus_decode: write %s failed\n", path); return 1; }
    return 0;
}

// Decode to a device buffer, copy back, write.
static int decode_out(const std::vector<uint8_t> &ogg, const char *out, bool f16k)
{
    char err[512];
    int64_t n = 0;
    int rc;
    std::vector<uint8_t> host;
    if (f16k) {
        int64_t n16 = 0;
        if ((rc = cuopus_probe(ogg.data(), ogg.size(), &n16, err, sizeof(err)))) {
            fprintf(stderr, "cuopus_decode: %s\n", err);
            return 1;
        }
        float *dev = nullptr;
        if (hipMalloc(&dev, (size_t)(n16 > 0 ? n16 : 1) * sizeof(float)) != hipSuccess) {
            fprintf(stderr, "cuopus_decode: hipMalloc failed\n");
            return 1;
        }
        rc = cuopus_decode_f16k_on(ogg.data(), ogg.size(), dev, (size_t)n16, &n, nullptr, err, sizeof(err));
        host.resize((size_t)n * sizeof(float));
        if (!rc && n && hipMemcpy(host.data(), dev, host.size(), hipMemcpyDeviceToHost) != hipSuccess) rc = 2;
        hipFree(dev);
    } else {
        cuopus_file *fh = nullptr;
        int64_t info[6] = {0};
        if ((rc = cuopus_file_open(ogg.data(), ogg.size(), &fh, info, err, sizeof(err)))) {
            fprintf(stderr, "cuopus_decode: %s\n", err);
            return 1;
        }
        cuopus_file_close(fh);
        const int64_t channels = info[2] > 0 ? info[2] : 1;
        const int64_t cap = info[1] * channels;
        int16_t *dev = nullptr;
        if (hipMalloc(&dev, (size_t)(cap > 0 ? cap : 1) * sizeof(int16_t)) != hipSuccess) {
            fprintf(stderr, "cuopus_decode: hipMalloc failed\n");
            return 1;
        }
        rc = cuopus_decode_s16_on(ogg.data(), ogg.size(), dev, (size_t)cap, &n, nullptr, err, sizeof(err));
        host.resize((size_t)(n * channels) * sizeof(int16_t));
        if (!rc && n && hipMemcpy(host.data(), dev, host.size(), hipMemcpyDeviceToHost) != hipSuccess) rc = 2;
        hipFree(dev);
    }
    if (rc) {
        fprintf(stderr, "cuopus_decode: %s\n", rc == 2 && !err[0] ? "device copy failed" : err);
        return 1;
    }
    FILE *f = strcmp(out, "-") ? fopen(out, "wb") : stdout;
    if (!f) { perror(out); return 1; }
    bool ok = host.empty() || fwrite(host.data(), 1, host.size(), f) == host.size();
    ok = (f == stdout ? fflush(f) == 0 : fclose(f) == 0) && ok;
    if (!ok) { fprintf(stderr, "cuopus_decode: write failed\n"); return 1; }
    fprintf(stderr, "cuopus_decode: %lld %s samples\n", (long long)n, f16k ? "f32 16 kHz" : "int16 48 kHz");
    return 0;
}

int main(int argc, char **argv)
{
    if (argc < 3 || strcmp(argv[1], "--ogg") != 0) return usage(argv[0]);
    const char *in = argv[2];
    const char *out = nullptr, *blob = nullptr, *chains = nullptr;
    bool f16k = false;
    int ai = 3;
    if (ai < argc && (argv[ai][0] != '-' || !strcmp(argv[ai], "-"))) out = argv[ai++];
    for (; ai < argc; ai++) {
        const char *a = argv[ai];
        if (!strcmp(a, "--seg") && ai + 1 < argc) setenv("CUOPUS_SEG_S", argv[++ai], 1);
        else if (!strcmp(a, "--preroll") && ai + 1 < argc) setenv("CUOPUS_PREROLL_MS", argv[++ai], 1);
        else if (!strcmp(a, "--stride") && ai + 1 < argc) ++ai;
        else if (!strcmp(a, "--f16k")) f16k = true;
        else if (!strcmp(a, "--dump-blob") && ai + 1 < argc) blob = argv[++ai];
        else if (!strcmp(a, "--dump-chains") && ai + 1 < argc) chains = argv[++ai];
        else {
            fprintf(stderr, "cuopus_decode: unknown option %s\n", a);
            return 2;
        }
    }
    std::vector<uint8_t> ogg;
    if (!read_file(in, &ogg)) { perror(in); return 1; }
    if (blob && dump_blob(ogg, blob)) return 1;
    if (chains && dump_chains(ogg, chains)) return 1;
    if (out) return decode_out(ogg, out, f16k);
    return 0;
}
