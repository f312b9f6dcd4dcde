// fuzz_gate.cpp -- malformed-input check of the decoder source on the CPU.
//
// Test infrastructure.  Built with AddressSanitizer and UndefinedBehavior-
// Sanitizer (cuopus_amd/build.py build_fuzz_gate): decodes every chain of a
// packed set (tests/malformed.py write_set) with the device decoder's source
// compiled for the host (the same-source gate) and with the oracle (libopus
// 1.5.2 fixed point, liboracle.so, not instrumented), and compares each
// packet's return value and final range and every decoded sample up to the
// first error of the chain.  Optional second argument: output channels (1
// or 2; 2 needs the stereo build, fuzz_gate_st).  Prints one line per
// mismatching chain; exit status 0 when everything matches, 1 on
// mismatches, 2 on bad input.
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <vector>

#include "hostgate.cpp"

extern "C" long orc_decode_chain_cfg(const unsigned char *data, const uint32_t *off, const uint32_t *len,
                                     const uint8_t *lost, int npk, int16_t *pcm, long cap, int32_t *nsamp,
                                     uint32_t *rng, int channels, int gain);

static bool rd(FILE *f, void *p, size_t n) { return n == 0 || fread(p, 1, n, f) == n; }

int main(int argc, char **argv)
{
    if (argc != 2 && argc != 3) { fprintf(stderr, "usage: %s set.bin [channels]\n", argv[0]); return 2; }
    const int C = argc == 3 ? atoi(argv[2]) : 1;
    if (C < 1 || C > kMaxCh) { fprintf(stderr, "channels must be 1..%d\n", kMaxCh); return 2; }
    FILE *f = fopen(argv[1], "rb");
    if (!f) { perror(argv[1]); return 2; }
    char magic[4];
    uint32_t hdr[4];
    if (!rd(f, magic, 4) || memcmp(magic, "CUFZ", 4) || !rd(f, hdr, sizeof(hdr))) return 2;
    const uint32_t npk = hdr[0], nch = hdr[1], nbytes = hdr[2], has_lost = hdr[3];
    std::vector<uint32_t> off(npk), len(npk), first(nch), npkts(nch), keep(nch);
    std::vector<uint8_t> lost(has_lost ? npk : 0), data(nbytes + 8, 0);
    bool ok = rd(f, off.data(), 4 * npk) && rd(f, len.data(), 4 * npk) && (!has_lost || rd(f, lost.data(), npk)) &&
              rd(f, first.data(), 4 * nch) && rd(f, npkts.data(), 4 * nch) && rd(f, keep.data(), 4 * nch) &&
              rd(f, data.data(), nbytes);
    fclose(f);
    if (!ok) return 2;
    int bad = 0;
    long packets = 0, samples = 0;
    for (uint32_t c = 0; c < nch; c++) {
        const uint32_t n = npkts[c], f0 = first[c];
        const long cap = 5760L * (n + 1);
        std::vector<int16_t> pa(cap * C), pb(cap * C);
        std::vector<int32_t> na(n, 0), nb(n, 0);
        std::vector<uint32_t> ra(n, 0), rb(n, 0);
        const uint8_t *lp = has_lost ? lost.data() + f0 : nullptr;
        long a = hg_decode_chain_cfg(data.data(), off.data() + f0, len.data() + f0, lp, (int)n, pa.data(), cap,
                                     na.data(), ra.data(), C, 0);
        long b = orc_decode_chain_cfg(data.data(), off.data() + f0, len.data() + f0, lp, (int)n, pb.data(), cap,
                                      nb.data(), rb.data(), C, 0);
        long done = 0;
        uint32_t k = 0;
        for (; k < n && nb[k] > 0; k++) done += nb[k];
        packets += k;
        samples += done;
        bool same = (a < 0) == (b < 0) && (a < 0 ? a == b : a == b) && memcmp(na.data(), nb.data(), 4 * n) == 0 &&
                    memcmp(ra.data(), rb.data(), 4 * n) == 0 &&
                    memcmp(pa.data(), pb.data(), sizeof(int16_t) * (size_t)(done * C)) == 0 &&
                    (uint32_t)done >= keep[c];
        if (!same) {
            if (bad < 20)
                printf("chain %u: gate %ld oracle %ld (packets ok %u of %u)\n", c, a, b, k, n);
            bad++;
        }
    }
    printf("fuzz_gate: %u chains, %ld packets decoded, %ld samples, %d mismatching chains\n", nch, packets, samples,
           bad);
    return bad ? 1 : 0;
}
