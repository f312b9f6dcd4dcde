// entdec.h -- RFC 6716 section 4.1 range decoder.
//
// Same state and update rules as libopus's ec_dec (vendor/celt/entdec.c:91-266,
// entcode.c:69-86, entcode.h:100-120): 32-bit rng/val, 8-bit symbols, a
// 1-byte "rem" carry, raw bits read backwards from the end of the frame into a
// 32-bit window.  ec_tell()/ec_tell_frac() derive from nbits_total exactly as
// libopus so every bit-budget decision downstream matches.
//
// When a whole wavefront decodes one chain, every field here is wave-uniform,
// so the compiler keeps the state in SGPRs and runs the updates on the SALU.
#pragma once
#include "common.h"

namespace CUO_NS {

struct RangeDec {
    const uint8_t *buf;
    uint32_t storage;
    uint32_t end_offs;
    uint32_t end_window;
    int nend_bits;
    int nbits_total;
    uint32_t offs;
    uint32_t rng;
    uint32_t val;
    uint32_t ext;
    int rem;
    int error;
};

constexpr uint32_t kEcCodeTop = 1u << 31;
constexpr uint32_t kEcCodeBot = kEcCodeTop >> 8;
constexpr int kEcCodeExtra = 7;  // (32-2)%8+1

// Every function loads the fields it needs as wave-uniform values (scalar
// registers), updates them and stores them back.
CUO_FI const uint8_t *ec_buf(const RangeDec *d)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return (const uint8_t *)(uintptr_t)uniform_u64((uint64_t)(uintptr_t)d->buf);
#else
    return d->buf;
#endif
}

// A register copy of a range decoder (every field wave-uniform): hot symbol
// loops decode on the copy and store it back once.
CUO_FI RangeDec ec_regs(const RangeDec *d)
{
    RangeDec r;
    r.buf = ec_buf(d);
    r.storage = uniform(d->storage);
    r.end_offs = uniform(d->end_offs);
    r.end_window = uniform(d->end_window);
    r.nend_bits = uniform(d->nend_bits);
    r.nbits_total = uniform(d->nbits_total);
    r.offs = uniform(d->offs);
    r.rng = uniform(d->rng);
    r.val = uniform(d->val);
    r.ext = uniform(d->ext);
    r.rem = uniform(d->rem);
    r.error = uniform(d->error);
    return r;
}

// Renormalise after a symbol: rng and val are the updated values.
CUO_FI void ec_normalize_rv(RangeDec *d, uint32_t rng, uint32_t val)
{
    if (rng <= kEcCodeBot) {
        int nbits = uniform(d->nbits_total);
        int rem = uniform(d->rem);
        uint32_t offs = uniform(d->offs), storage = uniform(d->storage);
        const uint8_t *buf = ec_buf(d);
        do {
            nbits += 8;
            rng <<= 8;
            int sym = rem;
            rem = offs < storage ? (int)ld_u8(buf, (int)offs++) : 0;
            sym = (sym << 8 | rem) >> (8 - kEcCodeExtra);
            val = ((val << 8) + (255 & ~(uint32_t)sym)) & (kEcCodeTop - 1);
        } while (rng <= kEcCodeBot);
        d->nbits_total = nbits;
        d->rem = rem;
        d->offs = offs;
    }
    d->rng = rng;
    d->val = val;
}
CUO_FI void ec_normalize(RangeDec *d) { ec_normalize_rv(d, uniform(d->rng), uniform(d->val)); }

CUO_FN void ec_init(RangeDec *d, const uint8_t *buf, uint32_t storage)
{
    d->buf = buf;
    d->storage = storage;
    d->end_offs = 0;
    d->end_window = 0;
    d->nend_bits = 0;
    d->nbits_total = 32 + 1 - ((32 - kEcCodeExtra) / 8) * 8;
    d->offs = storage > 0 ? 1 : 0;
    int rem = storage > 0 ? (int)ld_u8(buf, 0) : 0;
    d->rem = rem;
    uint32_t rng = 1u << kEcCodeExtra;
    d->error = 0;
    d->ext = 0;
    ec_normalize_rv(d, rng, rng - 1 - (uint32_t)(rem >> (8 - kEcCodeExtra)));
}

CUO_FI uint32_t ec_decode(RangeDec *d, uint32_t ft)
{
    uint32_t ext = uniform(d->rng) / ft;
    d->ext = ext;
    uint32_t s = uniform(d->val) / ext;
    uint32_t t = s + 1;
    return ft - (t < ft ? t : ft);
}

CUO_FI uint32_t ec_decode_bin(RangeDec *d, int bits)
{
    uint32_t ext = uniform(d->rng) >> bits;
    d->ext = ext;
    uint32_t s = uniform(d->val) / ext;
    uint32_t t = s + 1, ft = 1u << bits;
    return ft - (t < ft ? t : ft);
}

CUO_FI void ec_update(RangeDec *d, uint32_t fl, uint32_t fh, uint32_t ft)
{
    uint32_t ext = uniform(d->ext), rng = uniform(d->rng), val = uniform(d->val);
    uint32_t s = ext * (ft - fh);
    val -= s;
    rng = fl > 0 ? ext * (fh - fl) : rng - s;
    ec_normalize_rv(d, rng, val);
}

CUO_FI int ec_bit_logp(RangeDec *d, int logp)
{
    uint32_t r = uniform(d->rng), v = uniform(d->val);
    uint32_t s = r >> logp;
    int ret = v < s;
    if (!ret) v -= s;
    ec_normalize_rv(d, ret ? s : r - s, v);
    return ret;
}

template <typename T>
CUO_FN uint32_t icdf_at(const T *icdf, int i);
template <>
CUO_FI uint32_t icdf_at<uint8_t>(const uint8_t *icdf, int i) { return ld_u8(icdf, i); }

// icdf must point into a constant table (it is read with scalar loads).
template <typename T>
CUO_FI int ec_icdf_t(RangeDec *d, const T *icdf, int ftb)
{
    uint32_t s = uniform(d->rng), v = uniform(d->val);
    uint32_t r = s >> ftb;
    int ret = -1;
    uint32_t t;
#if defined(__HIP_DEVICE_COMPILE__)
    if (sizeof(T) == 1) {
        // same search, the table read a 32-bit word (four entries) per load
        const uint8_t *p = (const uint8_t *)icdf;
        const uint32_t mis = (uint32_t)(uintptr_t)p & 3u;
        uint32_t w = const_word(p) >> (8 * mis);
        int left = 4 - (int)mis;
        ret = 0;
        for (;;) {
            t = s;
            s = r * (w & 0xFFu);
            if (v >= s) break;
            ret++;
            if (--left == 0) {
                p += 4 - (int)((uintptr_t)p & 3u);
                w = const_word(p);
                left = 4;
            } else {
                w >>= 8;
            }
        }
        ec_normalize_rv(d, t - s, v - s);
        return ret;
    }
#endif
    do {
        t = s;
        s = r * icdf_at<T>(icdf, ++ret);
    } while (v < s);
    ec_normalize_rv(d, t - s, v - s);
    return ret;
}
CUO_FI int ec_icdf(RangeDec *d, const uint8_t *icdf, int ftb) { return ec_icdf_t<uint8_t>(d, icdf, ftb); }

// ec_icdf with the two-entry table {icdf0, 0}: symbol 0 takes [r*icdf0, rng),
// symbol 1 takes [0, r*icdf0).
CUO_FI int ec_icdf2(RangeDec *d, uint32_t icdf0, int ftb)
{
    uint32_t rng = uniform(d->rng), v = uniform(d->val);
    uint32_t s = (rng >> ftb) * icdf0;
    if (v >= s) {
        ec_normalize_rv(d, rng - s, v - s);
        return 0;
    }
    ec_normalize_rv(d, s, v);
    return 1;
}

CUO_FI uint32_t ec_bits(RangeDec *d, int bits)
{
    uint32_t window = uniform(d->end_window);
    int available = uniform(d->nend_bits);
    if (available < bits) {
        uint32_t end_offs = uniform(d->end_offs), storage = uniform(d->storage);
        const uint8_t *buf = ec_buf(d);
        do {
            uint32_t b = end_offs < storage ? ld_u8(buf, (int)(storage - ++end_offs)) : 0;
            window |= b << available;
            available += 8;
        } while (available <= 32 - 8);
        d->end_offs = end_offs;
    }
    uint32_t ret = window & ((1u << bits) - 1u);
    window >>= bits;
    available -= bits;
    d->end_window = window;
    d->nend_bits = available;
    d->nbits_total = uniform(d->nbits_total) + bits;
    return ret;
}

CUO_FI uint32_t ec_uint(RangeDec *d, uint32_t ft)
{
    ft--;
    int ftb = ilog(ft);
    if (ftb > 8) {
        ftb -= 8;
        uint32_t f = (ft >> ftb) + 1;
        uint32_t s = ec_decode(d, f);
        ec_update(d, s, s + 1, f);
        uint32_t t = s << ftb | ec_bits(d, ftb);
        if (t <= ft) return t;
        d->error = 1;
        return ft;
    }
    ft++;
    uint32_t s = ec_decode(d, ft);
    ec_update(d, s, s + 1, ft);
    return s;
}

CUO_FI int ec_tell(const RangeDec *d) { return uniform(d->nbits_total) - ilog(uniform(d->rng)); }

static constexpr uint32_t kTellCorrection[8] = {35733, 38967, 42495, 46340, 50535, 55109, 60097, 65535};
constexpr uint64_t kTellCorrLo = 0xB504A5FF98378B95ull, kTellCorrHi = 0xFFFFEAC1D745C567ull;
constexpr bool tell_corr_packed()
{
    for (int b = 0; b < 8; b++)
        if ((((b < 4 ? kTellCorrLo : kTellCorrHi) >> (16 * (b & 3))) & 0xFFFFu) != kTellCorrection[b]) return false;
    return true;
}
static_assert(tell_corr_packed(), "packed tell-correction constants differ from the table");

CUO_FI uint32_t ec_tell_frac(const RangeDec *d)
{
    // 1/8-bit resolution via a linear estimate plus threshold table
    // (vendor/celt/entcode.c:69-86)
    uint32_t nbits = (uint32_t)uniform(d->nbits_total) << kBitRes;
    uint32_t rng = uniform(d->rng);
    int l = ilog(rng);
    uint32_t r = rng >> (l - 16);
    uint32_t b = (r >> 12) - 8;
    // kTellCorrection[b] from two 64-bit constants (four 16-bit entries each)
    const uint64_t corr = b < 4 ? kTellCorrLo : kTellCorrHi;
    b += r > (uint32_t)((corr >> (16 * (b & 3))) & 0xFFFFu);
    l = (l << 3) + (int)b;
    return nbits - (uint32_t)l;
}

}  // namespace CUO_NS
