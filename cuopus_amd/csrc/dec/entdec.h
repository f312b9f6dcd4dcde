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

namespace cuo {

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

CUO_FN int ec_read_byte(RangeDec *d) { return d->offs < d->storage ? d->buf[d->offs++] : 0; }
CUO_FN int ec_read_byte_end(RangeDec *d)
{
    return d->end_offs < d->storage ? d->buf[d->storage - ++(d->end_offs)] : 0;
}

CUO_FN void ec_normalize(RangeDec *d)
{
    while (d->rng <= kEcCodeBot) {
        d->nbits_total += 8;
        d->rng <<= 8;
        int sym = d->rem;
        d->rem = ec_read_byte(d);
        sym = (sym << 8 | d->rem) >> (8 - kEcCodeExtra);
        d->val = ((d->val << 8) + (255 & ~(uint32_t)sym)) & (kEcCodeTop - 1);
    }
}

CUO_FN void ec_init(RangeDec *d, const uint8_t *buf, uint32_t storage)
{
    d->buf = buf;
    d->storage = storage;
    d->end_offs = 0;
    d->end_window = 0;
    d->nend_bits = 0;
    d->nbits_total = 32 + 1 - ((32 - kEcCodeExtra) / 8) * 8;
    d->offs = 0;
    d->rng = 1u << kEcCodeExtra;
    d->rem = ec_read_byte(d);
    d->val = d->rng - 1 - (d->rem >> (8 - kEcCodeExtra));
    d->error = 0;
    d->ext = 0;
    ec_normalize(d);
}

CUO_FN uint32_t ec_decode(RangeDec *d, uint32_t ft)
{
    d->ext = d->rng / ft;
    uint32_t s = d->val / d->ext;
    uint32_t t = s + 1;
    return ft - (t < ft ? t : ft);
}

CUO_FN uint32_t ec_decode_bin(RangeDec *d, int bits)
{
    d->ext = d->rng >> bits;
    uint32_t s = d->val / d->ext;
    uint32_t t = s + 1, ft = 1u << bits;
    return ft - (t < ft ? t : ft);
}

CUO_FN void ec_update(RangeDec *d, uint32_t fl, uint32_t fh, uint32_t ft)
{
    uint32_t s = d->ext * (ft - fh);
    d->val -= s;
    d->rng = fl > 0 ? d->ext * (fh - fl) : d->rng - s;
    ec_normalize(d);
}

CUO_FN int ec_bit_logp(RangeDec *d, int logp)
{
    uint32_t r = d->rng, v = d->val;
    uint32_t s = r >> logp;
    int ret = v < s;
    if (!ret) d->val = v - s;
    d->rng = ret ? s : r - s;
    ec_normalize(d);
    return ret;
}

template <typename T>
CUO_FN int ec_icdf_t(RangeDec *d, const T *icdf, int ftb)
{
    uint32_t s = d->rng, v = d->val;
    uint32_t r = s >> ftb;
    int ret = -1;
    uint32_t t;
    do {
        t = s;
        s = r * (uint32_t)icdf[++ret];
    } while (v < s);
    d->val = v - s;
    d->rng = t - s;
    ec_normalize(d);
    return ret;
}
CUO_FN int ec_icdf(RangeDec *d, const uint8_t *icdf, int ftb) { return ec_icdf_t<uint8_t>(d, icdf, ftb); }

CUO_FN uint32_t ec_bits(RangeDec *d, int bits)
{
    uint32_t window = d->end_window;
    int available = d->nend_bits;
    if (available < bits) {
        do {
            window |= (uint32_t)ec_read_byte_end(d) << available;
            available += 8;
        } while (available <= 32 - 8);
    }
    uint32_t ret = window & ((1u << bits) - 1u);
    window >>= bits;
    available -= bits;
    d->end_window = window;
    d->nend_bits = available;
    d->nbits_total += bits;
    return ret;
}

CUO_FN uint32_t ec_uint(RangeDec *d, uint32_t ft)
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

CUO_FN int ec_tell(const RangeDec *d) { return d->nbits_total - ilog(d->rng); }

static constexpr uint32_t kTellCorrection[8] = {35733, 38967, 42495, 46340, 50535, 55109, 60097, 65535};

CUO_FN uint32_t ec_tell_frac(const RangeDec *d)
{
    // 1/8-bit resolution via a linear estimate plus threshold table
    // (vendor/celt/entcode.c:69-86)
    uint32_t nbits = (uint32_t)d->nbits_total << kBitRes;
    int l = ilog(d->rng);
    uint32_t r = d->rng >> (l - 16);
    uint32_t b = (r >> 12) - 8;
    b += r > kTellCorrection[b];
    l = (l << 3) + (int)b;
    return nbits - (uint32_t)l;
}

}  // namespace cuo
