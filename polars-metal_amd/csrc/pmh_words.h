// pmh_words.h -- column loads as aligned 8-byte words, shared by the
// specialised scans of the hash and sort engines.
//
// A column value (or validity bit) is read by loading the aligned 8-byte word
// that holds it and cutting it out with a shift and mask whose parameters are
// wave-uniform.  A naturally aligned value never straddles words, and an
// aligned word never crosses a page, so the load is always mapped.  With no
// per-width switch and no lane predicate around a load, nothing forces the
// compiler to wait for one load before issuing the next.
#pragma once

#include "pmh_device.h"

namespace pmh {

// Stand-in validity bitmap for columns without one: every bit set.
__device__ const uint64_t pmh_all_valid[1] = {~0ull};

// Order-preserving int64 image of a double under Polars' TotalOrd (the order
// of tot_cmp): -0.0 and +0.0 map together, every NaN maps above +inf.
__host__ __device__ inline int64_t tot_key_bits(int64_t b, bool is_nan) {
    if (is_nan) return 0x7FFFFFFFFFFFFFFFll;
    if ((b & 0x7FFFFFFFFFFFFFFFll) == 0) return 0;
    return b >= 0 ? b : (b ^ 0x7FFFFFFFFFFFFFFFll);
}
__device__ __forceinline__ int64_t tot_key(double x) { return tot_key_bits(__double_as_longlong(x), x != x); }

// A column as one role of the query sees it, addressed by bit position:
// value r occupies bits [bit0 + (r << bshift), ... + width) of the word
// array at base; its validity bit is bit vbit0 + r * vstride of vbase.
struct RoleCol {
    const uint64_t *base, *vbase;
    uint64_t mask;
    uint32_t bit0, vbit0, vstride;
    int32_t bshift, tag, w;  // w: bytes per value, 0 for bit-packed Boolean
};
__device__ __forceinline__ RoleCol role_col(const QueryDesc *__restrict__ q, int c) {
    RoleCol r;
    const uintptr_t d = (uintptr_t)q->cols[c].data;
    const uintptr_t v = (uintptr_t)q->cols[c].valid;
    r.tag = q->cols[c].tag;
    r.w = r.tag == PMH_BOOL ? 0 : tag_size(r.tag);
    r.base = (const uint64_t *)(d & ~uintptr_t(7));
    r.bit0 = (uint32_t)(d & 7) * 8;
    r.bshift = r.w == 8 ? 6 : r.w == 4 ? 5 : r.w == 2 ? 4 : r.w == 1 ? 3 : 0;
    r.mask = r.w == 8 ? ~0ull : r.w == 0 ? 1ull : (1ull << (8 * r.w)) - 1;
    r.vbase = v ? (const uint64_t *)(v & ~uintptr_t(7)) : pmh_all_valid;
    r.vbit0 = v ? (uint32_t)(v & 7) * 8 : 0;
    r.vstride = v ? 1 : 0;
    return r;
}
__device__ __forceinline__ uint64_t vpos(const RoleCol &c, uint64_t r) { return c.bit0 + (r << c.bshift); }
__device__ __forceinline__ uint64_t bpos(const RoleCol &c, uint64_t r) { return c.vbit0 + r * c.vstride; }
__device__ __forceinline__ uint64_t word_at(const uint64_t *base, uint64_t pos) { return gld<uint64_t>(base, pos >> 6); }
__device__ __forceinline__ uint64_t cut(uint64_t word, uint64_t pos, uint64_t mask) { return (word >> (pos & 63)) & mask; }

// A column as one block of a scan sees it: the word bases advanced to the
// block's first row, so the bit positions of the block's rows (fewer than
// 2^25 of them) are 32-bit and a load is a uniform base plus a 32-bit offset.
struct BlockCol {
    const uint64_t *base, *vbase;
    uint64_t mask;
    uint32_t bit0, vbit0, vstride;
    int32_t bshift, tag;
};
__device__ __forceinline__ BlockCol block_col(const RoleCol &c, uint64_t r0) {
    BlockCol b;
    const uint64_t p = c.bit0 + (r0 << c.bshift), vp = c.vbit0 + r0 * c.vstride;
    b.base = c.base + (p >> 6);
    b.vbase = c.vbase + (vp >> 6);  // the all-valid stand-in stays put (vstride 0)
    b.bit0 = (uint32_t)(p & 63);
    b.vbit0 = (uint32_t)(vp & 63);
    b.vstride = c.vstride;
    b.mask = c.mask;
    b.bshift = c.bshift;
    b.tag = c.tag;
    return b;
}
__device__ __forceinline__ uint32_t vpos(const BlockCol &c, uint32_t r) { return c.bit0 + (r << c.bshift); }
__device__ __forceinline__ uint32_t bpos(const BlockCol &c, uint32_t r) { return c.vbit0 + r * c.vstride; }
__device__ __forceinline__ uint64_t word_at(const uint64_t *base, uint32_t pos) { return gld<uint64_t>(base, pos >> 6); }
// The same load marked non-temporal: a word read once (a gathered key or
// value) should not push out lines that are about to be read again.
__device__ __forceinline__ uint64_t word_at_nt(const uint64_t *base, uint32_t pos) {
    return __builtin_nontemporal_load((const PMH_G uint64_t *)base + (pos >> 6));
}

// Predicate images of U values of one column (raw bits, zero-extended): the
// column's TotalOrd image when flt, else its int64 value.  The tag switch is
// taken once for all U values.
template <int U>
__device__ __forceinline__ void pred_images(int tag, bool flt, const uint64_t (&raw)[U], int64_t (&img)[U]) {
    if (!flt && tag != PMH_F64 && tag != PMH_F32) {
        // integers: the raw bits are zero-extended, so one shift pair with a
        // wave-uniform count sign-extends the signed widths (0 for the rest)
        const int sh = tag == PMH_I32 ? 32 : tag == PMH_I16 ? 48 : tag == PMH_I8 ? 56 : 0;
#pragma unroll
        for (int u = 0; u < U; ++u) img[u] = (int64_t)(raw[u] << sh) >> sh;
        return;
    }
    if (flt) {
        if (tag == PMH_F64) {
#pragma unroll
            for (int u = 0; u < U; ++u) img[u] = tot_key(__longlong_as_double((long long)raw[u]));
        } else if (tag == PMH_F32) {
#pragma unroll
            for (int u = 0; u < U; ++u) img[u] = tot_key((double)__uint_as_float((uint32_t)raw[u]));
        } else {
#pragma unroll
            for (int u = 0; u < U; ++u) img[u] = tot_key((double)raw_to_i64(tag, raw[u]));
        }
        return;
    }
    switch (tag) {
    case PMH_I32:
#pragma unroll
        for (int u = 0; u < U; ++u) img[u] = (int32_t)(uint32_t)raw[u];
        break;
    case PMH_I16:
#pragma unroll
        for (int u = 0; u < U; ++u) img[u] = (int16_t)(uint16_t)raw[u];
        break;
    case PMH_I8:
#pragma unroll
        for (int u = 0; u < U; ++u) img[u] = (int8_t)(uint8_t)raw[u];
        break;
    case PMH_F64:
    case PMH_F32:
#pragma unroll
        for (int u = 0; u < U; ++u) img[u] = raw_to_i64(tag, raw[u]);
        break;
    default:  // I64, U64, U32, U16, U8, Utf8 codes, Bool: the raw bits
#pragma unroll
        for (int u = 0; u < U; ++u) img[u] = (int64_t)raw[u];
        break;
    }
}


}  // namespace pmh
