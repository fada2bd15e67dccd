// pmh_device.h -- device helpers shared by the gfx950 kernels: column loads,
// the stack-free predicate evaluator, key lanes, hashing, the exact f64
// accumulator and the state-merge "sinks" for LDS and global tables.
#pragma once

#include "pmh_internal.h"

namespace pmh {

constexpr uint64_t kEmpty = ~0ull;

// Global-memory accessors.  Column and table pointers reach the kernels
// through descriptors in device memory, so the compiler cannot prove their
// address space and would emit flat instructions, which also count against
// the LDS wait counter and serialise HBM loads behind LDS atomics.  These
// casts pin them to the global address space.
#define PMH_G __attribute__((address_space(1)))
// Read-only launch descriptors: a pointer that went through the constant
// address space lets the compiler treat every field load as invariant (no
// store in the kernel can change it), so fields are loaded once, not again
// after each atomic.
#define PMH_C __attribute__((address_space(4)))
template <class T> __device__ __forceinline__ const T *as_const(const T *p) {
    return (const T *)(const PMH_C T *)p;
}
template <class T> __device__ __forceinline__ T gld(const void *p, uint64_t i) { return ((const PMH_G T *)p)[i]; }
template <class T> __device__ __forceinline__ void gst(void *p, uint64_t i, T v) { ((PMH_G T *)p)[i] = v; }
// A load marked non-temporal: for data read once (spill records), so it does
// not push out lines still in use (a scatter's partially written runs).
template <class T> __device__ __forceinline__ T gld_nt(const void *p, uint64_t i) {
    return __builtin_nontemporal_load((const PMH_G T *)p + i);
}
template <class T> __device__ __forceinline__ PMH_G T *gp(void *p) { return (PMH_G T *)p; }
template <class T> __device__ __forceinline__ const PMH_G T *gp(const void *p) { return (const PMH_G T *)p; }

// Every per-group state array merges with one of these operators.
enum MergeOp : int32_t { MO_ADD32 = 0, MO_ADD64, MO_ADDF64, MO_MIN32, MO_MAX32, MO_MIN64, MO_MAX64, MO_OR32 };

__host__ __device__ inline int mo_size(int op) {
    return (op == MO_ADD32 || op == MO_MIN32 || op == MO_MAX32 || op == MO_OR32) ? 4 : 8;
}
__host__ __device__ inline uint64_t mo_identity(int op) {
    switch (op) {
    case MO_MIN32: return 0x7FFFFFFFull;
    case MO_MAX32: return 0x80000000ull;
    case MO_MIN64: return 0x7FFFFFFFFFFFFFFFull;
    case MO_MAX64: return 0x8000000000000000ull;
    default: return 0;
    }
}

// ---------------------------------------------------------------------------
// column loads (the tag is wave-uniform, so the switch does not diverge)
// ---------------------------------------------------------------------------
__device__ __forceinline__ bool dvalid(const uint8_t *v, uint64_t r) {
    return v ? ((gld<uint8_t>(v, r >> 3) >> (r & 7)) & 1) : true;
}

__device__ __forceinline__ int64_t dload_i64(const DevCol &c, uint64_t r) {
    switch (c.tag) {
    case PMH_I64: case PMH_U64: return gld<int64_t>(c.data, r);
    case PMH_I32: return gld<int32_t>(c.data, r);
    case PMH_U32: case PMH_UTF8: return gld<uint32_t>(c.data, r);
    case PMH_I16: return gld<int16_t>(c.data, r);
    case PMH_U16: return gld<uint16_t>(c.data, r);
    case PMH_I8: return gld<int8_t>(c.data, r);
    case PMH_U8: return gld<uint8_t>(c.data, r);
    case PMH_BOOL: return (gld<uint8_t>(c.data, r >> 3) >> (r & 7)) & 1;
    case PMH_F64: return (int64_t)gld<double>(c.data, r);
    case PMH_F32: return (int64_t)gld<float>(c.data, r);
    default: return 0;
    }
}

__device__ __forceinline__ double dload_f64(const DevCol &c, uint64_t r) {
    if (c.tag == PMH_F64) return gld<double>(c.data, r);
    if (c.tag == PMH_F32) return (double)gld<float>(c.data, r);
    return (double)dload_i64(c, r);
}

// Raw data bits, zero-extended: the key encoder's view of a value
// (crates/polars-metal-kernels/src/groupby.rs:226-292).
__device__ __forceinline__ uint64_t dload_bits(const DevCol &c, uint64_t r) {
    switch (c.tag) {
    case PMH_I64: case PMH_U64: case PMH_F64: return gld<uint64_t>(c.data, r);
    case PMH_I32: case PMH_U32: case PMH_F32: case PMH_UTF8: return gld<uint32_t>(c.data, r);
    case PMH_I16: case PMH_U16: return gld<uint16_t>(c.data, r);
    case PMH_I8: case PMH_U8: return gld<uint8_t>(c.data, r);
    case PMH_BOOL: return (gld<uint8_t>(c.data, r >> 3) >> (r & 7)) & 1;
    default: return 0;
    }
}

// ---------------------------------------------------------------------------
// row accessors: the predicate, key and aggregate code reads column values
// through one of these, either straight from memory (MemRow) or from a
// register cache filled by one batch of loads (CacheRow)
// ---------------------------------------------------------------------------
struct MemRow {
    const QueryDesc *__restrict__ q;
    uint64_t r;
    __device__ __forceinline__ bool valid(int c) const { return dvalid(q->cols[c].valid, r); }
    __device__ __forceinline__ int64_t i64(int c) const { return dload_i64(q->cols[c], r); }
    __device__ __forceinline__ double f64(int c) const { return dload_f64(q->cols[c], r); }
    __device__ __forceinline__ uint64_t bits(int c) const { return dload_bits(q->cols[c], r); }
};

// Raw bits of one row of a column, zero-extended to 64 bits.
__device__ __forceinline__ uint64_t load_raw(const DevCol &c, uint64_t r) { return dload_bits(c, r); }

__device__ __forceinline__ int64_t raw_to_i64(int tag, uint64_t x) {
    switch (tag) {
    case PMH_I32: return (int32_t)(uint32_t)x;
    case PMH_I16: return (int16_t)(uint16_t)x;
    case PMH_I8: return (int8_t)(uint8_t)x;
    case PMH_F64: return (int64_t)__longlong_as_double((long long)x);
    case PMH_F32: return (int64_t)__uint_as_float((uint32_t)x);
    default: return (int64_t)x;  // I64, U64, U32, U16, U8, Utf8 codes, Bool
    }
}

__device__ __forceinline__ double raw_to_f64(int tag, uint64_t x) {
    if (tag == PMH_F64) return __longlong_as_double((long long)x);
    if (tag == PMH_F32) return (double)__uint_as_float((uint32_t)x);
    return (double)raw_to_i64(tag, x);
}

// Register cache of up to four columns for one row.  Column indices are
// wave-uniform; the values live in four scalars (not an array) so the
// compiler cannot turn the selection back into an indexed private-memory
// access.
struct CacheRow {
    const QueryDesc *__restrict__ q;
    uint64_t v0, v1, v2, v3;
    uint32_t vm;  // bit c: column c valid
    __device__ __forceinline__ uint64_t pick(int c) const {
        return c == 0 ? v0 : (c == 1 ? v1 : (c == 2 ? v2 : v3));
    }
    __device__ __forceinline__ void set(int c, uint64_t x) {
        if (c == 0) v0 = x;
        else if (c == 1) v1 = x;
        else if (c == 2) v2 = x;
        else v3 = x;
    }
    __device__ __forceinline__ bool valid(int c) const { return (vm >> c) & 1; }
    __device__ __forceinline__ int64_t i64(int c) const { return raw_to_i64(q->cols[c].tag, pick(c)); }
    __device__ __forceinline__ double f64(int c) const { return raw_to_f64(q->cols[c].tag, pick(c)); }
    __device__ __forceinline__ uint64_t bits(int c) const { return pick(c); }
};

// ---------------------------------------------------------------------------
// predicate: node k's (data, valid) pair lives in bit k of two masks
// ---------------------------------------------------------------------------
__device__ __forceinline__ int tot_cmp(double a, double b) {
    // Polars TotalOrd: NaN == NaN, NaN above every number, -0.0 == +0.0.
    bool an = a != a, bn = b != b;
    if (an | bn) return (int)an - (int)bn;
    return (a < b) ? -1 : ((a > b) ? 1 : 0);
}

__device__ __forceinline__ bool apply_cmp(int op, int c) {
    switch (op) {
    case PMH_EQ: return c == 0;
    case PMH_NE: return c != 0;
    case PMH_LT: return c < 0;
    case PMH_LE: return c <= 0;
    case PMH_GT: return c > 0;
    default: return c >= 0;
    }
}

template <class Row>
__device__ __forceinline__ bool eval_pred(const QueryDesc *__restrict__ q, const Row &row) {
    const int n = q->n_pred;
    if (n == 0) return true;
    uint32_t D = 0, V = 0;
    for (int k = 0; k < n; ++k) {
        const PredNode &nd = q->pred[k];
        bool d = false, v = true;
        switch (nd.kind) {
        case PK_COLBOOL:
            v = row.valid(nd.lhs);
            d = row.bits(nd.lhs) & 1;
            break;
        case PK_CONST: d = nd.cval != 0; break;
        case PK_CMP: {
            bool lv = true, rv = true;
            int c;
            if (nd.is_f64) {
                double a = nd.lf, b = nd.rf;
                if (nd.lkind == OK_COL) { lv = row.valid(nd.lhs); a = row.f64(nd.lhs); }
                if (nd.rkind == OK_COL) { rv = row.valid(nd.rhs); b = row.f64(nd.rhs); }
                c = tot_cmp(a, b);
            } else {
                int64_t a = nd.li, b = nd.ri;
                if (nd.lkind == OK_COL) { lv = row.valid(nd.lhs); a = row.i64(nd.lhs); }
                if (nd.rkind == OK_COL) { rv = row.valid(nd.rhs); b = row.i64(nd.rhs); }
                c = (a < b) ? -1 : ((a > b) ? 1 : 0);
            }
            v = lv && rv;
            d = v && apply_cmp(nd.op, c);
            break;
        }
        default: {  // PK_AND / PK_OR, Kleene (shaders/logical_bool.metal:1-31)
            const bool dom = nd.kind == PK_OR;
            const bool ld = (D >> nd.lhs) & 1, lvv = (V >> nd.lhs) & 1;
            const bool rd = (D >> nd.rhs) & 1, rvv = (V >> nd.rhs) & 1;
            const bool lk = lvv && ld == dom, rk = rvv && rd == dom;
            if (lk || rk) { v = true; d = dom; }
            else if (lvv && rvv) { v = true; d = !dom; }
            else { v = false; d = false; }
            break;
        }
        }
        D |= (uint32_t)d << k;
        V |= (uint32_t)v << k;
    }
    return ((D & V) >> (n - 1)) & 1;
}

// ---------------------------------------------------------------------------
// keys
// ---------------------------------------------------------------------------
struct Key {
    uint64_t lo, hi;
    int special;  // 0: table key; 1: the all-ones lane; 2: null (KM_RAW64)
};

__device__ __forceinline__ void or_shift128(uint64_t &lo, uint64_t &hi, uint64_t v, int sh) {
    if (sh < 64) {
        lo |= v << sh;
        if (sh > 0) hi |= v >> (64 - sh);
    } else {
        hi |= v << (sh - 64);
    }
}

// The order-preserving unsigned image of a key value's raw bits under a
// narrowed field's transform.
__host__ __device__ __forceinline__ uint64_t key_image(int xform, int tag, uint64_t raw) {
    if (xform != KX_SIGNED) return raw;
    const int w = tag_data_bits(tag);
    const uint64_t s = w >= 64 ? raw : (uint64_t)((int64_t)(raw << (64 - w)) >> (64 - w));
    return s ^ (1ull << 63);
}
// The data bits of a field for a valid value's raw bits.
__device__ __forceinline__ uint64_t key_data(const KeyField &f, uint64_t raw) {
    if (f.xform == KX_RAW) return raw;
    const uint64_t d = key_image(f.xform, f.tag, raw) - f.bias;
    return f.bits >= 64 ? d : (d & ((1ull << f.bits) - 1));
}

// key_data, and in over the value's bits beyond the field: nonzero when its
// image lies outside a narrowed field's range [bias, bias + 2^bits).
__device__ __forceinline__ uint64_t key_data_over(const KeyField &f, uint64_t raw, uint64_t &over) {
    if (f.xform == KX_RAW) {
        over = 0;
        return raw;
    }
    const uint64_t d = key_image(f.xform, f.tag, raw) - f.bias;
    const uint64_t m = f.bits >= 64 ? d : (d & ((1ull << f.bits) - 1));
    over = d ^ m;
    return m;
}

template <class Row>
__device__ __forceinline__ Key load_key(const QueryDesc *__restrict__ q, const Row &row) {
    Key k{0, 0, 0};
    const int mode = q->key_mode;
    if (mode == KM_NONE) return k;
    if (mode == KM_RAW64) {
        const int c = q->keys[0].col;
        if (!row.valid(c)) { k.special = 2; return k; }
        k.lo = row.bits(c);
        if (k.lo == kEmpty) k.special = 1;
        return k;
    }
    for (int i = 0; i < q->n_keys; ++i) {
        const KeyField &f = q->keys[i];
        if (!row.valid(f.col)) {
            if (f.null_off >= 0) or_shift128(k.lo, k.hi, 1, f.null_off);
        } else {
            or_shift128(k.lo, k.hi, key_data(f, row.bits(f.col)), f.data_off);
        }
    }
    if (q->allones_possible && k.lo == kEmpty && (mode == KM_LANE64 || k.hi == kEmpty)) k.special = 1;
    return k;
}

__device__ __forceinline__ uint64_t hash_key(uint64_t lo, uint64_t hi) {
    uint64_t x = lo ^ (hi * 0x9E3779B97F4A7C15ull);
    x ^= x >> 31;
    x *= 0xd6e8feb86659fd93ull;
    x ^= x >> 32;
    x *= 0xd6e8feb86659fd93ull;
    x ^= x >> 32;
    return x;
}

// f64 -> int64 whose signed order is the IEEE order of non-NaN doubles
// (-0.0 sorts just below +0.0).
__device__ __forceinline__ int64_t f64_order_key(double x) {
    int64_t b = __double_as_longlong(x);
    return b >= 0 ? b : (b ^ 0x7FFFFFFFFFFFFFFFll);
}
__device__ __forceinline__ double f64_from_order_key(int64_t k) {
    int64_t b = k >= 0 ? k : (k ^ 0x7FFFFFFFFFFFFFFFll);
    return __longlong_as_double(b);
}

// ---------------------------------------------------------------------------
// state sinks
// ---------------------------------------------------------------------------
// Global table: state arrays of nslots entries each.
struct GSink {
    void *const *arr;
    uint64_t nslots;
    template <class T> __device__ __forceinline__ PMH_G T *at(int a, uint64_t s) { return gp<T>(arr[a]) + s; }
    __device__ __forceinline__ void add32(int a, uint64_t s, uint32_t v) {
        __hip_atomic_fetch_add(at<uint32_t>(a, s), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __device__ __forceinline__ void add64(int a, uint64_t s, int64_t v) {
        __hip_atomic_fetch_add(at<uint64_t>(a, s), (uint64_t)v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __device__ __forceinline__ void addf(int a, uint64_t s, double v) {
        __hip_atomic_fetch_add(at<double>(a, s), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __device__ __forceinline__ void min32(int a, uint64_t s, int32_t v) {
        __hip_atomic_fetch_min(at<int32_t>(a, s), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __device__ __forceinline__ void max32(int a, uint64_t s, int32_t v) {
        __hip_atomic_fetch_max(at<int32_t>(a, s), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __device__ __forceinline__ void min64(int a, uint64_t s, int64_t v) {
        __hip_atomic_fetch_min(at<int64_t>(a, s), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __device__ __forceinline__ void max64(int a, uint64_t s, int64_t v) {
        __hip_atomic_fetch_max(at<int64_t>(a, s), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __device__ __forceinline__ void or32(int a, uint64_t s, uint32_t v) {
        __hip_atomic_fetch_or(at<uint32_t>(a, s), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    // limb k of the exact accumulator whose limb 0 is array a (limbs are
    // allocated back to back, nslots apart)
    __device__ __forceinline__ void addlimb(int a, int k, uint64_t s, int64_t v) {
        __hip_atomic_fetch_add(gp<uint64_t>(arr[a]) + (uint64_t)k * nslots + s, (uint64_t)v, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
    }
};

// Block-local table in LDS: array a starts at byte offset off[a].
struct LSink {
    uint8_t *base;
    const uint32_t *off;
    uint32_t nslots;
    template <class T> __device__ __forceinline__ T *p(int a, uint32_t s) { return (T *)(base + off[a]) + s; }
    __device__ __forceinline__ void add32(int a, uint32_t s, uint32_t v) { atomicAdd(p<uint32_t>(a, s), v); }
    __device__ __forceinline__ void add64(int a, uint32_t s, int64_t v) {
        atomicAdd(p<unsigned long long>(a, s), (unsigned long long)v);
    }
    __device__ __forceinline__ void addf(int a, uint32_t s, double v) { unsafeAtomicAdd(p<double>(a, s), v); }
    __device__ __forceinline__ void min32(int a, uint32_t s, int32_t v) { atomicMin(p<int>(a, s), v); }
    __device__ __forceinline__ void max32(int a, uint32_t s, int32_t v) { atomicMax(p<int>(a, s), v); }
    __device__ __forceinline__ void min64(int a, uint32_t s, int64_t v) { atomicMin(p<long long>(a, s), (long long)v); }
    __device__ __forceinline__ void max64(int a, uint32_t s, int64_t v) { atomicMax(p<long long>(a, s), (long long)v); }
    __device__ __forceinline__ void or32(int a, uint32_t s, uint32_t v) { atomicOr(p<uint32_t>(a, s), v); }
    __device__ __forceinline__ void addlimb(int a, int k, uint32_t s, int64_t v) {
        atomicAdd(p<unsigned long long>(a, s) + (uint64_t)k * nslots, (unsigned long long)v);
    }
};

// Exact f64 accumulation.  A finite x = +-m * 2^e whose bits all lie in
// [2^-104, 2^56) is added as a 160-bit fixed-point integer split into five
// 32-bit chunks; each chunk goes to its own int64 limb, so limbs never carry
// and atomic adds commute exactly (2^31 additions of headroom per limb).
// Values outside that window, and Inf/NaN, go to an f64 residual.
template <class S, class Slot>
__device__ __forceinline__ void add_f64_exact(S &s, int a_limb, int a_res, Slot slot, double x) {
    uint64_t b = (uint64_t)__double_as_longlong(x);
    uint32_t E = (uint32_t)(b >> 52) & 0x7FFu;
    uint64_t m = b & 0xFFFFFFFFFFFFFull;
    int e;
    if (E == 0x7FFu) { s.addf(a_res, slot, x); return; }
    if (E == 0) {
        if (m == 0) return;
        e = -1074;
    } else {
        m |= 1ull << 52;
        e = (int)E - 1075;
    }
    if (e < -kFixScale) {
        int tz = __builtin_ctzll(m);
        if (e + tz < -kFixScale) { s.addf(a_res, slot, x); return; }
        m >>= (-kFixScale - e);
        e = -kFixScale;
    }
    int p = e + kFixScale;
    int msb = p + 63 - __builtin_clzll(m);
    if (msb >= 32 * kLimbs) { s.addf(a_res, slot, x); return; }
    int li = p >> 5, sh = p & 31;
    uint64_t lo = m << sh;
    uint64_t hi = sh ? (m >> (64 - sh)) : 0;
    int64_t c0 = (int64_t)(lo & 0xFFFFFFFFull), c1 = (int64_t)(lo >> 32), c2 = (int64_t)hi;
    if (b >> 63) { c0 = -c0; c1 = -c1; c2 = -c2; }
    if (c0) s.addlimb(a_limb, li, slot, c0);
    if (c1) s.addlimb(a_limb, li + 1, slot, c1);
    if (c2) s.addlimb(a_limb, li + 2, slot, c2);
}

// Exact f64 accumulation into LDS limbs (limb k of slot s at limb[k * SL + s])
// with one divergent region only for the rare values outside the fixed-point
// window (and Inf/NaN), which go to the residual.
__device__ __forceinline__ void lds_add_f64(unsigned long long *limb, double *res, uint32_t SL, uint32_t s, double x) {
    const uint64_t b = (uint64_t)__double_as_longlong(x);
    const uint32_t E = (uint32_t)(b >> 52) & 0x7FFu;
    uint64_t m = b & 0xFFFFFFFFFFFFFull;
    if (E) m |= 1ull << 52;
    const int e = E ? (int)E - 1075 : -1074;
    // in the window: every set bit of m * 2^e in [2^-kFixScale, 2^(32*kLimbs-kFixScale))
    const int sh0 = -kFixScale - e;  // > 0: low bits that must be zero
    const bool low_ok = m == 0 || sh0 <= 0 || (sh0 < 64 && (m & ((1ull << sh0) - 1)) == 0);
    const uint64_t mm = sh0 > 0 ? (sh0 < 64 ? m >> sh0 : 0) : m;
    const int p = sh0 > 0 ? 0 : e + kFixScale;
    const int msb = mm ? p + 63 - __builtin_clzll(mm) : 0;
    const bool in_window = E != 0x7FFu && low_ok && msb < 32 * kLimbs;
    if (in_window) {
        if (mm) {
            const int li = p >> 5, sh = p & 31;
            const uint64_t lo = mm << sh;
            const uint64_t hi = sh ? (mm >> (64 - sh)) : 0;
            int64_t c0 = (int64_t)(lo & 0xFFFFFFFFull), c1 = (int64_t)(lo >> 32), c2 = (int64_t)hi;
            if (b >> 63) { c0 = -c0; c1 = -c1; c2 = -c2; }
            // a chunk above the top limb is zero (msb < 32 * kLimbs): clamp
            // its limb index and add the zero
            const int l1 = min(li + 1, kLimbs - 1), l2 = min(li + 2, kLimbs - 1);
            atomicAdd(limb + (uint32_t)li * SL + s, (unsigned long long)c0);
            atomicAdd(limb + (uint32_t)l1 * SL + s, (unsigned long long)c1);
            atomicAdd(limb + (uint32_t)l2 * SL + s, (unsigned long long)c2);
        }
    } else {
        unsafeAtomicAdd(res + s, x);
    }
}

// Correctly rounded double of sum_j limb[j] * 2^(32 j - 104), plus residual.
__device__ __forceinline__ double limbs_to_double(const int64_t *L, double res) {
    // 256-bit two's complement accumulation
    uint64_t w[4] = {0, 0, 0, 0};
    for (int j = 0; j < kLimbs; ++j) {
        int64_t v = L[j];
        int sh = 32 * j;
        uint64_t ext = v < 0 ? ~0ull : 0ull;
        uint64_t t[4];
        // t = sign-extended v << sh
        int ws = sh >> 6, bs = sh & 63;
        uint64_t src[4] = {(uint64_t)v, ext, ext, ext};
        for (int i = 0; i < 4; ++i) {
            int si = i - ws;
            uint64_t cur = si >= 0 ? src[si] : 0;
            uint64_t prev = si - 1 >= 0 ? src[si - 1] : 0;
            t[i] = bs ? ((cur << bs) | (prev >> (64 - bs))) : cur;
        }
        uint64_t c = 0;
        for (int i = 0; i < 4; ++i) {
            uint64_t s1 = w[i] + t[i];
            uint64_t c1 = s1 < w[i];
            uint64_t s2 = s1 + c;
            uint64_t c2 = s2 < s1;
            w[i] = s2;
            c = c1 | c2;
        }
    }
    bool neg = (int64_t)w[3] < 0;
    if (neg) {
        uint64_t c = 1;
        for (int i = 0; i < 4; ++i) {
            uint64_t x = ~w[i] + c;
            c = (c && x == 0) ? 1 : 0;
            w[i] = x;
        }
    }
    int top = -1;
    for (int i = 3; i >= 0; --i)
        if (w[i]) { top = 64 * i + 63 - __builtin_clzll(w[i]); break; }
    double q = 0.0;
    if (top >= 0) {
        // 64-bit window whose bit 63 is the top bit, sticky-jammed into bit 0
        int lowpos = top - 63;
        uint64_t win, sticky = 0;
        if (lowpos <= 0) {
            int s = -lowpos;
            win = w[0] << s;
        } else {
            int wi = lowpos >> 6, bi = lowpos & 63;
            win = w[wi] >> bi;
            if (bi && wi + 1 < 4) win |= w[wi + 1] << (64 - bi);
            for (int i = 0; i < wi; ++i) sticky |= w[i];
            if (bi) sticky |= w[wi] & ((1ull << bi) - 1);
        }
        if (sticky) win |= 1;
        q = ldexp((double)win, lowpos - kFixScale);
    }
    if (neg) q = -q;
    return res != 0.0 ? q + res : q;
}

// Sum of one value per lane over the wave (every lane gets the sum).
__device__ __forceinline__ uint32_t wave_sum(uint32_t x) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
    return x;
}

// Exclusive scan of one value per thread across a block of at most 1024
// threads (whole waves); s_wave holds 16 words.  Every thread of the block
// must call it.
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t x, uint32_t *s_wave) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    uint32_t inc = x;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(inc, o);
        if (lane >= o) inc += y;
    }
    if (lane == 63) s_wave[wave] = inc;
    __syncthreads();
    if (wave == 0) {
        const uint32_t w = lane < nw ? s_wave[lane] : 0;
        uint32_t wi = w;
#pragma unroll
        for (int o = 1; o < 16; o <<= 1) {
            const uint32_t y = __shfl_up(wi, o);
            if (lane >= o) wi += y;
        }
        if (lane < nw) s_wave[lane] = wi - w;
    }
    __syncthreads();
    const uint32_t r = s_wave[wave] + inc - x;
    __syncthreads();  // s_wave may be reused right after
    return r;
}

}  // namespace pmh
