// pmh_utf8.hip -- Utf8 key columns: dictionary encoding and decoding on the
// device.
//
// The reference builds the dictionary on the host, one string at a time
// (build_dict_nullable, crates/polars-metal-buffer/src/dict.rs:40-66, called
// from crates/polars-metal-core/src/udf/groupby/mod.rs:818-936 after it has
// checked every row's offsets and UTF-8), groups the u32 codes like any key,
// and decodes the key column of the result (decode_dict, dict.rs:69-74) into
// its packed wire form (mod.rs:646-670).  The same semantics here:
//
//   * code k is the k-th distinct string in row order (first seen wins), so
//     codes equal the reference's; null rows get code 0 and stay null; ""
//     is a value of its own;
//   * a valid row whose offsets are decreasing or out of the byte buffer, or
//     whose bytes are not well-formed UTF-8, is an error naming the row.
//
// Device plan (rows in parallel):
//   u8_insert   each valid row hashes its string (8-byte words read from
//               aligned loads), checks its UTF-8, and finds or claims its
//               string's slot of an open-addressing table of (hash tag, row)
//               words, comparing strings on a tag match; atomicMin keeps the
//               string's first row per slot
//   scan        rows whose slot's first row is themselves are the distinct
//               strings in row order; an exclusive scan over them numbers
//               them (the codes) and records each one's byte range
//   u8_codes    every row reads its slot's code
// Decoding gathers each output row's string from the recorded byte ranges,
// with a scan over lengths for the output offsets.
#include <algorithm>
#include <cstring>

#include "pmh_device.h"
#include "pmh_internal.h"

struct pmh_utf8_dict {
    uint8_t *bytes = nullptr;      // device copy of the string bytes [offsets[0], offsets[n]), padded
    uint64_t bytes_len = 0;
    uint32_t *codes = nullptr;     // device, one per row
    uint64_t *rep_off = nullptr;   // device, per code: byte offset of the string in `bytes`
    uint32_t *rep_len = nullptr;   // device, per code: its length
    uint8_t *valid = nullptr;      // device bitmap of the rows (row i at bit i), or null: all valid
    uint64_t n_rows = 0, n_distinct = 0;
};

namespace pmh {

namespace {

constexpr uint32_t kNoSlot = 0xFFFFFFFFu;

struct U8In {
    const uint8_t *bytes;  // 8-byte aligned, readable 16 bytes past bytes_len
    int64_t bytes_len;
    const void *offsets;   // device, n + 1 entries
    int32_t large;
    int64_t base;          // offsets[0]
    const uint8_t *valid;  // device bitmap or null
    uint32_t vbit0;
    uint64_t n;
};

__device__ __forceinline__ int64_t u8_off(const U8In &a, uint64_t r) {
    return (a.large ? ((const int64_t *)a.offsets)[r] : (int64_t)((const int32_t *)a.offsets)[r]) - a.base;
}
__device__ __forceinline__ bool u8_valid(const U8In &a, uint64_t r) {
    if (!a.valid) return true;
    const uint64_t k = a.vbit0 + r;
    return (a.valid[k >> 3] >> (k & 7)) & 1u;
}
// 8 bytes starting at byte pos (any alignment) of an 8-byte aligned buffer.
__device__ __forceinline__ uint64_t load8(const uint8_t *p, int64_t pos) {
    const uint64_t *w = (const uint64_t *)p + (pos >> 3);
    const int sh = (int)(pos & 7) * 8;
    const uint64_t a = w[0];
    if (!sh) return a;
    return (a >> sh) | (w[1] << (64 - sh));
}
__device__ __forceinline__ uint64_t keep_low(uint64_t x, int64_t nbytes) {
    return nbytes >= 8 ? x : (nbytes <= 0 ? 0 : x & ((1ull << (nbytes * 8)) - 1));
}

// Well-formed UTF-8 (the Unicode standard's table of byte sequences: no
// overlong forms, no surrogates, nothing above U+10FFFF, nothing cut off).
struct Utf8Check {
    int need = 0;
    uint32_t lo = 0x80, hi = 0xBF;
    bool bad = false;
    __device__ __forceinline__ void byte(uint32_t b) {
        if (need == 0) {
            if (b < 0x80) return;
            lo = 0x80;
            hi = 0xBF;
            if (b >= 0xC2 && b <= 0xDF) need = 1;
            else if (b == 0xE0) { need = 2; lo = 0xA0; }
            else if ((b >= 0xE1 && b <= 0xEC) || b == 0xEE || b == 0xEF) need = 2;
            else if (b == 0xED) { need = 2; hi = 0x9F; }
            else if (b == 0xF0) { need = 3; lo = 0x90; }
            else if (b >= 0xF1 && b <= 0xF3) need = 3;
            else if (b == 0xF4) { need = 3; hi = 0x8F; }
            else bad = true;
            return;
        }
        if (b < lo || b > hi) bad = true;
        --need;
        lo = 0x80;
        hi = 0xBF;
    }
};

__device__ __forceinline__ uint64_t mix64(uint64_t x) {
    x ^= x >> 33;
    x *= 0xff51afd7ed558ccdull;
    x ^= x >> 33;
    x *= 0xc4ceb9fe1a85ec53ull;
    x ^= x >> 33;
    return x;
}

__device__ __forceinline__ bool same_string(const uint8_t *p, int64_t s1, int64_t s2, int64_t len) {
    for (int64_t i = 0; i < len; i += 8)
        if (keep_low(load8(p, s1 + i) ^ load8(p, s2 + i), len - i)) return false;
    return true;
}

__global__ __launch_bounds__(256) void u8_fill(uint64_t *table, uint32_t *minrow, uint64_t cap) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < cap; i += (uint64_t)gridDim.x * blockDim.x) {
        table[i] = 0;
        minrow[i] = kNoSlot;
    }
}

// err: (row << 8) | kind, kind 1 bad offsets, 2 bad UTF-8 (atomicMin keeps the first row)
__global__ __launch_bounds__(256) void u8_insert(U8In a, uint64_t *table, uint32_t *minrow, uint64_t mask,
                                                 uint32_t *slot_of, unsigned long long *err) {
    for (uint64_t r = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; r < a.n; r += (uint64_t)gridDim.x * blockDim.x) {
        slot_of[r] = kNoSlot;
        if (!u8_valid(a, r)) continue;
        const int64_t s = u8_off(a, r), e = u8_off(a, r + 1);
        if (s < 0 || e < s || e > a.bytes_len) {
            atomicMin(err, (unsigned long long)((r << 8) | 1));
            continue;
        }
        const int64_t len = e - s;
        uint64_t h = 0x9E3779B97F4A7C15ull ^ (uint64_t)len;
        Utf8Check chk;
        for (int64_t i = 0; i < len; i += 8) {
            const uint64_t w = keep_low(load8(a.bytes, s + i), len - i);
            h = (h ^ w) * 0x100000001b3ull;
            h ^= h >> 29;
            if (chk.need || (w & 0x8080808080808080ull)) {
                const int m = (int)min<int64_t>(8, len - i);
                for (int j = 0; j < m; ++j) chk.byte((uint32_t)(w >> (8 * j)) & 0xFF);
            }
        }
        if (chk.bad || chk.need) {
            atomicMin(err, (unsigned long long)((r << 8) | 2));
            continue;
        }
        h = mix64(h);
        const uint64_t tag = (h >> 32) | 1;  // never 0: a claimed slot is never empty
        const uint64_t mine = (tag << 32) | (r + 1);
        uint64_t i = h & mask;
        for (;;) {
            uint64_t cur = table[i];
            if (cur == 0) {
                const uint64_t old = atomicCAS((unsigned long long *)&table[i], 0ull, (unsigned long long)mine);
                if (old == 0) break;
                cur = old;
            }
            if ((cur >> 32) == tag) {
                const uint64_t q = (cur & 0xFFFFFFFFull) - 1;
                const int64_t qs = u8_off(a, q), qe = u8_off(a, q + 1);
                if (qe - qs == len && same_string(a.bytes, s, qs, len)) break;
            }
            i = (i + 1) & mask;
        }
        atomicMin(&minrow[i], (uint32_t)r);
        slot_of[r] = (uint32_t)i;
    }
}

// ---------------------------------------------------------------------------
// exclusive scan over a value per index, in three launches: per-block sums,
// a scan of the block sums (one block), per-block rescans that write
// ---------------------------------------------------------------------------
constexpr uint32_t kScanThreads = 1024, kScanPer = 8 * kScanThreads;

template <class F>
__global__ __launch_bounds__(kScanThreads) void scan_sums(F f, uint64_t n, uint64_t *bsum) {
    const uint64_t b0 = (uint64_t)blockIdx.x * kScanPer;
    uint64_t acc = 0;
    for (uint32_t k = threadIdx.x; k < kScanPer; k += kScanThreads)
        if (b0 + k < n) acc += f(b0 + k);
    // string lengths may sum past 32 bits: reduce in 64
    __shared__ unsigned long long s_t;
    if (threadIdx.x == 0) s_t = 0;
    __syncthreads();
    atomicAdd(&s_t, (unsigned long long)acc);
    __syncthreads();
    if (threadIdx.x == 0) bsum[blockIdx.x] = s_t;
}

__global__ __launch_bounds__(kScanThreads) void scan_blocks(uint64_t *bsum, uint64_t nb) {
    __shared__ uint64_t s_carry;
    __shared__ uint64_t s_x[kScanThreads];
    if (threadIdx.x == 0) s_carry = 0;
    __syncthreads();
    for (uint64_t c = 0; c < nb; c += kScanThreads) {
        const uint64_t i = c + threadIdx.x;
        const uint64_t x = i < nb ? bsum[i] : 0;
        s_x[threadIdx.x] = x;
        __syncthreads();
        // Hillis-Steele inclusive scan in LDS (64-bit values)
        for (uint32_t o = 1; o < kScanThreads; o <<= 1) {
            const uint64_t y = threadIdx.x >= o ? s_x[threadIdx.x - o] : 0;
            __syncthreads();
            s_x[threadIdx.x] += y;
            __syncthreads();
        }
        if (i < nb) bsum[i] = s_carry + s_x[threadIdx.x] - x;
        __syncthreads();
        if (threadIdx.x == kScanThreads - 1) s_carry += s_x[threadIdx.x];
        __syncthreads();
    }
    if (threadIdx.x == 0) bsum[nb] = s_carry;
}

template <class F, class W>
__global__ __launch_bounds__(kScanThreads) void scan_apply(F f, W w, uint64_t n, const uint64_t *bsum) {
    __shared__ uint32_t s_w[16];
    const uint64_t b0 = (uint64_t)blockIdx.x * kScanPer;
    uint64_t base = bsum[blockIdx.x];
    for (uint32_t t = 0; t < kScanPer; t += kScanThreads) {
        const uint64_t i = b0 + t + threadIdx.x;
        const uint32_t v = i < n ? f(i) : 0;
        const uint32_t ex = block_excl_scan(v, s_w);
        __shared__ uint32_t s_tile;
        if (threadIdx.x == kScanThreads - 1) s_tile = ex + v;
        if (i < n) w(i, base + ex, v);
        __syncthreads();
        base += s_tile;
        __syncthreads();
    }
}

uint64_t run_scan_blocks(uint64_t n) { return std::max<uint64_t>(1, ceil_div(n, kScanPer)); }

template <class F, class W>
const uint64_t *device_scan(F f, W w, uint64_t n, const char *slab_name) {
    Runtime &R = rt();
    const uint64_t nb = run_scan_blocks(n);
    auto *bsum = (uint64_t *)slab(slab_name, (nb + 1) * 8);
    hipLaunchKernelGGL(scan_sums<F>, dim3((uint32_t)nb), dim3(kScanThreads), 0, R.stream, f, n, bsum);
    PMH_HIP(hipGetLastError());
    hipLaunchKernelGGL(scan_blocks, dim3(1), dim3(kScanThreads), 0, R.stream, bsum, nb);
    PMH_HIP(hipGetLastError());
    hipLaunchKernelGGL((scan_apply<F, W>), dim3((uint32_t)nb), dim3(kScanThreads), 0, R.stream, f, w, n, bsum);
    PMH_HIP(hipGetLastError());
    return bsum + nb;  // the total, on the device once the launches complete
}

// the distinct strings: rows that are their slot's first row
struct RepFlag {
    const uint32_t *slot_of, *minrow;
    __device__ uint32_t operator()(uint64_t r) const {
        const uint32_t s = slot_of[r];
        return s != kNoSlot && minrow[s] == (uint32_t)r;
    }
};
struct RepWrite {
    U8In a;
    const uint32_t *slot_of;
    uint32_t *code_of_slot;
    uint64_t *rep_off;
    uint32_t *rep_len;
    __device__ void operator()(uint64_t r, uint64_t code, uint32_t v) const {
        if (!v) return;
        code_of_slot[slot_of[r]] = (uint32_t)code;
        const int64_t s = u8_off(a, r);
        rep_off[code] = (uint64_t)s;
        rep_len[code] = (uint32_t)(u8_off(a, r + 1) - s);
    }
};

__global__ __launch_bounds__(256) void u8_codes(const uint32_t *slot_of, const uint32_t *code_of_slot, uint64_t n,
                                                uint32_t *codes) {
    for (uint64_t r = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; r < n; r += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t s = slot_of[r];
        codes[r] = s == kNoSlot ? 0u : code_of_slot[s];
    }
}

// decoding: output row i's string length (0 for a null row)
struct OutLen {
    const uint32_t *codes;
    const uint8_t *valid;  // null: all valid
    const uint32_t *rep_len;
    uint64_t n_distinct;
    unsigned long long *err;
    __device__ uint32_t operator()(uint64_t i) const {
        if (valid && !((valid[i >> 3] >> (i & 7)) & 1u)) return 0;
        const uint32_t c = codes[i];
        if (c >= n_distinct) {
            atomicMin(err, (unsigned long long)i);
            return 0;
        }
        return rep_len[c];
    }
};
struct OutOff {
    int32_t *offs;
    __device__ void operator()(uint64_t i, uint64_t prefix, uint32_t) const { offs[i] = (int32_t)prefix; }
};

__global__ __launch_bounds__(256) void u8_gather(const uint32_t *codes, const uint8_t *valid, const int32_t *offs,
                                                 const uint64_t *rep_off, const uint32_t *rep_len, uint64_t n,
                                                 uint64_t n_distinct, const uint8_t *bytes, uint8_t *out) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        if (valid && !((valid[i >> 3] >> (i & 7)) & 1u)) continue;
        const uint32_t c = codes[i];
        if (c >= n_distinct) continue;
        const uint64_t s = rep_off[c];
        const uint32_t len = rep_len[c];
        uint8_t *d = out + offs[i];
        for (uint32_t j = 0; j < len; ++j) d[j] = bytes[s + j];
    }
}

// dst bit i = src bit (bit0 + i); bits from n on and the padding bytes zero
__global__ __launch_bounds__(256) void u8_rebase_bits(uint8_t *dst, const uint8_t *src, uint32_t bit0, uint64_t n,
                                                      uint64_t dst_bytes) {
    for (uint64_t b = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; b < dst_bytes; b += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t out = 0;
        for (int j = 0; j < 8; ++j) {
            const uint64_t i = b * 8 + j;
            if (i >= n) break;
            const uint64_t k = bit0 + i;
            out |= ((src[k >> 3] >> (k & 7)) & 1u) << j;
        }
        dst[b] = (uint8_t)out;
    }
}

uint64_t pow2_up(uint64_t x) {
    uint64_t p = 1;
    while (p < x) p <<= 1;
    return p;
}

uint32_t grid_for(uint64_t n) { return (uint32_t)std::min<uint64_t>(std::max<uint64_t>(1, ceil_div(n, 256)), 8192); }

void *dev_alloc(size_t bytes) {
    void *p = nullptr;
    PMH_HIP(hipMalloc(&p, bytes ? bytes : 1));
    return p;
}

}  // namespace

void utf8_dict_free(pmh_utf8_dict *d) {
    if (!d) return;
    Runtime &R = rt();
    (void)hipStreamSynchronize(R.stream);
    (void)hipFree(d->bytes);
    (void)hipFree(d->codes);
    (void)hipFree(d->rep_off);
    (void)hipFree(d->rep_len);
    (void)hipFree(d->valid);
    delete d;
}

void utf8_dict_codes(const pmh_utf8_dict &d, const uint32_t **codes, const uint8_t **valid, uint64_t *n_rows,
                     uint64_t *n_distinct) {
    if (codes) *codes = d.codes;
    if (valid) *valid = d.valid;
    if (n_rows) *n_rows = d.n_rows;
    if (n_distinct) *n_distinct = d.n_distinct;
}

pmh_utf8_dict *utf8_encode(const pmh_utf8_column &c) {
    Runtime &R = rt();
    if (R.scratch_mode) fail(PMH_E_RUNTIME, "polars_metal: Utf8 keys are not taken by the allocation-free entry");
    const uint64_t n = c.n_rows;
    if (n >= (uint64_t(1) << 31)) fail(PMH_E_VALUE, "polars_metal: Utf8 column has 2^31 rows or more");
    if (c.location != PMH_HOST && c.location != PMH_DEVICE) fail(PMH_E_VALUE, "polars_metal: bad Utf8 column location");
    if (!c.offsets) fail(PMH_E_VALUE, "polars_metal: Utf8 column has no offsets");
    if (c.valid_bit_offset > 7) fail(PMH_E_VALUE, "polars_metal: validity bit offset is 0..7");
    const size_t ow = c.large ? 8 : 4;
    // offsets[0] and offsets[n] bound the bytes to copy
    int64_t first = 0, last = 0;
    if (c.location == PMH_HOST) {
        first = c.large ? ((const int64_t *)c.offsets)[0] : ((const int32_t *)c.offsets)[0];
        last = c.large ? ((const int64_t *)c.offsets)[n] : ((const int32_t *)c.offsets)[n];
    } else {
        int64_t two[2] = {0, 0};
        PMH_HIP(hipMemcpyAsync(&two[0], c.offsets, ow, hipMemcpyDeviceToHost, R.stream));
        PMH_HIP(hipMemcpyAsync(&two[1], (const uint8_t *)c.offsets + n * ow, ow, hipMemcpyDeviceToHost, R.stream));
        PMH_HIP(hipStreamSynchronize(R.stream));
        first = c.large ? two[0] : (int32_t)two[0];
        last = c.large ? two[1] : (int32_t)two[1];
    }
    if (first < 0 || last < first || (uint64_t)last > c.bytes_len)
        fail(PMH_E_VALUE, "polars_metal: Utf8 offsets[0] = " + std::to_string(first) + ", offsets[n] = " +
                              std::to_string(last) + " do not lie in the " + std::to_string(c.bytes_len) +
                              "-byte string buffer");
    if (last > first && !c.bytes) fail(PMH_E_VALUE, "polars_metal: Utf8 column has no string bytes");
    std::unique_ptr<pmh_utf8_dict, void (*)(pmh_utf8_dict *)> d(new pmh_utf8_dict(), utf8_dict_free);
    d->n_rows = n;
    d->bytes_len = (uint64_t)(last - first);
    d->bytes = (uint8_t *)dev_alloc(((d->bytes_len + 7) & ~uint64_t(7)) + 16);
    d->codes = (uint32_t *)dev_alloc(n * 4);
    d->rep_off = (uint64_t *)dev_alloc(n * 8);
    d->rep_len = (uint32_t *)dev_alloc(n * 4);
    // the bytes (and, from the host, offsets and validity) on the device
    const void *doffs = c.offsets;
    const uint8_t *dvalid = c.valid;
    if (c.location == PMH_HOST) {
        upload_host(d->bytes, c.bytes + first, d->bytes_len);
        void *o = slab("u8:offsets", (n + 1) * ow);
        upload_host(o, c.offsets, (n + 1) * ow);
        doffs = o;
        if (c.valid) {
            const size_t vb = (c.valid_bit_offset + n + 7) / 8;
            void *v = slab("u8:valid", vb + 16);
            upload_host(v, c.valid, vb);
            dvalid = (const uint8_t *)v;
        }
    } else if (d->bytes_len) {
        PMH_HIP(hipMemcpyAsync(d->bytes, c.bytes + first, d->bytes_len, hipMemcpyDeviceToDevice, R.stream));
    }
    U8In a{d->bytes, (int64_t)d->bytes_len, doffs, c.large ? 1 : 0, first, dvalid, c.valid_bit_offset, n};
    if (dvalid) {  // the codes column's own bitmap, starting at bit 0
        const uint64_t vb = (n + 7) / 8 + 16;
        d->valid = (uint8_t *)dev_alloc(vb);
        hipLaunchKernelGGL(u8_rebase_bits, dim3(grid_for(vb)), dim3(256), 0, R.stream, d->valid, dvalid,
                           c.valid_bit_offset, n, vb);
        PMH_HIP(hipGetLastError());
    }
    // table of at least twice the rows, so probes stay short at any cardinality
    const uint64_t cap = pow2_up(std::max<uint64_t>(1024, 2 * n));
    auto *table = (uint64_t *)slab("u8:table", cap * 8);
    auto *minrow = (uint32_t *)slab("u8:minrow", cap * 4);
    auto *code_of_slot = (uint32_t *)slab("u8:code_of_slot", cap * 4);
    auto *slot_of = (uint32_t *)slab("u8:slot_of", n * 4 + 16);
    auto *err = (unsigned long long *)slab("u8:err", 64);
    const unsigned long long none = ~0ull;
    PMH_HIP(hipMemcpyAsync(err, &none, 8, hipMemcpyHostToDevice, R.stream));
    hipLaunchKernelGGL(u8_fill, dim3(grid_for(cap)), dim3(256), 0, R.stream, table, minrow, cap);
    PMH_HIP(hipGetLastError());
    if (n) {
        hipLaunchKernelGGL(u8_insert, dim3(grid_for(n)), dim3(256), 0, R.stream, a, table, minrow, cap - 1, slot_of, err);
        PMH_HIP(hipGetLastError());
    }
    const uint64_t *d_distinct = device_scan(RepFlag{slot_of, minrow},
                                             RepWrite{a, slot_of, code_of_slot, d->rep_off, d->rep_len}, n, "u8:bsum");
    if (n) {
        hipLaunchKernelGGL(u8_codes, dim3(grid_for(n)), dim3(256), 0, R.stream, slot_of, code_of_slot, n, d->codes);
        PMH_HIP(hipGetLastError());
    }
    unsigned long long e = none;
    uint64_t distinct = 0;
    PMH_HIP(hipMemcpyAsync(&e, err, 8, hipMemcpyDeviceToHost, R.stream));
    PMH_HIP(hipMemcpyAsync(&distinct, d_distinct, 8, hipMemcpyDeviceToHost, R.stream));
    PMH_HIP(hipStreamSynchronize(R.stream));
    if (e != none) {
        const uint64_t row = e >> 8;
        fail(PMH_E_VALUE, (e & 0xFF) == 1
                              ? "polars_metal: Utf8 row " + std::to_string(row) + " has invalid offsets"
                              : "polars_metal: Utf8 row " + std::to_string(row) + " is not valid UTF-8");
    }
    d->n_distinct = distinct;
    return d.release();
}

std::vector<uint8_t> utf8_decode(const pmh_utf8_dict &d, const uint32_t *codes, const uint8_t *valid, uint64_t n,
                                 int location) {
    Runtime &R = rt();
    if (location != PMH_HOST && location != PMH_DEVICE) fail(PMH_E_VALUE, "polars_metal: bad codes location");
    if (n >= (uint64_t(1) << 31)) fail(PMH_E_VALUE, "polars_metal: 2^31 codes or more");
    const uint32_t *dc = codes;
    const uint8_t *dv = valid;
    if (location == PMH_HOST && n) {
        auto *c = (uint32_t *)slab("u8d:codes", n * 4);
        upload_host(c, codes, n * 4);
        dc = c;
        if (valid) {
            auto *v = (uint8_t *)slab("u8d:valid", (n + 7) / 8 + 16);
            upload_host(v, valid, (n + 7) / 8);
            dv = v;
        }
    }
    auto *err = (unsigned long long *)slab("u8d:err", 64);
    const unsigned long long none = ~0ull;
    PMH_HIP(hipMemcpyAsync(err, &none, 8, hipMemcpyHostToDevice, R.stream));
    const size_t head = 4 + (n + 1) * 4;
    auto *offs = (int32_t *)slab("u8d:offs", (n + 1) * 4 + 16);
    uint64_t total = 0;
    unsigned long long e = none;
    if (n) {
        const uint64_t *d_total = device_scan(OutLen{dc, dv, d.rep_len, d.n_distinct, err}, OutOff{offs}, n, "u8d:bsum");
        PMH_HIP(hipMemcpyAsync(&total, d_total, 8, hipMemcpyDeviceToHost, R.stream));
        PMH_HIP(hipMemcpyAsync(&e, err, 8, hipMemcpyDeviceToHost, R.stream));
    }
    PMH_HIP(hipStreamSynchronize(R.stream));
    if (e != none) fail(PMH_E_VALUE, "polars_metal: code at row " + std::to_string(e) + " is not in the dictionary");
    if (total > (uint64_t)INT32_MAX) fail(PMH_E_VALUE, "polars_metal: decoded strings exceed 2^31 - 1 bytes");
    std::vector<uint8_t> out(head + total);
    const uint32_t nn = (uint32_t)n;
    memcpy(out.data(), &nn, 4);
    if (n) {
        auto *dout = (uint8_t *)slab("u8d:out", total + 16);
        hipLaunchKernelGGL(u8_gather, dim3(grid_for(n)), dim3(256), 0, R.stream, dc, dv, offs, d.rep_off, d.rep_len, n,
                           d.n_distinct, d.bytes, dout);
        PMH_HIP(hipGetLastError());
        download_host(out.data() + 4, offs, n * 4);
        download_host(out.data() + head, dout, total);
        PMH_HIP(hipStreamSynchronize(R.stream));
    }
    const int32_t end = (int32_t)total;
    memcpy(out.data() + 4 + n * 4, &end, 4);
    return out;
}

}  // namespace pmh
