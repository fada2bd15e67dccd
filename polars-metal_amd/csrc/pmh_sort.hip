// pmh_sort.hip -- the sort group-by path: selected rows' key lanes are
// radix-sorted and equal keys become contiguous runs that fold into one group
// each.
//
// Replaces the reference's A2 sort build (crates/polars-metal-kernels/src/
// groupby_build_sort/gpu.rs:24-252, shaders/groupby_sort_u128_lane.metal:
// 52-172, shaders/groupby_segments.metal:18-40; CPU restatement
// groupby_build_sort/reference.rs:3-35), which sorts (u128 lane, row) with a
// stable LSD radix over all 16 bytes of the lane and then marks segment
// starts.  Here:
//   1. so_bits: one pass over the rows computes the OR and AND of every
//      selected row's lane; the bits that differ between rows are the only
//      ones that can order two keys.  The host turns them into at most
//      kMaxRuns runs of lane bits.
//   2. so_encode: each selected row becomes a record (compressed key, row):
//      the runs' bits packed low to high, so compressed order is lane order.
//      For the 98-bit (i32, i64) lane of uniform [0,2^20) x [0,2^40) keys
//      that is 60 bits: eight 8-bit passes instead of thirteen.
//   3. so_hist / so_scan / so_scatter per 8-bit digit: a stable LSD pass with
//      per-tile digit counts, a per-digit scan over tiles, and a tile-local
//      stable rank (wave multisplit ballots + LDS counters) staged in LDS so
//      each digit run leaves as contiguous writes.
//   4. so_bound / so_bscan: group starts (compressed key differs from the
//      previous record), counted per tile and scanned, so a record's group is
//      its number of preceding starts: groups come out in key order, like
//      the reference's segment ids.
//   5. so_reduce: each record gathers its row's values and folds them into
//      its group's state arrays; a wave first reduces each run of equal
//      groups among its 64 lanes (segmented shuffle scans), so one atomic per
//      run and state array reaches memory.
//   6. so_emit / so_pack: keys decoded from the decompressed lane, aggregates
//      finished by the same code as the hash engines.
#include <algorithm>
#include <cstring>

#include "pmh_table.h"

namespace pmh {

namespace {

constexpr int kSortThreads = 1024;
constexpr int kSortRows = 4;                           // rows per thread per tile
constexpr uint32_t kSortTile = kSortThreads * kSortRows;  // 4096
constexpr int kDigitBits = 8;
constexpr int kDigits = 1 << kDigitBits;
constexpr int kMaxRuns = 16;
constexpr int kWaves = kSortThreads / 64;

// Lane bits [pos, pos + len) go to compressed bits [at, at + len); len <= 64.
struct Run {
    uint32_t pos, len, at, pad;
};

enum SortCtr : int32_t { SC_OR_LO = 0, SC_OR_HI, SC_AND_LO, SC_AND_HI, SC_SEL, SC_CURSOR, SC_GROUPS, SC_N = 8 };

struct SLaunch {
    QueryDesc q;
    int32_t n_out, kw;  // kw: 64-bit words of the compressed key (1 or 2)
    int32_t n_runs, pad;
    Run runs[kMaxRuns];
    uint64_t const_lo, const_hi;  // lane bits that do not vary (outside every run)
    uint64_t *ctr;                // SortCtr
    uint64_t n_sel, gcap;
    uint32_t ntiles, pad2;
    uint64_t *ka_lo, *ka_hi, *kb_lo, *kb_hi;
    uint32_t *ra, *rb;
    uint32_t *tile_hist;  // [kDigits][ntiles]
    uint32_t *digit_tot;  // [kDigits]
    uint32_t *tile_cnt;   // [ntiles + 1]: group starts per tile, then exclusive prefix
    uint64_t *gk_lo, *gk_hi;
    void *arr[kMaxArrs];
    OutCol out[kOutMax];
};

// The lane the sort orders by.  Lane modes: the key encoding itself.  One
// 64-bit key column: the reference's encoding of that key (null bit 0, data
// bits 1..64), built here because the hash engines keep such keys raw.
template <class Row>
__device__ __forceinline__ void sort_lane(const QueryDesc *__restrict__ q, const Row &row, uint64_t &lo, uint64_t &hi) {
    if (q->key_mode == KM_RAW64) {
        const int c = q->keys[0].col;
        if (!row.valid(c)) { lo = 1; hi = 0; return; }
        const uint64_t b = row.bits(c);
        lo = b << 1;
        hi = b >> 63;
        return;
    }
    const Key k = load_key(q, row);
    lo = k.lo;
    hi = k.hi;
}

__device__ __forceinline__ uint64_t bits_at(uint64_t lo, uint64_t hi, uint32_t pos, uint32_t len) {
    uint64_t v;
    if (pos == 0) v = lo;
    else if (pos < 64) v = (lo >> pos) | (hi << (64 - pos));
    else v = hi >> (pos - 64);
    return len >= 64 ? v : v & ((1ull << len) - 1);
}
__device__ __forceinline__ void put_bits(uint64_t &lo, uint64_t &hi, uint64_t v, uint32_t at) {
    if (at < 64) {
        lo |= v << at;
        if (at) hi |= v >> (64 - at);
    } else {
        hi |= v << (at - 64);
    }
}

__device__ __forceinline__ void compress(const SLaunch *__restrict__ L, uint64_t lo, uint64_t hi, uint64_t &clo,
                                         uint64_t &chi) {
    clo = chi = 0;
    for (int i = 0; i < L->n_runs; ++i) {
        const Run r = L->runs[i];
        put_bits(clo, chi, bits_at(lo, hi, r.pos, r.len), r.at);
    }
}
__device__ __forceinline__ void decompress(const SLaunch *__restrict__ L, uint64_t clo, uint64_t chi, uint64_t &lo,
                                           uint64_t &hi) {
    lo = L->const_lo;
    hi = L->const_hi;
    for (int i = 0; i < L->n_runs; ++i) {
        const Run r = L->runs[i];
        put_bits(lo, hi, bits_at(clo, chi, r.at, r.len), r.pos);
    }
}

__device__ __forceinline__ uint64_t wave_or(uint64_t x) {
    for (int o = 32; o; o >>= 1) x |= __shfl_xor(x, o);
    return x;
}
__device__ __forceinline__ uint64_t wave_and(uint64_t x) {
    for (int o = 32; o; o >>= 1) x &= __shfl_xor(x, o);
    return x;
}

// ---------------------------------------------------------------------------
// 1. varying lane bits
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void so_bits(const SLaunch *__restrict__ L) {
    const QueryDesc *__restrict__ q = &L->q;
    uint64_t olo = 0, ohi = 0, alo = ~0ull, ahi = ~0ull;
    uint32_t sel = 0;
    const uint64_t n = q->n_rows;
    for (uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n; r += (uint64_t)gridDim.x * blockDim.x) {
        const MemRow row{q, r};
        if (!eval_pred(q, row)) continue;
        uint64_t lo, hi;
        sort_lane(q, row, lo, hi);
        olo |= lo;
        ohi |= hi;
        alo &= lo;
        ahi &= hi;
        ++sel;
    }
    olo = wave_or(olo);
    ohi = wave_or(ohi);
    alo = wave_and(alo);
    ahi = wave_and(ahi);
    uint32_t s = sel;
    for (int o = 32; o; o >>= 1) s += __shfl_xor(s, o);
    if ((threadIdx.x & 63) == 0) {
        atomicOr((unsigned long long *)&L->ctr[SC_OR_LO], olo);
        atomicOr((unsigned long long *)&L->ctr[SC_OR_HI], ohi);
        atomicAnd((unsigned long long *)&L->ctr[SC_AND_LO], alo);
        atomicAnd((unsigned long long *)&L->ctr[SC_AND_HI], ahi);
        atomicAdd((unsigned long long *)&L->ctr[SC_SEL], (unsigned long long)s);
    }
}

__global__ void so_reset(const SLaunch *__restrict__ L) {
    if (threadIdx.x < SC_N) L->ctr[threadIdx.x] = (threadIdx.x == SC_AND_LO || threadIdx.x == SC_AND_HI) ? ~0ull : 0;
}

// ---------------------------------------------------------------------------
// 2. records
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void so_encode(const SLaunch *__restrict__ L) {
    const QueryDesc *__restrict__ q = &L->q;
    const uint64_t n = q->n_rows;
    const int lane = threadIdx.x & 63;
    const bool two = L->kw == 2;
    // uniform trip count so the ballots see whole waves
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    const uint64_t iters = (n + stride - 1) / stride;
    for (uint64_t it = 0; it < iters; ++it) {
        const uint64_t r = it * stride + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
        bool s = false;
        uint64_t clo = 0, chi = 0;
        if (r < n) {
            const MemRow row{q, r};
            if (eval_pred(q, row)) {
                s = true;
                uint64_t lo, hi;
                sort_lane(q, row, lo, hi);
                compress(L, lo, hi, clo, chi);
            }
        }
        const unsigned long long m = __ballot(s);
        if (!m) continue;
        const int leader = __ffsll((long long)m) - 1;
        uint64_t base = 0;
        if (lane == leader)
            base = atomicAdd((unsigned long long *)&L->ctr[SC_CURSOR], (unsigned long long)__popcll(m));
        base = __shfl(base, leader);
        if (!s) continue;
        const uint64_t pos = base + (uint64_t)__popcll(m & ((1ull << lane) - 1));
        gst<uint64_t>(L->ka_lo, pos, clo);
        if (two) gst<uint64_t>(L->ka_hi, pos, chi);
        gst<uint32_t>(L->ra, pos, (uint32_t)r);
    }
}

// ---------------------------------------------------------------------------
// 3. LSD radix passes over the compressed key
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t digit_of(uint64_t lo, uint64_t hi, int shift) {
    const uint64_t v = shift < 64 ? (lo >> shift) | (shift > 56 ? hi << (64 - shift) : 0) : hi >> (shift - 64);
    return (uint32_t)v & (kDigits - 1);
}

__global__ __launch_bounds__(kSortThreads) void so_hist(const SLaunch *__restrict__ L, const uint64_t *klo,
                                                      const uint64_t *khi, int shift) {
    __shared__ uint32_t hist[kWaves][kDigits];
    const int wave = threadIdx.x >> 6;
    const bool two = L->kw == 2;
    for (uint32_t t = blockIdx.x; t < L->ntiles; t += gridDim.x) {
        for (uint32_t i = threadIdx.x; i < kWaves * kDigits; i += kSortThreads) (&hist[0][0])[i] = 0;
        __syncthreads();
        const uint64_t b = (uint64_t)t * kSortTile;
#pragma unroll
        for (int k = 0; k < kSortRows; ++k) {
            const uint64_t i = b + (uint64_t)k * kSortThreads + threadIdx.x;
            if (i < L->n_sel) {
                const uint32_t d = digit_of(gld<uint64_t>(klo, i), two ? gld<uint64_t>(khi, i) : 0, shift);
                atomicAdd(&hist[wave][d], 1u);
            }
        }
        __syncthreads();
        if (threadIdx.x < kDigits) {
            uint32_t s = 0;
            for (int w = 0; w < kWaves; ++w) s += hist[w][threadIdx.x];
            L->tile_hist[(uint64_t)threadIdx.x * L->ntiles + t] = s;
        }
        __syncthreads();
    }
}

// One block per digit: exclusive scan of the digit's tile counts in place,
// the digit's total to digit_tot.
__global__ __launch_bounds__(1024) void so_scan(const SLaunch *__restrict__ L) {
    __shared__ uint32_t s[1024];
    __shared__ uint32_t carry;
    const uint32_t nt = L->ntiles;
    uint32_t *row = L->tile_hist + (uint64_t)blockIdx.x * nt;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (uint32_t t0 = 0; t0 < nt; t0 += 1024) {
        const uint32_t t = t0 + threadIdx.x;
        const uint32_t x = t < nt ? row[t] : 0;
        s[threadIdx.x] = x;
        __syncthreads();
        for (int off = 1; off < 1024; off <<= 1) {
            const uint32_t y = threadIdx.x >= (uint32_t)off ? s[threadIdx.x - off] : 0;
            __syncthreads();
            s[threadIdx.x] += y;
            __syncthreads();
        }
        if (t < nt) row[t] = carry + s[threadIdx.x] - x;
        __syncthreads();
        if (threadIdx.x == 1023) carry += s[1023];
        __syncthreads();
    }
    if (threadIdx.x == 0) L->digit_tot[blockIdx.x] = carry;
}

// Exclusive scan of kDigits values held by threads [0, kDigits) of a block.
__device__ __forceinline__ uint32_t block_excl_digits(uint32_t x, uint32_t *tmp) {
    if (threadIdx.x < kDigits) tmp[threadIdx.x] = x;
    __syncthreads();
    for (int off = 1; off < kDigits; off <<= 1) {
        uint32_t y = 0;
        if (threadIdx.x < kDigits && threadIdx.x >= (uint32_t)off) y = tmp[threadIdx.x - off];
        __syncthreads();
        if (threadIdx.x < kDigits) tmp[threadIdx.x] += y;
        __syncthreads();
    }
    const uint32_t r = threadIdx.x < kDigits ? tmp[threadIdx.x] - x : 0;
    __syncthreads();
    return r;
}

__global__ __launch_bounds__(kSortThreads) void so_scatter(const SLaunch *__restrict__ L, const uint64_t *slo,
                                                         const uint64_t *shi, const uint32_t *srow, uint64_t *dlo,
                                                         uint64_t *dhi, uint32_t *drow, int shift) {
    __shared__ uint32_t wcnt[kWaves][kDigits];
    __shared__ uint32_t run[kDigits];     // rows of each digit ranked so far in this tile
    __shared__ uint32_t dstart[kDigits];  // LDS start of each digit run
    __shared__ uint32_t gbase[kDigits];   // global start of each digit run of this tile
    __shared__ uint32_t dbase[kDigits];   // global start of each digit
    __shared__ uint32_t tmp[kDigits];
    __shared__ uint64_t st_lo[kSortTile];
    __shared__ uint32_t st_row[kSortTile];
    __shared__ uint8_t st_d[kSortTile];
    extern __shared__ __align__(16) uint8_t smem[];
    uint64_t *st_hi = (uint64_t *)smem;  // kSortTile entries when kw == 2
    const bool two = L->kw == 2;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const unsigned long long lt = (1ull << lane) - 1;
    {
        const uint32_t tot = threadIdx.x < kDigits ? L->digit_tot[threadIdx.x] : 0;
        const uint32_t e = block_excl_digits(tot, tmp);
        if (threadIdx.x < kDigits) dbase[threadIdx.x] = e;
    }
    __syncthreads();
    for (uint32_t t = blockIdx.x; t < L->ntiles; t += gridDim.x) {
        const uint64_t b = (uint64_t)t * kSortTile;
        const uint32_t rows = (uint32_t)min<uint64_t>(kSortTile, L->n_sel - b);
        if (threadIdx.x < kDigits) run[threadIdx.x] = 0;
        __syncthreads();
        uint32_t my_d[kSortRows], my_pos[kSortRows];
        uint64_t my_lo[kSortRows], my_hi[kSortRows];
        uint32_t my_row[kSortRows];
#pragma unroll
        for (int k = 0; k < kSortRows; ++k) {
            const uint32_t i = k * kSortThreads + threadIdx.x;
            const bool live = i < rows;
            uint32_t d = 0;
            my_lo[k] = my_hi[k] = 0;
            my_row[k] = 0;
            if (live) {
                my_lo[k] = gld<uint64_t>(slo, b + i);
                if (two) my_hi[k] = gld<uint64_t>(shi, b + i);
                my_row[k] = gld<uint32_t>(srow, b + i);
                d = digit_of(my_lo[k], my_hi[k], shift);
            }
            // multisplit: lanes holding the same digit
            unsigned long long peers = __ballot(live);
#pragma unroll
            for (int bit = 0; bit < kDigitBits; ++bit) {
                const unsigned long long m = __ballot((d >> bit) & 1);
                peers &= ((d >> bit) & 1) ? m : ~m;
            }
            const uint32_t rank = (uint32_t)__popcll(peers & lt);
            for (uint32_t x = threadIdx.x; x < kWaves * kDigits; x += kSortThreads) (&wcnt[0][0])[x] = 0;
            __syncthreads();
            if (live && rank == 0) wcnt[wave][d] = (uint32_t)__popcll(peers);
            __syncthreads();
            if (threadIdx.x < kDigits) {
                const uint32_t dd = threadIdx.x;
                uint32_t acc = run[dd];
                for (int w = 0; w < kWaves; ++w) {
                    const uint32_t x = wcnt[w][dd];
                    wcnt[w][dd] = acc;
                    acc += x;
                }
                run[dd] = acc;
            }
            __syncthreads();
            my_d[k] = d;
            my_pos[k] = live ? wcnt[wave][d] + rank : 0xFFFFFFFFu;
            __syncthreads();
        }
        {
            const uint32_t cnt = threadIdx.x < kDigits ? run[threadIdx.x] : 0;
            const uint32_t e = block_excl_digits(cnt, tmp);
            if (threadIdx.x < kDigits) {
                dstart[threadIdx.x] = e;
                gbase[threadIdx.x] = dbase[threadIdx.x] + L->tile_hist[(uint64_t)threadIdx.x * L->ntiles + t];
            }
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < kSortRows; ++k) {
            if (my_pos[k] == 0xFFFFFFFFu) continue;
            const uint32_t sp = dstart[my_d[k]] + my_pos[k];
            st_lo[sp] = my_lo[k];
            if (two) st_hi[sp] = my_hi[k];
            st_row[sp] = my_row[k];
            st_d[sp] = (uint8_t)my_d[k];
        }
        __syncthreads();
        for (uint32_t sp = threadIdx.x; sp < rows; sp += kSortThreads) {
            const uint32_t d = st_d[sp];
            const uint64_t g = (uint64_t)gbase[d] + (sp - dstart[d]);
            gst<uint64_t>(dlo, g, st_lo[sp]);
            if (two) gst<uint64_t>(dhi, g, st_hi[sp]);
            gst<uint32_t>(drow, g, st_row[sp]);
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// 4. group starts
// ---------------------------------------------------------------------------
__device__ __forceinline__ bool is_start(const uint64_t *klo, const uint64_t *khi, bool two, uint64_t i) {
    if (i == 0) return true;
    return gld<uint64_t>(klo, i) != gld<uint64_t>(klo, i - 1) ||
           (two && gld<uint64_t>(khi, i) != gld<uint64_t>(khi, i - 1));
}

__global__ __launch_bounds__(256) void so_bound(const SLaunch *__restrict__ L, const uint64_t *klo,
                                              const uint64_t *khi) {
    __shared__ uint32_t s_cnt;
    const bool two = L->kw == 2;
    for (uint32_t t = blockIdx.x; t < L->ntiles; t += gridDim.x) {
        if (threadIdx.x == 0) s_cnt = 0;
        __syncthreads();
        const uint64_t b = (uint64_t)t * kSortTile;
        uint32_t c = 0;
        for (uint32_t k = threadIdx.x; k < kSortTile; k += blockDim.x) {
            const uint64_t i = b + k;
            if (i < L->n_sel && is_start(klo, khi, two, i)) ++c;
        }
        for (int o = 32; o; o >>= 1) c += __shfl_xor(c, o);
        if ((threadIdx.x & 63) == 0) atomicAdd(&s_cnt, c);
        __syncthreads();
        if (threadIdx.x == 0) L->tile_cnt[t] = s_cnt;
        __syncthreads();
    }
}

__global__ __launch_bounds__(1024) void so_bscan(const SLaunch *__restrict__ L) {
    __shared__ uint32_t s[1024];
    __shared__ uint32_t carry;
    const uint32_t nt = L->ntiles;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (uint32_t t0 = 0; t0 < nt; t0 += 1024) {
        const uint32_t t = t0 + threadIdx.x;
        const uint32_t x = t < nt ? L->tile_cnt[t] : 0;
        s[threadIdx.x] = x;
        __syncthreads();
        for (int off = 1; off < 1024; off <<= 1) {
            const uint32_t y = threadIdx.x >= (uint32_t)off ? s[threadIdx.x - off] : 0;
            __syncthreads();
            s[threadIdx.x] += y;
            __syncthreads();
        }
        if (t < nt) L->tile_cnt[t] = carry + s[threadIdx.x] - x;
        __syncthreads();
        if (threadIdx.x == 1023) carry += s[1023];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        L->tile_cnt[nt] = carry;
        L->ctr[SC_GROUPS] = carry;
    }
}

__global__ void so_init_state(const SLaunch *__restrict__ L, uint64_t ng) {
    const Layout &lay = L->q.layout;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < ng; i += stride)
        for (int a = 0; a < lay.n_arrs; ++a) {
            const int op = lay.init[a];
            if (mo_size(op) == 4) gst<uint32_t>(L->arr[a], i, (uint32_t)mo_identity(op));
            else gst<uint64_t>(L->arr[a], i, mo_identity(op));
        }
}

// ---------------------------------------------------------------------------
// 5. fold sorted runs into group state
// ---------------------------------------------------------------------------
// Inclusive scan within segments of the wave (head: first lane of a
// segment); the segment's last lane ends with the segment's total.
template <class T, class Op>
__device__ __forceinline__ T seg_scan(T x, int seg_start, int lane, Op op) {
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const T y = __shfl_up(x, off);
        if (lane - off >= seg_start) x = op(x, y);
    }
    return x;
}

__global__ __launch_bounds__(kSortThreads) void so_reduce(const SLaunch *__restrict__ L, const uint64_t *klo,
                                                        const uint64_t *khi, const uint32_t *krow) {
    __shared__ uint32_t wsum[kWaves];
    __shared__ uint32_t carry;
    const QueryDesc *__restrict__ q = &L->q;
    const Layout &lay = q->layout;
    const bool two = L->kw == 2;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    GSink g{L->arr, L->gcap};
    auto add64 = [](int64_t a, int64_t b) { return a + b; };
    auto addu = [](uint32_t a, uint32_t b) { return a + b; };
    for (uint32_t t = blockIdx.x; t < L->ntiles; t += gridDim.x) {
        const uint64_t b = (uint64_t)t * kSortTile;
        if (threadIdx.x == 0) carry = 0;
        __syncthreads();
        for (int k = 0; k < kSortRows; ++k) {
            const uint64_t i = b + (uint64_t)k * kSortThreads + threadIdx.x;
            const bool live = i < L->n_sel;
            const bool start = live && is_start(klo, khi, two, i);
            // group id = starts at indices <= i, minus one
            const unsigned long long sm = __ballot(start);
            if (lane == 0) wsum[wave] = (uint32_t)__popcll(sm);
            __syncthreads();
            uint32_t before = carry;
            for (int w = 0; w < wave; ++w) before += wsum[w];
            __syncthreads();
            if (threadIdx.x == kSortThreads - 1) {
                uint32_t tot = 0;
                for (int w = 0; w < kWaves; ++w) tot += wsum[w];
                carry += tot;
            }
            const uint64_t gid = (uint64_t)L->tile_cnt[t] + before + (uint32_t)__popcll(sm & ((2ull << lane) - 1)) - 1;
            if (start) {
                gst<uint64_t>(L->gk_lo, gid, gld<uint64_t>(klo, i));
                if (two) gst<uint64_t>(L->gk_hi, gid, gld<uint64_t>(khi, i));
            }
            // segments of equal group inside the wave
            const unsigned long long heads = sm | 1ull;
            const int seg_start = 63 - __builtin_clzll(heads & ((2ull << lane) - 1));
            const unsigned long long lv = __ballot(live);
            const bool last = live && (lane == 63 || ((heads >> (lane + 1)) & 1) || !((lv >> (lane + 1)) & 1));
            const uint64_t row = live ? gld<uint32_t>(krow, i) : 0;
            const MemRow R{q, row};
            // len
            {
                const uint32_t c = seg_scan<uint32_t>(live ? 1u : 0u, seg_start, lane, addu);
                if (last) g.add32(lay.a_len, gid, c);
            }
            for (int j = 0; j < q->n_vals; ++j) {
                const ValSpec &vs = q->vals[j];
                const bool ok = live && R.valid(vs.col);
                const uint32_t c = seg_scan<uint32_t>(ok ? 1u : 0u, seg_start, lane, addu);
                if (last && c) g.add32(vs.a_cnt, gid, c);
                if (vs.kind == VK_F) {
                    const double x = ok ? R.f64(vs.col) : 0.0;
                    if (vs.need & NEED_SUM) {
                        // exact: the row's chunks into five limb partials
                        int64_t lb[kLimbs] = {0, 0, 0, 0, 0};
                        bool resid = false;
                        if (ok) {
                            const uint64_t bb = (uint64_t)__double_as_longlong(x);
                            const uint32_t E = (uint32_t)(bb >> 52) & 0x7FFu;
                            uint64_t m = bb & 0xFFFFFFFFFFFFFull;
                            if (E) m |= 1ull << 52;
                            const int e = E ? (int)E - 1075 : -1074;
                            const int sh0 = -kFixScale - e;
                            const bool low_ok = m == 0 || sh0 <= 0 || (sh0 < 64 && (m & ((1ull << sh0) - 1)) == 0);
                            const uint64_t mm = sh0 > 0 ? (sh0 < 64 ? m >> sh0 : 0) : m;
                            const int p = sh0 > 0 ? 0 : e + kFixScale;
                            const int msb = mm ? p + 63 - __builtin_clzll(mm) : 0;
                            if (E != 0x7FFu && low_ok && msb < 32 * kLimbs) {
                                if (mm) {
                                    const int li = p >> 5, sh = p & 31;
                                    const uint64_t lo = mm << sh;
                                    const uint64_t hi = sh ? (mm >> (64 - sh)) : 0;
                                    int64_t c0 = (int64_t)(lo & 0xFFFFFFFFull), c1 = (int64_t)(lo >> 32), c2 = (int64_t)hi;
                                    if (bb >> 63) { c0 = -c0; c1 = -c1; c2 = -c2; }
#pragma unroll
                                    for (int z = 0; z < kLimbs; ++z)
                                        lb[z] = (z == li ? c0 : 0) + (z == li + 1 ? c1 : 0) + (z == li + 2 ? c2 : 0);
                                }
                            } else {
                                resid = true;
                            }
                        }
#pragma unroll
                        for (int z = 0; z < kLimbs; ++z) {
                            const int64_t s = seg_scan<int64_t>(lb[z], seg_start, lane, add64);
                            if (last && s) g.addlimb(vs.a_limb, z, gid, s);
                        }
                        if (resid) g.addf(vs.a_res, gid, x);
                    }
                    if (vs.need & NEED_MINMAX) {
                        const bool nan = ok && x != x;
                        const int64_t kk = f64_order_key(x);
                        const int64_t mn = seg_scan<int64_t>(ok && !nan ? kk : INT64_MAX, seg_start, lane,
                                                             [](int64_t a, int64_t b) { return a < b ? a : b; });
                        const int64_t mx = seg_scan<int64_t>(ok && !nan ? kk : INT64_MIN, seg_start, lane,
                                                             [](int64_t a, int64_t b) { return a > b ? a : b; });
                        const uint32_t nn = seg_scan<uint32_t>(nan ? 1u : 0u, seg_start, lane, addu);
                        if (last) {
                            if (mn != INT64_MAX) g.min64(vs.a_min, gid, mn);
                            if (mx != INT64_MIN) g.max64(vs.a_max, gid, mx);
                            if (nn) g.or32(vs.a_nan, gid, 1u);
                        }
                    }
                } else if (vs.kind == VK_I64) {
                    const int64_t x = ok ? R.i64(vs.col) : 0;
                    if (vs.need & NEED_SUM) {
                        const int64_t s = seg_scan<int64_t>(x, seg_start, lane, add64);
                        if (last && c) g.add64(vs.a_sum, gid, s);
                    }
                    if (vs.need & NEED_MINMAX) {
                        const int64_t mn = seg_scan<int64_t>(ok ? x : INT64_MAX, seg_start, lane,
                                                             [](int64_t a, int64_t b) { return a < b ? a : b; });
                        const int64_t mx = seg_scan<int64_t>(ok ? x : INT64_MIN, seg_start, lane,
                                                             [](int64_t a, int64_t b) { return a > b ? a : b; });
                        if (last && c) { g.min64(vs.a_min, gid, mn); g.max64(vs.a_max, gid, mx); }
                    }
                } else {
                    const int32_t x = ok ? (int32_t)R.i64(vs.col) : 0;
                    if (vs.need & NEED_SUM) {
                        const int64_t s = seg_scan<int64_t>((int64_t)x, seg_start, lane, add64);
                        if (last && c) g.add64(vs.a_sum, gid, s);
                    }
                    if (vs.need & NEED_MINMAX) {
                        const int32_t mn = seg_scan<int32_t>(ok ? x : INT32_MAX, seg_start, lane,
                                                             [](int32_t a, int32_t b) { return a < b ? a : b; });
                        const int32_t mx = seg_scan<int32_t>(ok ? x : INT32_MIN, seg_start, lane,
                                                             [](int32_t a, int32_t b) { return a > b ? a : b; });
                        if (last && c) { g.min32(vs.a_min, gid, mn); g.max32(vs.a_max, gid, mx); }
                    }
                }
            }
            __syncthreads();
        }
    }
}

// ---------------------------------------------------------------------------
// 6. outputs
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void so_emit(const SLaunch *__restrict__ L) {
    const QueryDesc *__restrict__ q = &L->q;
    const uint64_t ng = L->ctr[SC_GROUPS];
    const bool two = L->kw == 2;
    const GReader rd{L->arr};
    for (uint64_t gi = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; gi < ng; gi += (uint64_t)gridDim.x * blockDim.x) {
        uint64_t lo, hi;
        decompress(L, gld<uint64_t>(L->gk_lo, gi), two ? gld<uint64_t>(L->gk_hi, gi) : 0, lo, hi);
        int special = 0;
        if (q->key_mode == KM_RAW64) {
            if (lo & 1) { special = 2; lo = 0; }
            else {
                lo = (lo >> 1) | (hi << 63);
                special = lo == kEmpty ? 1 : 0;
            }
            hi = 0;
        }
        emit_group(q, L->out, L->n_out, rd, gi, lo, hi, special, gi);
    }
}

__global__ void so_pack(const SLaunch *__restrict__ L) {
    const uint64_t ng = L->ctr[SC_GROUPS];
    const uint64_t nbytes = (ng + 7) / 8;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nbytes; i += (uint64_t)gridDim.x * blockDim.x)
        for (int o = 0; o < L->n_out; ++o) {
            const OutCol &oc = L->out[o];
            if (oc.vbytes && oc.valid) {
                uint8_t bt = 0;
                for (int k = 0; k < 8; ++k) {
                    const uint64_t r = i * 8 + k;
                    if (r < ng && oc.vbytes[r]) bt |= (uint8_t)(1u << k);
                }
                ((uint8_t *)oc.valid)[i] = bt;
            }
            if (oc.src == OS_KEY && oc.tag == PMH_BOOL && oc.data_bits) {
                uint8_t bt = 0;
                for (int k = 0; k < 8; ++k) {
                    const uint64_t r = i * 8 + k;
                    if (r < ng && ((const uint8_t *)oc.data)[r]) bt |= (uint8_t)(1u << k);
                }
                oc.data_bits[i] = bt;
            }
        }
}

void launch_check() { PMH_HIP(hipGetLastError()); }

// Runs of varying lane bits (at most kMaxRuns; closest runs merge when
// there are more, taking the constant bits between them along).
void plan_runs(SLaunch &L, uint64_t vlo, uint64_t vhi) {
    std::vector<std::pair<uint32_t, uint32_t>> rs;  // [pos, end)
    auto bit = [&](int b) { return b < 64 ? (vlo >> b) & 1 : (vhi >> (b - 64)) & 1; };
    for (int b = 0; b < 128;) {
        if (!bit(b)) { ++b; continue; }
        int e = b;
        while (e < 128 && bit(e)) ++e;
        rs.push_back({(uint32_t)b, (uint32_t)e});
        b = e;
    }
    while (rs.size() > (size_t)kMaxRuns - 2) {  // room to split runs longer than 64 bits
        size_t best = 0;
        uint32_t gap = ~0u;
        for (size_t i = 0; i + 1 < rs.size(); ++i)
            if (rs[i + 1].first - rs[i].second < gap) { gap = rs[i + 1].first - rs[i].second; best = i; }
        rs[best].second = rs[best + 1].second;
        rs.erase(rs.begin() + (long)best + 1);
    }
    L.n_runs = 0;
    uint32_t at = 0;
    for (auto &r : rs) {
        for (uint32_t p = r.first; p < r.second;) {
            const uint32_t len = std::min<uint32_t>(64, r.second - p);
            L.runs[L.n_runs++] = Run{p, len, at, 0};
            at += len;
            p += len;
        }
    }
    L.kw = at > 64 ? 2 : 1;
}

}  // namespace

pmh_result *run_sort(Prep &P, const GroupByPlan &plan, uint32_t flags) {
    (void)plan;
    Runtime &R = rt();
    const uint64_t n = P.n_rows;
    const int cus = R.prop.multiProcessorCount > 0 ? R.prop.multiProcessorCount : 256;
    auto Lh = std::make_unique<SLaunch>();
    memset(Lh.get(), 0, sizeof(SLaunch));
    SLaunch &L = *Lh;
    L.q = P.q;
    L.n_out = P.n_out;
    L.ctr = (uint64_t *)slab("so:ctr", SC_N * 8);
    SLaunch *dL = (SLaunch *)slab("so:launch", sizeof(SLaunch));
    hipEvent_t ev0, ev1;
    PMH_HIP(hipEventCreate(&ev0));
    PMH_HIP(hipEventCreate(&ev1));
    PMH_HIP(hipEventRecord(ev0, R.stream));

    // ---- 1. varying lane bits and the selected-row count
    PMH_HIP(hipMemcpyAsync(dL, &L, sizeof(SLaunch), hipMemcpyHostToDevice, R.stream));
    hipLaunchKernelGGL(so_reset, dim3(1), dim3(64), 0, R.stream, dL);
    {
        KernelTimer kt("so_bits");
        hipLaunchKernelGGL(so_bits, dim3(cus * 8), dim3(256), 0, R.stream, dL);
        launch_check();
    }
    uint64_t c[SC_N];
    PMH_HIP(hipMemcpyAsync(c, L.ctr, sizeof(c), hipMemcpyDeviceToHost, R.stream));
    PMH_HIP(hipStreamSynchronize(R.stream));
    const uint64_t n_sel = c[SC_SEL];
    L.n_sel = n_sel;
    plan_runs(L, c[SC_OR_LO] ^ c[SC_AND_LO], c[SC_OR_HI] ^ c[SC_AND_HI]);
    uint64_t vlo = 0, vhi = 0;  // bits covered by the runs
    for (int i = 0; i < L.n_runs; ++i)
        for (uint32_t b = L.runs[i].pos; b < L.runs[i].pos + L.runs[i].len; ++b) {
            if (b < 64) vlo |= 1ull << b;
            else vhi |= 1ull << (b - 64);
        }
    L.const_lo = c[SC_AND_LO] & ~vlo;
    L.const_hi = c[SC_AND_HI] & ~vhi;
    uint32_t width = 0;
    for (int i = 0; i < L.n_runs; ++i) width += L.runs[i].len;
    const int passes = (int)((width + kDigitBits - 1) / kDigitBits);

    // ---- buffers
    const uint64_t cap = std::max<uint64_t>(n_sel, 1);
    L.ntiles = (uint32_t)std::max<uint64_t>(1, ceil_div(n_sel, kSortTile));
    L.ka_lo = (uint64_t *)slab("so:ka_lo", cap * 8 + 64);
    L.kb_lo = (uint64_t *)slab("so:kb_lo", cap * 8 + 64);
    L.ka_hi = L.kw == 2 ? (uint64_t *)slab("so:ka_hi", cap * 8 + 64) : nullptr;
    L.kb_hi = L.kw == 2 ? (uint64_t *)slab("so:kb_hi", cap * 8 + 64) : nullptr;
    L.ra = (uint32_t *)slab("so:ra", cap * 4 + 64);
    L.rb = (uint32_t *)slab("so:rb", cap * 4 + 64);
    L.tile_hist = (uint32_t *)slab("so:tile_hist", (uint64_t)kDigits * L.ntiles * 4 + 64);
    L.digit_tot = (uint32_t *)slab("so:digit_tot", kDigits * 4 + 64);
    L.tile_cnt = (uint32_t *)slab("so:tile_cnt", ((uint64_t)L.ntiles + 1) * 4 + 64);
    PMH_HIP(hipMemcpyAsync(dL, &L, sizeof(SLaunch), hipMemcpyHostToDevice, R.stream));

    // ---- 2. records, 3. radix passes
    const int pgrid = cus * 2;
    uint64_t *slo = L.ka_lo, *shi = L.ka_hi, *dlo = L.kb_lo, *dhi = L.kb_hi;
    uint32_t *srow = L.ra, *drow = L.rb;
    if (n_sel) {
        {
            KernelTimer kt("so_encode");
            hipLaunchKernelGGL(so_encode, dim3(cus * 8), dim3(256), 0, R.stream, dL);
            launch_check();
        }
        const size_t dyn = L.kw == 2 ? (size_t)kSortTile * 8 : 0;
        PMH_HIP(hipFuncSetAttribute((const void *)so_scatter, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)std::max<size_t>(dyn, 16)));
        for (int ps = 0; ps < passes; ++ps) {
            const int shift = ps * kDigitBits;
            {
                KernelTimer kt("so_hist");
                hipLaunchKernelGGL(so_hist, dim3(pgrid), dim3(kSortThreads), 0, R.stream, dL, slo, shi, shift);
                launch_check();
            }
            {
                KernelTimer kt("so_scan");
                hipLaunchKernelGGL(so_scan, dim3(kDigits), dim3(1024), 0, R.stream, dL);
                launch_check();
            }
            {
                KernelTimer kt("so_scatter");
                hipLaunchKernelGGL(so_scatter, dim3(pgrid), dim3(kSortThreads), std::max<size_t>(dyn, 16), R.stream,
                                   dL, slo, shi, srow, dlo, dhi, drow, shift);
                launch_check();
            }
            std::swap(slo, dlo);
            std::swap(shi, dhi);
            std::swap(srow, drow);
        }
        // ---- 4. group starts
        {
            KernelTimer kt("so_bound");
            hipLaunchKernelGGL(so_bound, dim3(pgrid), dim3(256), 0, R.stream, dL, slo, shi);
            launch_check();
        }
        hipLaunchKernelGGL(so_bscan, dim3(1), dim3(1024), 0, R.stream, dL);
        launch_check();
    }
    PMH_HIP(hipMemcpyAsync(c, L.ctr, sizeof(c), hipMemcpyDeviceToHost, R.stream));
    PMH_HIP(hipStreamSynchronize(R.stream));
    const uint64_t ng = n_sel ? c[SC_GROUPS] : 0;

    // ---- group state and outputs, sized for ng
    const uint64_t gcap = std::max<uint64_t>(ng, 1);
    L.gcap = gcap;
    L.gk_lo = (uint64_t *)slab("so:gk_lo", gcap * 8 + 64);
    L.gk_hi = (uint64_t *)slab("so:gk_hi", gcap * 8 + 64);
    {
        uint8_t *hb = (uint8_t *)slab("so:state", state_bytes(P, gcap) + 4096);
        size_t o = 0;
        carve_state(P, hb, o, gcap, L.arr);
    }
    for (int oi = 0; oi < P.n_out; ++oi) {
        OutCol &oc = L.out[oi];
        oc = P.outs[oi];
        const size_t esz = std::max(1, tag_size(oc.tag));
        oc.data = slab("so:out" + std::to_string(oi), gcap * esz + 64);
        const bool need_valid = !(oc.src == OS_AGG && (oc.op == AGG_LEN || oc.op == AGG_COUNT));
        oc.vbytes = need_valid ? (uint8_t *)slab("so:outvb" + std::to_string(oi), gcap + 64) : nullptr;
        oc.valid = need_valid ? (uint64_t *)slab("so:outv" + std::to_string(oi), gcap / 8 + 64) : nullptr;
        oc.data_bits = (oc.src == OS_KEY && oc.tag == PMH_BOOL)
                           ? (uint8_t *)slab("so:outbits" + std::to_string(oi), gcap / 8 + 64)
                           : nullptr;
    }
    PMH_HIP(hipMemcpyAsync(dL, &L, sizeof(SLaunch), hipMemcpyHostToDevice, R.stream));
    if (ng) {
        {
            KernelTimer kt("so_init_state");
            hipLaunchKernelGGL(so_init_state, dim3((uint32_t)std::min<uint64_t>(ceil_div(ng, 256), cus * 16)),
                               dim3(256), 0, R.stream, dL, ng);
            launch_check();
        }
        {
            KernelTimer kt("so_reduce");
            hipLaunchKernelGGL(so_reduce, dim3(pgrid), dim3(kSortThreads), 0, R.stream, dL, slo, shi, srow);
            launch_check();
        }
        {
            KernelTimer kt("so_emit");
            hipLaunchKernelGGL(so_emit, dim3((uint32_t)std::min<uint64_t>(ceil_div(ng, 256), cus * 16)), dim3(256), 0,
                               R.stream, dL);
            launch_check();
        }
        hipLaunchKernelGGL(so_pack, dim3(cus * 4), dim3(256), 0, R.stream, dL);
        launch_check();
    }
    PMH_HIP(hipEventRecord(ev1, R.stream));
    PMH_HIP(hipEventSynchronize(ev1));
    float dev_ms = 0;
    PMH_HIP(hipEventElapsedTime(&dev_ms, ev0, ev1));
    PMH_HIP(hipEventDestroy(ev0));
    PMH_HIP(hipEventDestroy(ev1));

    pmh_result *res = new_result(P.n_out);
    res->n_rows = ng;
    res->stats.method = PMH_METHOD_SORT;
    res->stats.n_passes = passes;
    res->stats.initial_capacity = cap;
    res->stats.final_capacity = cap;
    res->stats.n_groups = ng;
    res->stats.n_selected = n_sel;
    res->stats.device_ms = dev_ms;
    const uint64_t vbytes = bitmap_bytes_padded(ng);
    const uint64_t packed = (ng + 7) / 8;
    for (int oi = 0; oi < P.n_out; ++oi) {
        const OutCol &oc = L.out[oi];
        pmh_out_column &col = res->columns[oi];
        set_out_name(col, P.names[oi]);
        col.tag = oc.tag;
        const bool is_bool = oc.src == OS_KEY && oc.tag == PMH_BOOL;
        col.data_len = is_bool ? vbytes : ng * tag_size(oc.tag);
        col.valid_len = vbytes;
        const void *dsrc = is_bool ? (const void *)oc.data_bits : oc.data;
        if (flags & PMH_OUT_DEVICE) {
            col.location = PMH_DEVICE;
            col.data = (void *)dsrc;
            col.valid = (uint8_t *)oc.valid;
        } else {
            col.location = PMH_HOST;
            col.data = malloc(col.data_len ? col.data_len : 1);
            col.valid = (uint8_t *)malloc(vbytes);
            memset(col.valid, 0, vbytes);
            if (is_bool) memset(col.data, 0, col.data_len);
            const uint64_t nb_copy = is_bool ? packed : col.data_len;
            if (nb_copy) PMH_HIP(hipMemcpyAsync(col.data, dsrc, nb_copy, hipMemcpyDeviceToHost, R.stream));
            if (oc.valid) {
                if (packed) PMH_HIP(hipMemcpyAsync(col.valid, oc.valid, packed, hipMemcpyDeviceToHost, R.stream));
            } else {
                memset(col.valid, 0xFF, vbytes);  // counts / lens are never null
            }
        }
    }
    PMH_HIP(hipStreamSynchronize(R.stream));
    collect_kernel_times();
    return res;
}

}  // namespace pmh
