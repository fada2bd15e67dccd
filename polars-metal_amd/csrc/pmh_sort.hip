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
//      each digit run leaves as contiguous writes.  Large queries sort only
//      the key's top bits, most significant digit first where that takes
//      at most two passes (so_msd_*), and finish each bucket those bits
//      define in LDS (3b: so_bucket_fold folds its groups there).
//   4. so_bound / so_count_sum / so_count_scan: group starts (compressed key
//      differs from the previous record), counted per tile and scanned, so a
//      record's group is its number of preceding starts: groups come out in
//      key order, like the reference's segment ids.
//   5. so_reduce: each record gathers its row's values and folds them into
//      its group's state arrays; a wave first reduces each run of equal
//      groups among its 64 lanes (segmented shuffle scans), so one atomic per
//      run and state array reaches memory.
//   6. so_emit / so_pack: keys decoded from the decompressed lane, aggregates
//      finished by the same code as the hash engines.
#include <algorithm>
#include <cstring>

#include "pmh_table.h"
#include "pmh_words.h"

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

enum SortCtr : int32_t {
    SC_OR_LO = 0, SC_OR_HI, SC_AND_LO, SC_AND_HI, SC_SEL, SC_CURSOR, SC_GROUPS, SC_BIG,
    SC_PROBE_D,  // distinct top-bit patterns in the skew probe's sample
    SC_MISS,     // sampled lane bits: a record has a bit outside them
    SC_N = 10
};

// Bucket finish (MSD over the top bits, then LDS sorts): a bucket of at most
// kBucketCap records is sorted in one block's LDS; the radix passes first
// order the records by the key's top tbits bits, which split them into
// buckets of about kBucketMean records.
constexpr uint32_t kBucketCap = 4096;   // keys of one word (2048 for two-word keys)
constexpr uint32_t kBucketMean = 3072;  // (1536 for two-word keys)
__host__ __device__ constexpr uint32_t bucket_cap(bool two) { return two ? kBucketCap / 2 : kBucketCap; }
constexpr int kBucketThreads = 512;

struct SLaunch {
    QueryDesc q;
    int32_t n_out, kw;  // kw: 64-bit words of the compressed key (1 or 2)
    int32_t n_runs, pad;
    Run runs[kMaxRuns];
    uint64_t const_lo, const_hi;  // lane bits that do not vary (outside every run)
    uint64_t vmask_lo, vmask_hi;  // lane bits inside the runs
    int32_t sampled, pad6;        // the runs come from a sample: the encode checks every record
    uint64_t *ctr;                // SortCtr
    uint64_t n_sel, gcap;
    uint32_t ntiles, pad2;
    uint64_t *ka_lo, *ka_hi, *kb_lo, *kb_hi;
    uint32_t *ra, *rb;
    // carry mode: the records hold the value bits of the (at most two)
    // aggregated columns and their validity instead of the row index, so
    // the fold streams them instead of gathering through the row
    int32_t carry, pad5;
    uint64_t *va[2], *vb[2];  // value bits, ping and pong
    uint8_t *fa, *fb;         // bit j: value j valid
    uint32_t *tile_hist;  // [kDigits][ntiles]
    uint32_t *digit_tot;  // [kDigits]
    uint32_t *tile_cnt;   // [ntiles + 1]: group starts per tile, then exclusive prefix
    uint64_t *gk_lo, *gk_hi;
    void *arr[kMaxArrs];
    OutCol out[kOutMax];
    FastScan fs;  // the predicate as intervals, when fast is set
    int32_t fast, pad3;
    // per-block partials of the bits pass (or_lo, or_hi, and_lo, and_hi,
    // selected rows), then each block's first record for the encode pass:
    // both passes map rows to blocks the same way, so no pass needs a
    // global cursor (one contended address would serialise every wave)
    uint64_t *bpart;  // [nblk][5]
    uint64_t *boff;   // [nblk]
    uint32_t nblk, pad4;
};

// Wave partials of the bits pass into the block's entry of bpart.
__device__ __forceinline__ void block_bits_out(const SLaunch *__restrict__ L, uint64_t olo, uint64_t ohi,
                                               uint64_t alo, uint64_t ahi, uint32_t cnt) {
    __shared__ unsigned long long s[5];
    if (threadIdx.x == 0) { s[0] = s[1] = 0; s[2] = s[3] = ~0ull; s[4] = 0; }
    __syncthreads();
    for (int o = 32; o; o >>= 1) {
        olo |= __shfl_xor(olo, o);
        ohi |= __shfl_xor(ohi, o);
        alo &= __shfl_xor(alo, o);
        ahi &= __shfl_xor(ahi, o);
        cnt += __shfl_xor(cnt, o);
    }
    if ((threadIdx.x & 63) == 0) {
        atomicOr(&s[0], olo);
        atomicOr(&s[1], ohi);
        atomicAnd(&s[2], alo);
        atomicAnd(&s[3], ahi);
        atomicAdd(&s[4], (unsigned long long)cnt);
    }
    __syncthreads();
    if (threadIdx.x < 5) gst<uint64_t>(L->bpart, (uint64_t)blockIdx.x * 5 + threadIdx.x, s[threadIdx.x]);
}

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
    block_bits_out(L, olo, ohi, alo, ahi, sel);
}

// Reduces the blocks' partials into ctr and turns the blocks' selected
// counts into their first record (exclusive prefix).
__global__ __launch_bounds__(1024) void so_bits_final(const SLaunch *__restrict__ L) {
    __shared__ unsigned long long s[4];
    __shared__ uint64_t sc[1024];
    __shared__ uint64_t carry;
    if (threadIdx.x == 0) { s[0] = s[1] = 0; s[2] = s[3] = ~0ull; carry = 0; }
    __syncthreads();
    uint64_t olo = 0, ohi = 0, alo = ~0ull, ahi = ~0ull;
    for (uint32_t b = threadIdx.x; b < L->nblk; b += blockDim.x) {
        olo |= L->bpart[b * 5 + 0];
        ohi |= L->bpart[b * 5 + 1];
        alo &= L->bpart[b * 5 + 2];
        ahi &= L->bpart[b * 5 + 3];
    }
    atomicOr(&s[0], olo);
    atomicOr(&s[1], ohi);
    atomicAnd(&s[2], alo);
    atomicAnd(&s[3], ahi);
    for (uint32_t b0 = 0; b0 < L->nblk; b0 += 1024) {
        const uint32_t b = b0 + threadIdx.x;
        const uint64_t x = b < L->nblk ? L->bpart[b * 5 + 4] : 0;
        sc[threadIdx.x] = x;
        __syncthreads();
        for (int off = 1; off < 1024; off <<= 1) {
            const uint64_t y = threadIdx.x >= (uint32_t)off ? sc[threadIdx.x - off] : 0;
            __syncthreads();
            sc[threadIdx.x] += y;
            __syncthreads();
        }
        if (b < L->nblk) L->boff[b] = carry + sc[threadIdx.x] - x;
        __syncthreads();
        if (threadIdx.x == 1023) carry += sc[1023];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        L->ctr[SC_OR_LO] = s[0];
        L->ctr[SC_OR_HI] = s[1];
        L->ctr[SC_AND_LO] = s[2];
        L->ctr[SC_AND_HI] = s[3];
        L->ctr[SC_SEL] = carry;
    }
}

// Varying lane bits from kBitsSample evenly spaced rows of a query without a
// predicate (every row is a record; the encode checks each against them),
// one row per thread, and the encode's per-block first record, which needs
// no count then: the encode's block b takes 1024 rows of each full round of
// nblk * 1024 rows and min(1024, max(0, rest - 1024 b)) of the last.  The
// OR/AND counters start at 0 and all ones (so_reset).
constexpr uint32_t kBitsSample = 65536;
__global__ __launch_bounds__(1024) void so_bits_sample(const SLaunch *__restrict__ L) {
    const QueryDesc *__restrict__ q = &L->q;
    const uint64_t n = q->n_rows;
    const uint32_t S = (uint32_t)min<uint64_t>(n, kBitsSample);
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    uint64_t olo = 0, ohi = 0, alo = ~0ull, ahi = ~0ull;
    if (i < S) {
        const MemRow row{q, (uint64_t)i * n / S};
        sort_lane(q, row, olo, ohi);
        alo = olo;
        ahi = ohi;
    }
    olo = wave_or(olo);
    ohi = wave_or(ohi);
    alo = wave_and(alo);
    ahi = wave_and(ahi);
    if ((threadIdx.x & 63) == 0) {
        atomicOr((unsigned long long *)&L->ctr[SC_OR_LO], olo);
        atomicOr((unsigned long long *)&L->ctr[SC_OR_HI], ohi);
        atomicAnd((unsigned long long *)&L->ctr[SC_AND_LO], alo);
        atomicAnd((unsigned long long *)&L->ctr[SC_AND_HI], ahi);
    }
    if (blockIdx.x == 0) {
        const uint64_t per = (uint64_t)L->nblk * 1024, full = n / per, rest = n % per;
        for (uint32_t b = threadIdx.x; b < L->nblk; b += blockDim.x)
            L->boff[b] = (uint64_t)b * full * 1024 + min<uint64_t>(rest, (uint64_t)b * 1024);
        if (threadIdx.x == 0) L->ctr[SC_SEL] = n;
    }
}

__global__ void so_reset_big(const SLaunch *__restrict__ L) {
    if (threadIdx.x == 0) L->ctr[SC_BIG] = 0;
}

__global__ void so_reset(const SLaunch *__restrict__ L) {
    if (threadIdx.x < SC_N) L->ctr[threadIdx.x] = (threadIdx.x == SC_AND_LO || threadIdx.x == SC_AND_HI) ? ~0ull : 0;
}

// ---------------------------------------------------------------------------
// 1+2, the common shape: an interval predicate (see FastScan) and at most
// four key fields, read as aligned words (pmh_words.h), two rows per lane,
// every load issued before any value is used.  ENCODE false: varying bits
// (as so_bits); true: records (as so_encode).
// ---------------------------------------------------------------------------
template <bool ENCODE>
__global__ __launch_bounds__(512) void so_scan_fast(const SLaunch *__restrict__ L) {
    // 1024 rows per block and round (the encode's per-block first records
    // count the bits pass's blocks): two rows per lane over 512 threads keeps
    // the words in flight within four waves per SIMD
    constexpr int U = 2;
    __shared__ Run s_runs[kMaxRuns];
    __shared__ uint32_t s_cur;
    const QueryDesc *__restrict__ q = &L->q;
    if (threadIdx.x < kMaxRuns) s_runs[threadIdx.x] = L->runs[threadIdx.x];
    if (threadIdx.x == 0) s_cur = 0;
    __syncthreads();
    const uint64_t bbase = ENCODE ? L->boff[blockIdx.x] : 0;
    const int nruns = L->n_runs;
    const int nr = L->fs.nr;
    RoleCol pc[2];
    int pflt[2];
    int64_t plo[2], phi[2];
#pragma unroll
    for (int r = 0; r < 2; ++r) {
        pc[r] = role_col(q, r < nr ? L->fs.rcol[r] : 0);
        pflt[r] = L->fs.rflt[r];
        plo[r] = L->fs.rlo[r];
        phi[r] = L->fs.rhi[r];
    }
    const bool raw = q->key_mode == KM_RAW64;
    const int nk = q->key_mode == KM_NONE ? 0 : q->n_keys;
    RoleCol kc[4];
    int knull[4], kdata[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        kc[i] = role_col(q, i < nk ? q->keys[i].col : 0);
        knull[i] = i < nk ? q->keys[i].null_off : 0;
        kdata[i] = i < nk ? q->keys[i].data_off : 0;
    }
    const bool two = L->kw == 2;
    const bool carry = ENCODE && L->carry;
    const bool all = q->n_pred == 0;  // every row selected
    const int nv = carry ? q->n_vals : 0;
    RoleCol vc[2];
#pragma unroll
    for (int j = 0; j < 2; ++j) vc[j] = role_col(q, q->vals[j < q->n_vals ? j : 0].col);
    const int lane = threadIdx.x & 63;
    const uint64_t n = q->n_rows;
    const uint64_t per = (uint64_t)gridDim.x * blockDim.x * U;
    const uint64_t iters = (n + per - 1) / per;  // uniform: ballots see whole waves
    uint64_t olo = 0, ohi = 0, alo = ~0ull, ahi = ~0ull;
    uint64_t miss = 0;  // ENCODE: lane bits of records outside the runs' and the constant bits
    uint32_t cnt = 0;
    for (uint64_t it = 0; it < iters; ++it) {
        const uint64_t base = it * per + (uint64_t)blockIdx.x * blockDim.x * U + threadIdx.x;
        uint64_t row[U];
        uint32_t sel = 0;
#pragma unroll
        for (int u = 0; u < U; ++u) {
            row[u] = base + (uint64_t)u * blockDim.x;
            sel |= (uint32_t)(row[u] < n) << u;
            row[u] = row[u] < n ? row[u] : n - 1;
        }
        uint64_t pw[2][U], pv[2][U], kw[4][U], kv[4][U], vw[2][U], vv[2][U];
#pragma unroll
        for (int j = 0; j < 2; ++j) {
#pragma unroll
            for (int u = 0; u < U; ++u) vw[j][u] = vv[j][u] = 0;
            if (j >= nv) continue;
#pragma unroll
            for (int u = 0; u < U; ++u) {
                vw[j][u] = word_at(vc[j].base, vpos(vc[j], row[u]));
                vv[j][u] = word_at(vc[j].vbase, bpos(vc[j], row[u]));
            }
        }
#pragma unroll
        for (int r = 0; r < 2; ++r) {
            if (r >= nr) continue;
#pragma unroll
            for (int u = 0; u < U; ++u) {
                pw[r][u] = word_at(pc[r].base, vpos(pc[r], row[u]));
                pv[r][u] = word_at(pc[r].vbase, bpos(pc[r], row[u]));
            }
        }
#pragma unroll
        for (int i = 0; i < 4; ++i) {
#pragma unroll
            for (int u = 0; u < U; ++u) kw[i][u] = kv[i][u] = 0;
            if (i >= nk) continue;
#pragma unroll
            for (int u = 0; u < U; ++u) {
                kw[i][u] = word_at(kc[i].base, vpos(kc[i], row[u]));
                kv[i][u] = word_at(kc[i].vbase, bpos(kc[i], row[u]));
            }
        }
#pragma unroll
        for (int r = 0; r < 2; ++r) {
            if (r >= nr) continue;
            uint64_t v[U];
            uint32_t vb = 0;
#pragma unroll
            for (int u = 0; u < U; ++u) {
                v[u] = cut(pw[r][u], vpos(pc[r], row[u]), pc[r].mask);
                vb |= (uint32_t)cut(pv[r][u], bpos(pc[r], row[u]), 1) << u;
            }
            int64_t img[U];
            pred_images<U>(pc[r].tag, pflt[r] != 0, v, img);
#pragma unroll
            for (int u = 0; u < U; ++u)
                if (img[u] < plo[r] || img[u] > phi[r]) vb &= ~(1u << u);
            sel &= vb;
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            uint64_t lo = 0, hi = 0;
            if (raw) {
                const bool ok = cut(kv[0][u], bpos(kc[0], row[u]), 1);
                lo = ok ? kw[0][u] << 1 : 1;
                hi = ok ? kw[0][u] >> 63 : 0;
            } else {
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    if (i >= nk) continue;
                    if (cut(kv[i][u], bpos(kc[i], row[u]), 1))
                        put_bits(lo, hi, cut(kw[i][u], vpos(kc[i], row[u]), kc[i].mask), kdata[i]);
                    else
                        put_bits(lo, hi, 1, knull[i]);
                }
            }
            const bool s = (sel >> u) & 1;
            if (ENCODE && s) miss |= ((lo ^ L->const_lo) & ~L->vmask_lo) | ((hi ^ L->const_hi) & ~L->vmask_hi);
            if (!ENCODE) {
                if (s) {
                    olo |= lo;
                    ohi |= hi;
                    alo &= lo;
                    ahi &= hi;
                    ++cnt;
                }
                continue;
            }
            uint64_t clo = 0, chi = 0;
            for (int k = 0; k < nruns; ++k) {
                const Run rr = s_runs[k];
                put_bits(clo, chi, bits_at(lo, hi, rr.pos, rr.len), rr.at);
            }
            uint64_t pos;
            if (all) {
                // no predicate: every row is a record, in the block's row order
                if (!s) continue;
                pos = bbase + it * (uint64_t)(blockDim.x * U) + (uint64_t)u * blockDim.x + threadIdx.x;
            } else {
                const unsigned long long m = __ballot(s);
                if (!m) continue;
                const int leader = __ffsll((long long)m) - 1;
                uint64_t b = 0;
                if (lane == leader) b = bbase + atomicAdd(&s_cur, (uint32_t)__popcll(m));
                b = __shfl(b, leader);
                pos = b + (uint64_t)__popcll(m & ((1ull << lane) - 1));
            }
            if (s) {
                gst<uint64_t>(L->ka_lo, pos, clo);
                if (two) gst<uint64_t>(L->ka_hi, pos, chi);
                if (carry) {
                    uint32_t f = 0;
#pragma unroll
                    for (int j = 0; j < 2; ++j) {
                        if (j >= nv) continue;
                        gst<uint64_t>(L->va[j], pos, cut(vw[j][u], vpos(vc[j], row[u]), vc[j].mask));
                        f |= (uint32_t)cut(vv[j][u], bpos(vc[j], row[u]), 1) << j;
                    }
                    gst<uint8_t>(L->fa, pos, (uint8_t)f);
                } else {
                    gst<uint32_t>(L->ra, pos, (uint32_t)row[u]);
                }
            }
        }
    }
    if (!ENCODE) block_bits_out(L, olo, ohi, alo, ahi, cnt);
    // runs planned from a sample: a record with a bit outside them sends
    // the query through the exact pass
    if (ENCODE && L->sampled && __ballot(miss != 0) && lane == 0)
        atomicOr((unsigned long long *)&L->ctr[SC_MISS], 1ull);
}

// ---------------------------------------------------------------------------
// 2. records
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void so_encode(const SLaunch *__restrict__ L) {
    __shared__ uint32_t s_cur;
    if (threadIdx.x == 0) s_cur = 0;
    __syncthreads();
    const uint64_t bbase = L->boff[blockIdx.x];
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
        if (lane == leader) base = bbase + atomicAdd(&s_cur, (uint32_t)__popcll(m));
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
                const uint32_t d = digit_of(gld_nt<uint64_t>(klo, i), two ? gld_nt<uint64_t>(khi, i) : 0, shift);
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

// The payload a record carries beside its key: the row index, or (carry
// mode) the value bits of up to two columns and their validity bits.
struct Payload {
    uint32_t *row;
    uint64_t *v[2];
    uint8_t *f;
};

// CARRY: the records carry values (Payload v, f) instead of rows.
template <bool CARRY>
__global__ __launch_bounds__(kSortThreads) void so_scatter(const SLaunch *__restrict__ L, const uint64_t *slo,
                                                         const uint64_t *shi, const Payload sp_, uint64_t *dlo,
                                                         uint64_t *dhi, const Payload dp_, int shift) {
    // per (row step, wave, digit): the wave's rows of that digit in that
    // step, then (in place) the rows of that digit before them in the tile
    __shared__ uint16_t wcnt[kSortRows][kWaves][kDigits];
    __shared__ uint32_t run[kDigits];     // rows of each digit in this tile
    __shared__ uint32_t dstart[kDigits];  // LDS start of each digit run
    __shared__ uint32_t gbase[kDigits];   // global start of each digit run of this tile
    __shared__ uint32_t dbase[kDigits];   // global start of each digit
    __shared__ uint32_t tmp[kDigits];
    __shared__ uint32_t s_wave[16];
    __shared__ uint64_t st_lo[kSortTile];
    __shared__ uint8_t st_d[kSortTile];
    // dynamic: [hi keys when kw == 2], then rows, or values and validity
    extern __shared__ __align__(16) uint8_t smem[];
    const bool two = L->kw == 2;
    const int nv = CARRY ? L->q.n_vals : 0;
    uint64_t *st_hi = (uint64_t *)smem;
    uint8_t *dyn = smem + (two ? (size_t)kSortTile * 8 : 0);
    uint32_t *st_row = (uint32_t *)dyn;
    uint64_t *st_v0 = (uint64_t *)dyn;
    uint64_t *st_v1 = st_v0 + kSortTile;
    uint8_t *st_f = (uint8_t *)(st_v0 + (size_t)kSortTile * (nv > 1 ? 2 : 1));
    const uint32_t *srow = sp_.row;
    uint32_t *drow = dp_.row;
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
        uint32_t my_d[kSortRows], my_pos[kSortRows];
        uint64_t my_lo[kSortRows], my_hi[kSortRows];
        uint32_t my_row[kSortRows];
        uint64_t my_v0[kSortRows], my_v1[kSortRows];
        uint8_t my_f[kSortRows];
        // every record of the tile loads before the first rank (one wait on
        // memory per tile), non-temporally: each is read once, and the
        // tile's runs are still being merged in L2
#pragma unroll
        for (int k = 0; k < kSortRows; ++k) {
            const uint32_t i = k * kSortThreads + threadIdx.x;
            const uint64_t src = b + (i < rows ? i : 0);  // rows >= 1
            my_lo[k] = gld_nt<uint64_t>(slo, src);
            my_hi[k] = two ? gld_nt<uint64_t>(shi, src) : 0;
            if (CARRY) {
                my_row[k] = 0;
                my_v0[k] = gld_nt<uint64_t>(sp_.v[0], src);
                my_v1[k] = nv > 1 ? gld_nt<uint64_t>(sp_.v[1], src) : 0;
                my_f[k] = gld_nt<uint8_t>(sp_.f, src);
            } else {
                my_row[k] = gld_nt<uint32_t>(srow, src);
                my_v0[k] = my_v1[k] = 0;
                my_f[k] = 0;
            }
        }
        // every step's multisplit first (lanes holding the same digit),
        // then one count per (step, wave, digit) and one scan per digit in
        // (step, wave) order, which is row order: the rank stays stable
        uint32_t my_rank[kSortRows];
        for (uint32_t x = threadIdx.x; x < kSortRows * kWaves * kDigits / 2; x += kSortThreads)
            ((uint32_t *)&wcnt[0][0][0])[x] = 0;
#pragma unroll
        for (int k = 0; k < kSortRows; ++k) {
            const uint32_t i = k * kSortThreads + threadIdx.x;
            const bool live = i < rows;
            const uint32_t d = live ? digit_of(my_lo[k], my_hi[k], shift) : 0;
            unsigned long long peers = __ballot(live);
#pragma unroll
            for (int bit = 0; bit < kDigitBits; ++bit) {
                const unsigned long long m = __ballot((d >> bit) & 1);
                peers &= ((d >> bit) & 1) ? m : ~m;
            }
            my_d[k] = d;
            my_rank[k] = live ? (uint32_t)__popcll(peers & lt) : 0xFFFFFFFFu;
            my_pos[k] = live ? (uint32_t)__popcll(peers) : 0;  // the peers' count, for the leader
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < kSortRows; ++k)
            if (my_rank[k] == 0) wcnt[k][wave][my_d[k]] = (uint16_t)my_pos[k];
        __syncthreads();
        if (threadIdx.x < kDigits) {
            const uint32_t dd = threadIdx.x;
            uint32_t acc = 0;
#pragma unroll
            for (int k = 0; k < kSortRows; ++k)
                for (int w = 0; w < kWaves; ++w) {
                    const uint32_t x = wcnt[k][w][dd];
                    wcnt[k][w][dd] = (uint16_t)acc;
                    acc += x;
                }
            run[dd] = acc;
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < kSortRows; ++k)
            my_pos[k] = my_rank[k] != 0xFFFFFFFFu ? wcnt[k][wave][my_d[k]] + my_rank[k] : 0xFFFFFFFFu;
        {
            const uint32_t cnt = threadIdx.x < kDigits ? run[threadIdx.x] : 0;
            const uint32_t e = block_excl_scan(cnt, s_wave);
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
            if (CARRY) {
                st_v0[sp] = my_v0[k];
                if (nv > 1) st_v1[sp] = my_v1[k];
                st_f[sp] = my_f[k];
            } else {
                st_row[sp] = my_row[k];
            }
            st_d[sp] = (uint8_t)my_d[k];
        }
        __syncthreads();
        for (uint32_t sp = threadIdx.x; sp < rows; sp += kSortThreads) {
            const uint32_t d = st_d[sp];
            const uint64_t g = (uint64_t)gbase[d] + (sp - dstart[d]);
            gst<uint64_t>(dlo, g, st_lo[sp]);
            if (two) gst<uint64_t>(dhi, g, st_hi[sp]);
            if (CARRY) {
                gst<uint64_t>(dp_.v[0], g, st_v0[sp]);
                if (nv > 1) gst<uint64_t>(dp_.v[1], g, st_v1[sp]);
                gst<uint8_t>(dp_.f, g, st_f[sp]);
            } else {
                gst<uint32_t>(drow, g, st_row[sp]);
            }
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
// 3b. bucket finish: after the passes over the top tbits key bits the records
// are grouped by those bits in order (buckets); each bucket is sorted on the
// whole key in LDS and written back, which completes the sort.
// ---------------------------------------------------------------------------
// The key's bits [shift, shift + 64) as a 64-bit value.
__device__ __forceinline__ uint64_t key_bits_from(uint64_t lo, uint64_t hi, int shift) {
    return shift == 0 ? lo : shift < 64 ? (lo >> shift) | (hi << (64 - shift)) : hi >> (shift - 64);
}

// Start of every bucket: bstart[b] = first record whose top bits are >= b,
// by a binary search per bucket over the records, which the passes left in
// bucket order (bstart[nb] = n_sel).
__global__ __launch_bounds__(256) void so_bucket_bounds(const SLaunch *__restrict__ L, const uint64_t *klo,
                                                      const uint64_t *khi, uint64_t *bstart, uint32_t nb,
                                                      int shift) {
    const bool two = L->kw == 2;
    const uint64_t n = L->n_sel;
    for (uint32_t b = blockIdx.x * blockDim.x + threadIdx.x; b <= nb; b += gridDim.x * blockDim.x) {
        uint64_t lo = 0, hi = b == nb ? n : 0;
        if (b < nb) {
            hi = n;
            while (lo < hi) {
                const uint64_t mid = (lo + hi) / 2;
                const uint64_t top = key_bits_from(gld<uint64_t>(klo, mid), two ? gld<uint64_t>(khi, mid) : 0, shift);
                if (top < b) lo = mid + 1;
                else hi = mid;
            }
        }
        bstart[b] = b == nb ? n : lo;
    }
}

// MSD top-bit passes of the bucket finish (unstable: the bucket fold orders
// each bucket itself).  The first pass groups the records by the top
// tbits - 8 key bits, the second groups each of those buckets by the next 8
// bits inside it; bucket b of the fold is then contiguous, in key order, and
// its start comes from the counts.  A tile ranks its records by LDS
// counters and its runs claim their place with one atomic per digit on the
// cursor of (digit, XCD slot): the grid is a multiple of 8 and block k takes
// tiles k, k + grid, ..., so runs that meet in one sub-region are written
// through one L2.  Second-pass tiles lie inside one first-pass bucket.
struct MsdPass {
    int shift;               // the digit: key bits [shift, shift + bits)
    uint32_t nd;             // digits (at most kDigits)
    uint32_t second;         // second pass: tiles inside the buckets below
    uint32_t pad;
    const uint64_t *bst;     // second pass: first-pass bucket starts [buckets + 1]
    const uint32_t *toff;    // second pass: first tile of each bucket [buckets + 1]
    const uint32_t *treg;    // second pass: bucket of each tile
    uint32_t *xtot;          // [(bucket * nd + digit) * 8 + slot]: records
    uint64_t *xcur;          // same layout: next free record of each sub-region
    uint32_t ntiles;         // first pass: tiles of kSortTile records
    uint32_t pad2;
};

__device__ __forceinline__ void msd_tile(const SLaunch *__restrict__ L, const MsdPass &M, uint32_t t, uint32_t &b,
                                         uint64_t &begin, uint32_t &rows) {
    if (M.second) {
        b = M.treg[t];
        begin = M.bst[b] + (uint64_t)(t - M.toff[b]) * kSortTile;
        rows = (uint32_t)min<uint64_t>(kSortTile, M.bst[b + 1] - begin);
    } else {
        b = 0;
        begin = (uint64_t)t * kSortTile;
        rows = (uint32_t)min<uint64_t>(kSortTile, L->n_sel - begin);
    }
}

__device__ __forceinline__ uint32_t msd_ntiles(const MsdPass &M, uint32_t nbk) {
    return M.second ? M.toff[nbk] : M.ntiles;
}

// Records per (bucket, digit, XCD slot).  nbk: first-pass buckets (second
// pass).  The first pass takes tile t on block t mod grid, as the scatter
// does, and each wave counts into its own LDS histogram over all the block's
// tiles.  The second pass only has to credit each tile to the slot of the
// scatter block that takes it ((t mod grid) mod 8), so a block counts a
// contiguous range of tiles (one or two buckets) into per-slot LDS counts
// and flushes them when the bucket changes.
__global__ __launch_bounds__(kSortThreads) void so_msd_hist(const SLaunch *__restrict__ L, const uint64_t *klo,
                                                          const uint64_t *khi, const MsdPass M, uint32_t nbk) {
    __shared__ uint32_t hist[kWaves][kDigits];  // first pass: [wave][digit]; second: [slot][digit]
    const bool two = L->kw == 2;
    const uint32_t tid = threadIdx.x, wave = tid >> 6;
    const uint32_t ntiles = msd_ntiles(M, nbk);
    const uint64_t dmask = M.nd - 1;
    for (uint32_t i = tid; i < kWaves * kDigits; i += kSortThreads) (&hist[0][0])[i] = 0;
    __syncthreads();
    auto count = [&](uint32_t t, uint32_t row) {
        uint32_t b, rows;
        uint64_t begin;
        msd_tile(L, M, t, b, begin, rows);
        uint32_t d[kSortRows];
#pragma unroll
        for (int k = 0; k < kSortRows; ++k) {
            const uint32_t i = k * kSortThreads + tid;
            const uint64_t s = begin + (i < rows ? i : 0);
            d[k] = (uint32_t)(key_bits_from(gld_nt<uint64_t>(klo, s), two ? gld_nt<uint64_t>(khi, s) : 0, M.shift) &
                              dmask);
        }
#pragma unroll
        for (int k = 0; k < kSortRows; ++k)
            if (k * kSortThreads + tid < rows) atomicAdd(&hist[row][d[k]], 1u);
    };
    if (!M.second) {
        for (uint32_t t = blockIdx.x; t < ntiles; t += gridDim.x) count(t, wave);
        __syncthreads();
        if (tid < M.nd) {
            uint32_t c = 0;
            for (int w = 0; w < kWaves; ++w) c += hist[w][tid];
            if (c) atomicAdd(&M.xtot[tid * 8 + (blockIdx.x & 7)], c);
        }
        return;
    }
    const uint32_t per = (ntiles + gridDim.x - 1) / gridDim.x;
    const uint32_t t0 = blockIdx.x * per, t1 = min(ntiles, t0 + per);
    uint32_t cur = 0xFFFFFFFFu;  // the bucket the counts belong to
    auto flush = [&]() {
        __syncthreads();
        for (uint32_t i = tid; i < 8 * kDigits; i += kSortThreads) {
            const uint32_t c = (&hist[0][0])[i];
            if (c) {
                atomicAdd(&M.xtot[((uint64_t)cur * M.nd + (i % kDigits)) * 8 + i / kDigits], c);
                (&hist[0][0])[i] = 0;
            }
        }
        __syncthreads();
    };
    for (uint32_t t = t0; t < t1; ++t) {
        const uint32_t b = M.treg[t];
        if (b != cur) {
            if (cur != 0xFFFFFFFFu) flush();
            cur = b;
        }
        count(t, (t % gridDim.x) & 7);
    }
    if (cur != 0xFFFFFFFFu) flush();
}

// Per 1024 (bucket, digit) pairs, one thread each, the records they hold.
__global__ __launch_bounds__(1024) void so_msd_sum(const uint32_t *xtot, uint32_t np, uint32_t *bsum) {
    __shared__ uint32_t s_wave[16];
    const uint32_t c = blockIdx.x * 1024 + threadIdx.x;
    uint32_t s = 0;
    if (c < np) {
        const uint4 a = *(const uint4 *)(xtot + (uint64_t)c * 8);
        const uint4 b = *(const uint4 *)(xtot + (uint64_t)c * 8 + 4);
        s = a.x + a.y + a.z + a.w + b.x + b.y + b.z + b.w;
    }
    const uint32_t e = block_excl_scan(s, s_wave);
    if (threadIdx.x == 1023) bsum[blockIdx.x] = e + s;
}

// Each block of 1024 pairs: its first record (the blocks before it, by one
// wave), then each pair's start (into starts, with starts[np] the total) and
// its slots' cursors: the exclusive prefix of the counts in order.
__global__ __launch_bounds__(1024) void so_msd_scan(const uint32_t *xtot, uint64_t *xcur, uint32_t np,
                                                  const uint32_t *bsum, uint64_t *starts) {
    __shared__ uint32_t s_wave[16];
    __shared__ uint64_t s_off;
    if (threadIdx.x < 64) {
        uint64_t o = 0;
        for (uint32_t j = threadIdx.x; j < blockIdx.x; j += 64) o += bsum[j];
#pragma unroll
        for (int k = 32; k; k >>= 1) o += __shfl_xor(o, k);
        if (threadIdx.x == 0) s_off = o;
    }
    __syncthreads();
    const uint32_t c = blockIdx.x * 1024 + threadIdx.x;
    uint32_t cnt[8], s = 0;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        cnt[k] = c < np ? xtot[(uint64_t)c * 8 + k] : 0;
        s += cnt[k];
    }
    const uint64_t base = s_off + block_excl_scan(s, s_wave);
    if (c < np) {
        starts[c] = base;
        uint64_t o = base;
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            xcur[(uint64_t)c * 8 + k] = o;
            o += cnt[k];
        }
    }
    if (c + 1 == np) starts[np] = base + s;
}

// The second pass's tiles inside the first pass's nbk buckets (nbk < 1024):
// every block scans the buckets' tile counts, block 0 writes each bucket's
// first tile, and the blocks share the map tile -> bucket and clearing the
// second pass's nclear counts.  nbk = 0 only clears.
__global__ __launch_bounds__(1024) void so_msd_tiles(const uint64_t *bst, uint32_t nbk, uint32_t *toff, uint32_t *treg,
                                                   uint32_t *xtot, uint32_t nclear) {
    __shared__ uint32_t s_wave[16];
    __shared__ uint32_t s_toff[1025];
    const uint32_t gt = blockIdx.x * blockDim.x + threadIdx.x, gn = gridDim.x * blockDim.x;
    for (uint32_t i = gt; i < nclear; i += gn) xtot[i] = 0;
    if (nbk == 0) return;
    const uint32_t b = threadIdx.x;
    const uint32_t nt = b < nbk ? (uint32_t)((bst[b + 1] - bst[b] + kSortTile - 1) / kSortTile) : 0;
    const uint32_t t0 = block_excl_scan(nt, s_wave);
    if (b < nbk) s_toff[b] = t0;
    if (b + 1 == nbk) s_toff[nbk] = t0 + nt;
    __syncthreads();
    if (blockIdx.x == 0 && b <= nbk) toff[b] = s_toff[b];
    const uint32_t total = s_toff[nbk];
    for (uint32_t t = gt; t < total; t += gn) {
        // the last bucket whose first tile is at or before t
        uint32_t lo = 0, hi = nbk;
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) / 2;
            if (s_toff[mid] <= t) lo = mid;
            else hi = mid;
        }
        treg[t] = lo;
    }
}

template <bool CARRY, bool TWO>
__global__ __launch_bounds__(kSortThreads, TWO ? 4 : 8) void so_msd_scatter(const SLaunch *__restrict__ L, const uint64_t *slo,
                                                             const uint64_t *shi, const Payload sp_, uint64_t *dlo,
                                                             uint64_t *dhi, const Payload dp_, const MsdPass M,
                                                             uint32_t nbk) {
    __shared__ uint32_t run[kDigits];     // records of each digit in this tile
    __shared__ uint32_t dstart[kDigits];  // LDS start of each digit run
    __shared__ uint64_t gbase[kDigits];   // global start of each digit run of this tile
    __shared__ uint32_t s_wave[16];
    __shared__ uint64_t st_lo[kSortTile];
    __shared__ uint8_t st_d[kSortTile];
    // dynamic: [hi keys when kw == 2], then rows, or values and validity
    extern __shared__ __align__(16) uint8_t smem[];
    constexpr bool two = TWO;  // two-word keys
    const int nv = CARRY ? L->q.n_vals : 0;
    uint64_t *st_hi = (uint64_t *)smem;
    uint8_t *dyn = smem + (two ? (size_t)kSortTile * 8 : 0);
    uint32_t *st_row = (uint32_t *)dyn;
    uint64_t *st_v0 = (uint64_t *)dyn;
    uint64_t *st_v1 = st_v0 + kSortTile;
    uint8_t *st_f = (uint8_t *)(st_v0 + (size_t)kSortTile * (nv > 1 ? 2 : 1));
    const uint32_t tid = threadIdx.x, slot = blockIdx.x & 7;
    const uint32_t ntiles = msd_ntiles(M, nbk);
    const uint64_t dmask = M.nd - 1;
    for (uint32_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        uint32_t b, rows;
        uint64_t begin;
        msd_tile(L, M, t, b, begin, rows);
        if (tid < kDigits) run[tid] = 0;
        __syncthreads();
        // every record of the tile loads before the first rank
        uint64_t my_lo[kSortRows], my_hi[kSortRows], my_v0[kSortRows], my_v1[kSortRows];
        uint32_t my_row[kSortRows], my_rk[kSortRows];
        uint8_t my_f[kSortRows];
#pragma unroll
        for (int k = 0; k < kSortRows; ++k) {
            const uint32_t i = k * kSortThreads + tid;
            const uint64_t src = begin + (i < rows ? i : 0);
            my_lo[k] = gld_nt<uint64_t>(slo, src);
            my_hi[k] = two ? gld_nt<uint64_t>(shi, src) : 0;
            if (CARRY) {
                my_row[k] = 0;
                my_v0[k] = gld_nt<uint64_t>(sp_.v[0], src);
                my_v1[k] = nv > 1 ? gld_nt<uint64_t>(sp_.v[1], src) : 0;
                my_f[k] = gld_nt<uint8_t>(sp_.f, src);
            } else {
                my_row[k] = gld_nt<uint32_t>(sp_.row, src);
                my_v0[k] = my_v1[k] = 0;
                my_f[k] = 0;
            }
        }
#pragma unroll
        for (int k = 0; k < kSortRows; ++k) {
            const uint32_t d = (uint32_t)(key_bits_from(my_lo[k], my_hi[k], M.shift) & dmask);
            my_rk[k] = k * kSortThreads + tid < rows ? atomicAdd(&run[d], 1u) : 0xFFFFFFFFu;
        }
        __syncthreads();
        {
            const uint32_t c = tid < kDigits ? run[tid] : 0;
            const uint32_t e = block_excl_scan(c, s_wave);
            if (tid < kDigits) {
                dstart[tid] = e;
                gbase[tid] = c ? atomicAdd((unsigned long long *)&M.xcur[((uint64_t)b * M.nd + tid) * 8 + slot],
                                           (unsigned long long)c)
                               : 0;
            }
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < kSortRows; ++k) {
            if (my_rk[k] == 0xFFFFFFFFu) continue;
            const uint32_t d = (uint32_t)(key_bits_from(my_lo[k], my_hi[k], M.shift) & dmask);
            const uint32_t sp = dstart[d] + my_rk[k];
            st_lo[sp] = my_lo[k];
            if (two) st_hi[sp] = my_hi[k];
            if (CARRY) {
                st_v0[sp] = my_v0[k];
                if (nv > 1) st_v1[sp] = my_v1[k];
                st_f[sp] = my_f[k];
            } else {
                st_row[sp] = my_row[k];
            }
            st_d[sp] = (uint8_t)d;
        }
        __syncthreads();
        for (uint32_t sp = tid; sp < rows; sp += kSortThreads) {
            const uint32_t d = st_d[sp];
            const uint64_t g = gbase[d] + (sp - dstart[d]);
            gst<uint64_t>(dlo, g, st_lo[sp]);
            if (two) gst<uint64_t>(dhi, g, st_hi[sp]);
            if (CARRY) {
                gst<uint64_t>(dp_.v[0], g, st_v0[sp]);
                if (nv > 1) gst<uint64_t>(dp_.v[1], g, st_v1[sp]);
                gst<uint8_t>(dp_.f, g, st_f[sp]);
            } else {
                gst<uint32_t>(dp_.row, g, st_row[sp]);
            }
        }
        __syncthreads();
    }
}

// Skew probe before the top-bit passes: the largest count of one bucket
// among up to kProbeSample evenly spaced records (any order: the records
// are still in row order) into ctr[SC_BIG], and the number of distinct
// buckets they fall in into ctr[SC_PROBE_D].  A bucket far above the LDS
// capacity shows as a large count; buckets near it may not (they still
// fall back after the bucket kernel flags them).
constexpr uint32_t kProbeSample = 8192;
__global__ __launch_bounds__(1024) void so_bucket_probe(const SLaunch *__restrict__ L, const uint64_t *klo,
                                                      const uint64_t *khi, int shift) {
    constexpr uint32_t kSlots = 2 * kProbeSample;
    __shared__ uint32_t s_id[kSlots], s_cnt[kSlots];
    __shared__ uint32_t s_max, s_distinct;
    const bool two = L->kw == 2;
    const uint64_t n = L->n_sel;
    const uint32_t S = (uint32_t)min<uint64_t>(n, kProbeSample);
    for (uint32_t i = threadIdx.x; i < kSlots; i += blockDim.x) { s_id[i] = 0; s_cnt[i] = 0; }
    if (threadIdx.x == 0) s_max = s_distinct = 0;
    __syncthreads();
    uint32_t mx = 0;
    for (uint32_t i = threadIdx.x; i < S; i += blockDim.x) {
        const uint64_t r = (uint64_t)i * (n / S) + ((uint64_t)i * (n % S)) / S;
        const uint64_t top = key_bits_from(gld<uint64_t>(klo, r), two ? gld<uint64_t>(khi, r) : 0, shift);
        const uint32_t id = (uint32_t)top + 1u;  // buckets are fewer than 2^32 - 1
        uint32_t h = (uint32_t)(((uint64_t)id * 0x9E3779B97F4A7C15ull) >> 49) & (kSlots - 1);
        for (uint32_t p = 0; p < kSlots; ++p) {  // at most S < kSlots ids: a free slot exists
            const uint32_t cur = atomicCAS(&s_id[h], 0u, id);
            if (cur == 0u || cur == id) {
                if (cur == 0u) atomicAdd(&s_distinct, 1u);
                mx = max(mx, atomicAdd(&s_cnt[h], 1u) + 1u);
                break;
            }
            h = (h + 1) & (kSlots - 1);
        }
    }
    atomicMax(&s_max, mx);
    __syncthreads();
    if (threadIdx.x == 0) {
        L->ctr[SC_BIG] = s_max;
        L->ctr[SC_PROBE_D] = s_distinct;
    }
}

__device__ __forceinline__ bool rec_less(uint64_t ahi, uint64_t alo, uint32_t arow, uint64_t bhi, uint64_t blo,
                                         uint32_t brow) {
    if (ahi != bhi) return ahi < bhi;
    if (alo != blo) return alo < blo;
    return arow < brow;
}

// Bitonic sort of m (a power of two) records in LDS on (key, row).
template <bool TWO>
__device__ __forceinline__ void lds_bitonic(uint64_t *s_lo, uint64_t *s_hi, uint32_t *s_row, uint32_t m) {
    for (uint32_t k = 2; k <= m; k <<= 1) {
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            for (uint32_t t = threadIdx.x; t < m / 2; t += blockDim.x) {
                const uint32_t a = 2 * j * (t / j) + (t % j), c = a + j;
                const bool up = (a & k) == 0;
                const uint64_t alo = s_lo[a], clo = s_lo[c];
                const uint64_t ahi = TWO ? s_hi[a] : 0, chi = TWO ? s_hi[c] : 0;
                const uint32_t arow = s_row[a], crow = s_row[c];
                if (rec_less(chi, clo, crow, ahi, alo, arow) == up) {
                    s_lo[a] = clo;
                    s_lo[c] = alo;
                    if (TWO) { s_hi[a] = chi; s_hi[c] = ahi; }
                    s_row[a] = crow;
                    s_row[c] = arow;
                }
            }
            __syncthreads();
        }
    }
}

constexpr int kBinBits = 12;
constexpr uint32_t kBins = 1u << kBinBits;
constexpr uint32_t kBinSortMax = 64;

// LDS of the bucket kernels: the bucket's records grouped by bin (keys,
// [hi], rows), the bin counters, and the sorted order (place -> record).
// Small enough (72 KB for one-word keys) for two blocks per CU, so one
// block's loads overlap the other's LDS work.
template <bool TWO>
struct BucketBufs {
    uint64_t *lo, *hi;
    uint32_t *row;
    uint32_t *bin;   // kBins: counts, then starts
    uint16_t *perm;  // sorted place -> record, when the bins were ranked
};
template <bool TWO>
__device__ __forceinline__ BucketBufs<TWO> bucket_bufs(uint8_t *smem) {
    constexpr uint32_t cap = bucket_cap(TWO);
    BucketBufs<TWO> B;
    B.lo = (uint64_t *)smem;
    B.hi = B.lo + cap;
    B.row = (uint32_t *)(TWO ? B.hi + cap : B.hi);
    B.bin = B.row + cap;
    B.perm = (uint16_t *)(B.bin + kBins);
    return B;
}
__host__ __device__ constexpr size_t bucket_lds_bytes(bool two) {
    return (size_t)bucket_cap(two) * ((two ? 16 : 8) + 4 + 2) + (size_t)kBins * 4;
}

// Sorts one bucket, the n (1..cap) records [beg, beg + n) (they share the
// key's top bits), on (key, row) in LDS: a counting sort on the next
// bin_bits key bits places the records by bin, then every record finds its
// place inside its bin by counting the bin's records that order before it
// (equal keys always share a bin) and writes itself into B.perm at that
// place.  A bin of more than kBinSortMax records sends the bucket to a
// bitonic sort of the records themselves instead.  CARRY records sort on
// (key, place in the bucket) and take their values from that place later;
// row records sort on (key, row).  Returns true when the sorted order is
// B.perm, false when the records themselves are sorted; the block is
// synchronised on return.
template <bool TWO, bool CARRY>
__device__ __forceinline__ bool bucket_sort_lds(const uint64_t *klo, const uint64_t *khi, const Payload &sp_,
                                                uint64_t beg, uint32_t n, const BucketBufs<TWO> &B, uint32_t nbins,
                                                int bin_shift, uint32_t *s_wave, uint32_t *s_crowded) {
    constexpr uint32_t cap = bucket_cap(TWO);
    constexpr uint32_t kPer = cap / kBucketThreads;
    const uint32_t tid = threadIdx.x;
    const uint64_t bmask = nbins - 1;
    for (uint32_t i = tid; i < nbins; i += kBucketThreads) B.bin[i] = 0;
    if (tid == 0) *s_crowded = 0;
    // every load is issued before the first is used
    uint64_t lo[kPer], hi[kPer];
    uint32_t row[kPer], br[kPer];  // br: bin << 16 | rank in the bin, ~0u: no record
#pragma unroll
    for (uint32_t k = 0; k < kPer; ++k) {
        const uint32_t i = k * kBucketThreads + tid;
        lo[k] = hi[k] = 0;
        row[k] = 0;
        if (i < n) {
            lo[k] = gld_nt<uint64_t>(klo, beg + i);
            if (TWO) hi[k] = gld_nt<uint64_t>(khi, beg + i);
            row[k] = CARRY ? i : gld_nt<uint32_t>(sp_.row, beg + i);
        }
    }
    __syncthreads();
#pragma unroll
    for (uint32_t k = 0; k < kPer; ++k) {
        const uint32_t i = k * kBucketThreads + tid;
        br[k] = ~0u;
        if (i < n) {
            const uint32_t bn = (uint32_t)(key_bits_from(lo[k], hi[k], bin_shift) & bmask);
            br[k] = bn << 16 | atomicAdd(&B.bin[bn], 1u);
        }
    }
    __syncthreads();
    // bin starts: each thread scans its consecutive run of bins
    {
        constexpr uint32_t kRun = kBins / kBucketThreads;
        uint32_t c[kRun], tot = 0;
#pragma unroll
        for (uint32_t r = 0; r < kRun; ++r) {
            const uint32_t i = tid * kRun + r;
            c[r] = i < nbins ? B.bin[i] : 0;
            tot += c[r];
            if (c[r] > kBinSortMax) *s_crowded = 1;
        }
        uint32_t base = block_excl_scan(tot, s_wave);
#pragma unroll
        for (uint32_t r = 0; r < kRun; ++r) {
            const uint32_t i = tid * kRun + r;
            if (i < nbins) B.bin[i] = base;
            base += c[r];
        }
    }
    __syncthreads();
#pragma unroll
    for (uint32_t k = 0; k < kPer; ++k) {
        if (br[k] == ~0u) continue;
        const uint32_t d = B.bin[br[k] >> 16] + (br[k] & 0xFFFFu);
        B.lo[d] = lo[k];
        if (TWO) B.hi[d] = hi[k];
        B.row[d] = row[k];
    }
    __syncthreads();
    if (*s_crowded) {
        // a crowded bin: sort the whole bucket (padded) instead
        uint32_t m = 2;
        while (m < n) m <<= 1;
        for (uint32_t i = n + tid; i < m; i += kBucketThreads) {
            B.lo[i] = ~0ull;
            if (TWO) B.hi[i] = ~0ull;
            B.row[i] = ~0u;
        }
        __syncthreads();
        lds_bitonic<TWO>(B.lo, B.hi, B.row, m);
        return false;
    }
    // inside each bin, a record's place is the number of the bin's records
    // that order before it (keys, then rows: all distinct)
#pragma unroll
    for (uint32_t k = 0; k < kPer; ++k) {
        if (br[k] == ~0u) continue;
        const uint32_t bn = br[k] >> 16;
        const uint32_t s0 = B.bin[bn], s1 = bn + 1 < nbins ? B.bin[bn + 1] : n;
        uint32_t r = 0;
        for (uint32_t x = s0; x < s1; ++x)
            r += rec_less(TWO ? B.hi[x] : 0, B.lo[x], B.row[x], hi[k], lo[k], row[k]) ? 1u : 0u;
        B.perm[s0 + r] = (uint16_t)(s0 + (br[k] & 0xFFFFu));
    }
    __syncthreads();
    return true;
}

// A sorted bucket out to the destination buffer: the keys, and the row or
// (CARRY) the values and validity found at the record's place in the
// source bucket.
template <bool TWO, bool CARRY>
__device__ __forceinline__ void bucket_out(const SLaunch *__restrict__ L, const Payload &sp_, uint64_t *dlo,
                                           uint64_t *dhi, const Payload &dp_, uint64_t beg, uint32_t n,
                                           const BucketBufs<TWO> &B, bool ranked) {
    const int nv = CARRY ? L->q.n_vals : 0;
    for (uint32_t i = threadIdx.x; i < n; i += kBucketThreads) {
        const uint32_t x = ranked ? B.perm[i] : i;
        gst<uint64_t>(dlo, beg + i, B.lo[x]);
        if (TWO) gst<uint64_t>(dhi, beg + i, B.hi[x]);
        if (CARRY) {
            const uint64_t from = beg + B.row[x];
            gst<uint64_t>(dp_.v[0], beg + i, gld<uint64_t>(sp_.v[0], from));
            if (nv > 1) gst<uint64_t>(dp_.v[1], beg + i, gld<uint64_t>(sp_.v[1], from));
            gst<uint8_t>(dp_.f, beg + i, gld<uint8_t>(sp_.f, from));
        } else {
            gst<uint32_t>(dp_.row, beg + i, B.row[x]);
        }
    }
}

// One bucket per block iteration: sorted in LDS and written, keys and
// payload, to the other buffer.  A bucket larger than its LDS capacity is
// left as it is and flags SC_BIG: the host then sorts with whole-key LSD
// passes.
template <bool TWO, bool CARRY>
__global__ __launch_bounds__(kBucketThreads) void so_bucket_sort(const SLaunch *__restrict__ L, const uint64_t *klo,
                                                               const uint64_t *khi, const Payload sp_,
                                                               uint64_t *dlo, uint64_t *dhi, const Payload dp_,
                                                               const uint64_t *bstart, uint32_t nb, int bin_shift,
                                                               int bin_bits) {
    extern __shared__ __align__(16) uint8_t smem[];
    const BucketBufs<TWO> B = bucket_bufs<TWO>(smem);
    __shared__ uint32_t s_wave[16];
    __shared__ uint32_t s_crowded;
    for (uint32_t b = blockIdx.x; b < nb; b += gridDim.x) {
        const uint64_t beg = bstart[b], end = bstart[b + 1];
        const uint64_t n64 = end - beg;  // uniform across the block
        if (n64 <= 1) continue;
        if (n64 > bucket_cap(TWO)) {
            if (threadIdx.x == 0) atomicAdd((unsigned long long *)&L->ctr[SC_BIG], 1ull);
            continue;
        }
        const uint32_t n = (uint32_t)n64;
        const bool ranked = bucket_sort_lds<TWO, CARRY>(klo, khi, sp_, beg, n, B, 1u << bin_bits, bin_shift, s_wave,
                                                        &s_crowded);
        bucket_out<TWO, CARRY>(L, sp_, dlo, dhi, dp_, beg, n, B, ranked);
        __syncthreads();
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

// seg_scan with nr rounds: complete when every segment of the step has
// fewer than 2^nr lanes.
template <class T, class Op>
__device__ __forceinline__ T seg_scan_n(T x, int seg_start, int lane, Op op, int nr) {
    for (int r = 0; r < nr; ++r) {
        const int off = 1 << r;
        const T y = __shfl_up(x, off);
        if (lane - off >= seg_start) x = op(x, y);
    }
    return x;
}

// Lane src's value (src uniform across the wave).
__device__ __forceinline__ uint32_t lane_u32(uint32_t x, int src) {
    return (uint32_t)__builtin_amdgcn_readlane((int)x, src);
}
__device__ __forceinline__ uint64_t lane_u64(uint64_t x, int src) {
    return (uint64_t)lane_u32((uint32_t)x, src) | (uint64_t)lane_u32((uint32_t)(x >> 32), src) << 32;
}
__device__ __forceinline__ int64_t lane_i64(int64_t x, int src) { return (int64_t)lane_u64((uint64_t)x, src); }

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

// The common shape (at most two aggregated columns): runs of one wave step
// of 64 records fold with segmented scans as in so_reduce; a step's last run
// stays open in registers (Carry) and joins the next step's first run.
// res: the values outside the exact window (and Inf/NaN), summed in f64 in
// record order, as a sequential fold sums them (their sum depends on the
// order: 1e308 - 1e308 + 5e-324 is 5e-324 only in that order).
struct Carry {
    uint32_t len, cnt[2], nan[2];
    int64_t limb[2][kLimbs], sum[2], mn[2], mx[2];
    double res[2];
};

__device__ __forceinline__ void carry_reset(Carry &c) {
    c.len = 0;
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        c.cnt[j] = c.nan[j] = 0;
        c.sum[j] = 0;
        c.mn[j] = INT64_MAX;
        c.mx[j] = INT64_MIN;
        c.res[j] = 0.0;
#pragma unroll
        for (int z = 0; z < kLimbs; ++z) c.limb[j][z] = 0;
    }
}

// x (this step's run that began at lane 0) continues the run carried in c.
// EXACT: some column keeps exact f64 sums (limbs and residual); else those
// fields stay zero and are skipped.  The residual is not joined here: the
// step's fold starts its first run from the carried one (fold_records).
template <bool EXACT>
__device__ __forceinline__ void carry_join(Carry &x, const Carry &c) {
    x.len += c.len;
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        x.cnt[j] += c.cnt[j];
        x.nan[j] += c.nan[j];
        x.sum[j] += c.sum[j];
        x.mn[j] = x.mn[j] < c.mn[j] ? x.mn[j] : c.mn[j];
        x.mx[j] = x.mx[j] > c.mx[j] ? x.mx[j] : c.mx[j];
        if (!EXACT) continue;
#pragma unroll
        for (int z = 0; z < kLimbs; ++z) x.limb[j][z] += c.limb[j][z];
    }
}

// c = lane src's x (every lane; src uniform).
template <bool EXACT>
__device__ __forceinline__ void carry_take(Carry &c, const Carry &x, int src) {
    c.len = lane_u32(x.len, src);
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        c.cnt[j] = lane_u32(x.cnt[j], src);
        c.nan[j] = lane_u32(x.nan[j], src);
        c.sum[j] = lane_i64(x.sum[j], src);
        c.mn[j] = lane_i64(x.mn[j], src);
        c.mx[j] = lane_i64(x.mx[j], src);
        if (!EXACT) continue;
        c.res[j] = __longlong_as_double((long long)lane_u64((uint64_t)__double_as_longlong(x.res[j]), src));
#pragma unroll
        for (int z = 0; z < kLimbs; ++z) c.limb[j][z] = lane_i64(x.limb[j][z], src);
    }
}

// The aggregated columns of the fast fold.
struct FoldVals {
    int nv;
    ValSpec vs[2];
    RoleCol vc[2];
};
__device__ __forceinline__ FoldVals fold_vals(const QueryDesc *__restrict__ q) {
    FoldVals F;
    F.nv = q->n_vals;
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        F.vs[j] = q->vals[j < F.nv ? j : 0];
        F.vc[j] = role_col(q, F.vs[j].col);
    }
    return F;
}

// A record's value bits and validity: CARRY from the payload at index at,
// else from the source columns at row at.  Every word is loaded before any
// is used.
template <bool CARRY>
__device__ __forceinline__ void record_values(const FoldVals &F, const Payload &pl, bool live, uint64_t at,
                                              uint64_t (&raw)[2], bool (&ok)[2]) {
    uint64_t w[2], vb[2];
    const uint32_t fl = CARRY && live ? gld<uint8_t>(pl.f, at) : 0;
    const uint64_t row = live ? at : 0;
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        w[j] = vb[j] = 0;
        if (j >= F.nv) continue;
        if (CARRY) {
            w[j] = live ? gld<uint64_t>(pl.v[j], at) : 0;
        } else {
            w[j] = word_at(F.vc[j].base, vpos(F.vc[j], row));
            vb[j] = word_at(F.vc[j].vbase, bpos(F.vc[j], row));
        }
    }
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        raw[j] = 0;
        ok[j] = false;
        if (j >= F.nv) continue;
        raw[j] = CARRY ? w[j] : cut(w[j], vpos(F.vc[j], row), F.vc[j].mask);
        ok[j] = live && (CARRY ? ((fl >> j) & 1) != 0 : cut(vb[j], bpos(F.vc[j], row), 1) != 0);
    }
}

// x = each lane's run totals from its segment's first lane (seg_start) up to
// itself; the segment's last lane holds the run's totals.  res0: the
// residual the first segment continues from (the carried run's, or zero).
// Counts come from ballots (no shuffles); sums, limbs and min/max scan
// with only the rounds the step's longest segment needs.
template <bool EXACT>
__device__ __forceinline__ void fold_records(Carry &x, const FoldVals &F, bool live, int seg_start, int lane,
                                             const uint64_t (&raw)[2], const bool (&ok)[2], const double (&res0)[2]) {
    auto add64 = [](int64_t a, int64_t b) { return a + b; };
    auto mn64 = [](int64_t a, int64_t b) { return a < b ? a : b; };
    auto mx64 = [](int64_t a, int64_t b) { return a > b ? a : b; };
    carry_reset(x);
    // the lanes of this lane's segment up to itself
    const unsigned long long seg = ((2ull << lane) - 1) & ~((1ull << seg_start) - 1);
    int nr = 0;
    while (nr < 6 && __ballot(live && lane - seg_start >= (1 << nr))) ++nr;
    x.len = (uint32_t)__popcll(__ballot(live) & seg);
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        if (j >= F.nv) continue;
        const ValSpec &S = F.vs[j];
        x.cnt[j] = (uint32_t)__popcll(__ballot(ok[j]) & seg);
        if (S.kind == VK_F) {
            const double v = F.vc[j].tag == PMH_F64 ? __longlong_as_double((long long)raw[j])
                                                    : (double)__uint_as_float((uint32_t)raw[j]);
            if (EXACT && (S.need & NEED_SUM)) {
                int64_t lb[kLimbs] = {0, 0, 0, 0, 0};
                bool out = false;
                if (ok[j]) {
                    const uint64_t bb = (uint64_t)__double_as_longlong(v);
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
                        const int li = p >> 5, sh = p & 31;
                        const uint64_t lo = mm << sh;
                        const uint64_t hi = sh ? (mm >> (64 - sh)) : 0;
                        int64_t c0 = (int64_t)(lo & 0xFFFFFFFFull), c1 = (int64_t)(lo >> 32), c2 = (int64_t)hi;
                        if (bb >> 63) { c0 = -c0; c1 = -c1; c2 = -c2; }
#pragma unroll
                        for (int z = 0; z < kLimbs; ++z)
                            lb[z] = (z == li ? c0 : 0) + (z == li + 1 ? c1 : 0) + (z == li + 2 ? c2 : 0);
                    } else {
                        out = true;  // outside the window (rare): the residual
                    }
                }
#pragma unroll
                for (int z = 0; z < kLimbs; ++z) x.limb[j][z] = seg_scan_n<int64_t>(lb[z], seg_start, lane, add64, nr);
                // the rare residual values one at a time, in record order
                double r = seg_start == 0 ? res0[j] : 0.0;
                for (unsigned long long om = __ballot(out); om; om &= om - 1) {
                    const int src = __builtin_ctzll(om);
                    const double val = __longlong_as_double((long long)lane_u64((uint64_t)__double_as_longlong(v), src));
                    const int sseg = (int)lane_u32((uint32_t)seg_start, src);
                    if (lane >= src && seg_start == sseg) r += val;
                }
                x.res[j] = r;
            }
            if (S.need & NEED_MINMAX) {
                const bool nan = ok[j] && v != v;
                const int64_t kk = f64_order_key(v);
                x.mn[j] = seg_scan_n<int64_t>(ok[j] && !nan ? kk : INT64_MAX, seg_start, lane, mn64, nr);
                x.mx[j] = seg_scan_n<int64_t>(ok[j] && !nan ? kk : INT64_MIN, seg_start, lane, mx64, nr);
                x.nan[j] = (uint32_t)__popcll(__ballot(nan) & seg);
            }
        } else {
            const int64_t v = S.kind == VK_I64 ? (int64_t)raw[j] : (int64_t)(int32_t)(uint32_t)raw[j];
            if (S.need & NEED_SUM) x.sum[j] = seg_scan_n<int64_t>(ok[j] ? v : 0, seg_start, lane, add64, nr);
            if (S.need & NEED_MINMAX) {
                x.mn[j] = seg_scan_n<int64_t>(ok[j] ? v : INT64_MAX, seg_start, lane, mn64, nr);
                x.mx[j] = seg_scan_n<int64_t>(ok[j] ? v : INT64_MIN, seg_start, lane, mx64, nr);
            }
        }
    }
}

// One wave per tile of kSortTile sorted records.  A group spanning many
// records costs one atomic per tile and state array instead of one per wave.
template <bool CARRY, bool EXACT>
__global__ __launch_bounds__(256) void so_reduce_fast(const SLaunch *__restrict__ L, const uint64_t *klo,
                                                     const uint64_t *khi, const Payload pl) {
    const QueryDesc *__restrict__ q = &L->q;
    const Layout &lay = q->layout;
    const bool two = L->kw == 2;
    const int lane = threadIdx.x & 63;
    const uint32_t t = blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64;
    if (t >= L->ntiles) return;  // whole waves leave together
    const FoldVals F = fold_vals(q);
    const int nv = F.nv;
    GSink g{L->arr, L->gcap};
    // one lane writes the totals of one run of group gid: with plain stores
    // when the run is the group's only one (it starts and ends inside this
    // tile, so no other tile touches the group; the arrays hold identities),
    // else with atomics
    auto flush = [&](uint64_t gid, const Carry &x, bool only) {
        auto w32 = [&](int a, uint32_t v) {
            if (only) gst<uint32_t>(L->arr[a], gid, v);
            else g.add32(a, gid, v);
        };
        if (x.len) w32(lay.a_len, x.len);
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            if (j >= nv || !x.cnt[j]) continue;
            const ValSpec &S = F.vs[j];
            w32(S.a_cnt, x.cnt[j]);
            if (S.need & NEED_SUM) {
                if (S.kind == VK_F) {
                    if (!EXACT) continue;  // not reached: EXACT is set when a column sums floats
#pragma unroll
                    for (int z = 0; z < kLimbs; ++z) {
                        if (!x.limb[j][z]) continue;
                        if (only) gst<int64_t>(L->arr[S.a_limb], (uint64_t)z * L->gcap + gid, x.limb[j][z]);
                        else g.addlimb(S.a_limb, z, gid, x.limb[j][z]);
                    }
                    if (x.res[j] != 0.0 || x.res[j] != x.res[j]) {
                        if (only) gst<double>(L->arr[S.a_res], gid, x.res[j]);
                        else g.addf(S.a_res, gid, x.res[j]);
                    }
                } else if (only) {
                    gst<int64_t>(L->arr[S.a_sum], gid, x.sum[j]);
                } else {
                    g.add64(S.a_sum, gid, x.sum[j]);
                }
            }
            if (S.need & NEED_MINMAX) {
                if (S.kind == VK_I32) {
                    if (only) {
                        gst<int32_t>(L->arr[S.a_min], gid, (int32_t)x.mn[j]);
                        gst<int32_t>(L->arr[S.a_max], gid, (int32_t)x.mx[j]);
                    } else {
                        if (x.mn[j] != INT64_MAX) g.min32(S.a_min, gid, (int32_t)x.mn[j]);
                        if (x.mx[j] != INT64_MIN) g.max32(S.a_max, gid, (int32_t)x.mx[j]);
                    }
                } else if (only) {
                    gst<int64_t>(L->arr[S.a_min], gid, x.mn[j]);
                    gst<int64_t>(L->arr[S.a_max], gid, x.mx[j]);
                } else {
                    if (x.mn[j] != INT64_MAX) g.min64(S.a_min, gid, x.mn[j]);
                    if (x.mx[j] != INT64_MIN) g.max64(S.a_max, gid, x.mx[j]);
                }
                if (S.kind == VK_F && x.nan[j]) {
                    if (only) gst<uint32_t>(L->arr[S.a_nan], gid, 1u);
                    else g.or32(S.a_nan, gid, 1u);
                }
            }
        }
    };
    Carry c;
    carry_reset(c);
    bool have = false;
    uint64_t cgid = 0;
    const uint64_t tb = (uint64_t)t * kSortTile;
    uint64_t before = 0;  // group starts seen in this tile before the current iteration
    bool c_own = false;   // the carried run started inside this tile
    for (uint32_t k = 0; k < kSortTile; k += 64) {
        const uint64_t i = tb + k + lane;
        const bool live = i < L->n_sel;
        const unsigned long long lv = __ballot(live);
        if (!lv) break;
        const bool start = live && is_start(klo, khi, two, i);
        const unsigned long long sm = __ballot(start);
        const uint64_t gid = (uint64_t)L->tile_cnt[t] + before + (uint32_t)__popcll(sm & ((2ull << lane) - 1)) - 1;
        before += (uint32_t)__popcll(sm);
        if (start) {
            gst<uint64_t>(L->gk_lo, gid, gld<uint64_t>(klo, i));
            if (two) gst<uint64_t>(L->gk_hi, gid, gld<uint64_t>(khi, i));
        }
        // the open run of the previous iteration ends where this one starts
        if (have && (sm & 1ull)) {
            if (lane == 0) flush(cgid, c, c_own);
            have = false;
        }
        const unsigned long long heads = sm | 1ull;
        const int seg_start = 63 - __builtin_clzll(heads & ((2ull << lane) - 1));
        const int last_live = 63 - __builtin_clzll(lv);
        const bool seg_last = live && (lane == last_live || ((heads >> (lane + 1)) & 1));
        uint64_t raw[2];
        bool ok[2];
        record_values<CARRY>(F, pl, live, CARRY ? i : (live ? gld<uint32_t>(pl.row, i) : 0), raw, ok);
        Carry x;
        const double res0[2] = {have ? c.res[0] : 0.0, have ? c.res[1] : 0.0};
        fold_records<EXACT>(x, F, live, seg_start, lane, raw, ok, res0);
        // a continuing run: the carry joins this iteration's first run
        if (have && seg_start == 0 && seg_last) carry_join<EXACT>(x, c);
        // closed runs go out; the last run of the iteration becomes the carry.
        // A run owns its group unless it began at lane 0 as the carried run
        // of a tile whose first records continue the previous tile's group.
        const bool own = seg_start > 0 || (sm & 1ull) || (have && c_own);
        if (seg_last && lane != last_live) flush(gid, x, own);
        c_own = lane_u32(own ? 1u : 0u, last_live) != 0;
        cgid = lane_u64(gid, last_live);
        carry_take<EXACT>(c, x, last_live);
        have = true;
    }
    // the tile's last run is the group's only one if the next tile starts a
    // new group
    const uint64_t nx = tb + kSortTile;
    const bool closed = nx >= L->n_sel || is_start(klo, khi, two, nx);
    if (have && lane == 0) flush(cgid, c, c_own && closed);
}

// Every state field of group gid, with plain stores (the fold's state
// arrays are not initialised: a group's one writer stores all of them).
template <bool EXACT>
__device__ __forceinline__ void fold_store(const SLaunch *__restrict__ L, const FoldVals &F, uint64_t gid,
                                           const Carry &x) {
    const Layout &lay = L->q.layout;
    gst<uint32_t>(L->arr[lay.a_len], gid, x.len);
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        if (j >= F.nv) continue;
        const ValSpec &S = F.vs[j];
        gst<uint32_t>(L->arr[S.a_cnt], gid, x.cnt[j]);
        if (S.need & NEED_SUM) {
            if (S.kind == VK_F) {
                if (!EXACT) continue;  // not reached: EXACT is set when a column sums floats
#pragma unroll
                for (int z = 0; z < kLimbs; ++z) gst<int64_t>(L->arr[S.a_limb + z], gid, x.limb[j][z]);
                gst<double>(L->arr[S.a_res], gid, x.res[j]);
            } else {
                gst<int64_t>(L->arr[S.a_sum], gid, x.sum[j]);
            }
        }
        if (S.need & NEED_MINMAX) {
            if (S.kind == VK_I32) {
                gst<int32_t>(L->arr[S.a_min], gid, (int32_t)x.mn[j]);
                gst<int32_t>(L->arr[S.a_max], gid, (int32_t)x.mx[j]);
            } else {
                gst<int64_t>(L->arr[S.a_min], gid, x.mn[j]);
                gst<int64_t>(L->arr[S.a_max], gid, x.mx[j]);
            }
            if (S.kind == VK_F) gst<uint32_t>(L->arr[S.a_nan], gid, x.nan[j] ? 1u : 0u);
        }
    }
}

// Bucket fold: each bucket is sorted in LDS as in so_bucket_sort and its
// groups are folded straight from LDS; no sorted record is written back.  A
// bucket holds whole groups (its records share the key's top bits), and
// each wave folds whole groups: wave w takes the groups that start in
// [w n / W, (w + 1) n / W) of the bucket's n sorted records (W waves).  A
// group's state (the fast fold's shape, at most two aggregated columns) and
// compressed key go to sparse index bstart[b] + (its rank among the
// bucket's groups), with plain stores; bg[b] counts the bucket's groups.
// The emit reads them back in bucket order.  Four waves per SIMD (at most
// 128 registers): two blocks per CU.
template <bool TWO, bool CARRY, bool EXACT>
__global__ __launch_bounds__(kBucketThreads, 4) void so_bucket_fold(const SLaunch *__restrict__ L, const uint64_t *klo,
                                                               const uint64_t *khi, const Payload sp_,
                                                               const uint64_t *bstart, uint32_t nb, int bin_shift,
                                                               int bin_bits, uint32_t *bg) {
    constexpr int kWB = kBucketThreads / 64;
    constexpr uint32_t cap = bucket_cap(TWO);
    extern __shared__ __align__(16) uint8_t smem[];
    const BucketBufs<TWO> B = bucket_bufs<TWO>(smem);
    __shared__ uint32_t s_wave[16];
    __shared__ uint32_t s_crowded;
    __shared__ uint32_t s_split[kWB], s_gcnt[kWB];
    __shared__ unsigned long long s_start[cap / 64];  // bit i: sorted place i starts a group
    const FoldVals F = fold_vals(&L->q);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (uint32_t b = blockIdx.x; b < nb; b += gridDim.x) {
        const uint64_t beg = bstart[b], n64 = bstart[b + 1] - beg;  // uniform across the block
        if (n64 == 0 || n64 > cap) {
            if (threadIdx.x == 0) {
                bg[b] = 0;
                if (n64) atomicAdd((unsigned long long *)&L->ctr[SC_BIG], 1ull);
            }
            continue;
        }
        const uint32_t n = (uint32_t)n64;
        const bool ranked = bucket_sort_lds<TWO, CARRY>(klo, khi, sp_, beg, n, B, 1u << bin_bits, bin_shift, s_wave,
                                                        &s_crowded);
        auto rec = [&](uint32_t i) -> uint32_t { return ranked ? B.perm[i] : i; };
        // group starts, one bit per sorted place
        for (uint32_t k = (uint32_t)w * 64; k < n; k += kBucketThreads) {
            const uint32_t i = k + lane;
            bool st = false;
            if (i < n) {
                const uint32_t x = rec(i);
                st = i == 0;
                if (!st) {
                    const uint32_t p = rec(i - 1);
                    st = B.lo[x] != B.lo[p] || (TWO && B.hi[x] != B.hi[p]);
                }
            }
            const unsigned long long m = __ballot(st);
            if (lane == 0) s_start[k >> 6] = m;
        }
        __syncthreads();
        auto starts = [&](uint32_t k) -> unsigned long long { return k < n ? s_start[k >> 6] : 0ull; };
        // the wave's records: from the first group start at or after w n / W
        // to the next wave's first (n for the last wave)
        {
            const uint32_t t0 = (uint32_t)((uint64_t)w * n / kWB);
            uint32_t s = n;
            for (uint32_t k = t0 & ~63u; k < n; k += 64) {
                unsigned long long m = starts(k);
                if (k < t0) m &= ~0ull << (t0 - k);
                if (k + 64 > n) m &= (n - k) == 64 ? ~0ull : ((1ull << (n - k)) - 1);
                if (m) {
                    s = k + (uint32_t)__builtin_ctzll(m);
                    break;
                }
            }
            if (lane == 0) s_split[w] = s;
        }
        __syncthreads();
        const uint32_t r0 = s_split[w], r1 = w + 1 < kWB ? s_split[w + 1] : n;
        // groups of each wave, then of the waves before it
        {
            uint32_t cnt = 0;
            for (uint32_t k = r0 & ~63u; k < r1; k += 64) {
                unsigned long long m = starts(k);
                if (k < r0) m &= ~0ull << (r0 - k);
                if (k + 64 > r1) m &= (r1 - k) == 64 ? ~0ull : ((1ull << (r1 - k)) - 1);
                cnt += (uint32_t)__popcll(m);
            }
            if (lane == 0) s_gcnt[w] = cnt;
        }
        __syncthreads();
        uint32_t gbase = 0, total = 0;
#pragma unroll
        for (int v = 0; v < kWB; ++v) {
            const uint32_t cv = s_gcnt[v];
            gbase += v < w ? cv : 0;
            total += cv;
        }
        if (threadIdx.x == 0) bg[b] = total;
        Carry c;
        carry_reset(c);
        bool have = false;
        uint64_t cgid = 0;
        uint32_t before = 0;
        for (uint32_t k = r0; k < r1; k += 64) {
            const uint32_t i = k + lane;
            const bool live = i < r1;
            const unsigned long long lv = __ballot(live);
            // starts among this step's places k .. k + 63 (k need not be a multiple of 64)
            const uint32_t sh = k & 63u;
            const unsigned long long w0 = starts(k & ~63u), w1 = starts((k & ~63u) + 64);
            const unsigned long long sm = (sh ? (w0 >> sh) | (w1 << (64 - sh)) : w0) & lv;
            const bool start = (sm >> lane) & 1;
            const uint64_t gid = beg + gbase + before + (uint32_t)__popcll(sm & ((2ull << lane) - 1)) - 1;
            before += (uint32_t)__popcll(sm);
            const uint32_t x = live ? rec(i) : 0;
            if (start) {
                gst<uint64_t>(L->gk_lo, gid, B.lo[x]);
                if (TWO) gst<uint64_t>(L->gk_hi, gid, B.hi[x]);
            }
            if (have && (sm & 1ull)) {
                if (lane == 0) fold_store<EXACT>(L, F, cgid, c);
                have = false;
            }
            const unsigned long long heads = sm | 1ull;
            const int seg_start = 63 - __builtin_clzll(heads & ((2ull << lane) - 1));
            const int last_live = 63 - __builtin_clzll(lv);
            const bool seg_last = live && (lane == last_live || ((heads >> (lane + 1)) & 1));
            uint64_t raw[2];
            bool ok[2];
            const uint32_t at = live ? B.row[x] : 0;
            record_values<CARRY>(F, sp_, live, CARRY ? beg + at : (uint64_t)at, raw, ok);
            Carry xs;
            const double res0[2] = {have ? c.res[0] : 0.0, have ? c.res[1] : 0.0};
            fold_records<EXACT>(xs, F, live, seg_start, lane, raw, ok, res0);
            if (have && seg_start == 0 && seg_last) carry_join<EXACT>(xs, c);
            if (seg_last && lane != last_live) fold_store<EXACT>(L, F, gid, xs);
            cgid = lane_u64(gid, last_live);
            carry_take<EXACT>(c, xs, last_live);
            have = true;
        }
        if (have && lane == 0) fold_store<EXACT>(L, F, cgid, c);
        __syncthreads();  // the next bucket reuses the LDS
    }
}

// Exclusive prefix of n counts (pre may be cnt: in place) and their total,
// by blocks of 4096 counts, four per thread read as one 16-byte word (cnt is
// 16-byte aligned): so_count_sum leaves each block's sum in bsum, then
// so_count_scan takes the sums of the blocks before its own (one wave) and
// writes its block's prefixes; the last block writes the total.
__device__ __forceinline__ uint4 count_quad(const uint32_t *cnt, uint32_t n, uint64_t i) {
    if (i + 4 <= n) return *(const uint4 *)(cnt + i);
    uint4 v = make_uint4(0, 0, 0, 0);
    if (i < n) v.x = cnt[i];
    if (i + 1 < n) v.y = cnt[i + 1];
    if (i + 2 < n) v.z = cnt[i + 2];
    return v;
}

__global__ __launch_bounds__(1024) void so_count_sum(const uint32_t *cnt, uint32_t n, uint32_t *bsum) {
    __shared__ uint32_t s_wave[16];
    const uint4 v = count_quad(cnt, n, ((uint64_t)blockIdx.x * 1024 + threadIdx.x) * 4);
    const uint32_t s = v.x + v.y + v.z + v.w;
    const uint32_t e = block_excl_scan(s, s_wave);
    if (threadIdx.x == 1023) bsum[blockIdx.x] = e + s;
}

__global__ __launch_bounds__(1024) void so_count_scan(const uint32_t *cnt, uint32_t *pre, uint32_t n,
                                                    const uint32_t *bsum, uint64_t *total) {
    __shared__ uint32_t s_wave[16];
    __shared__ uint32_t s_off;
    if (threadIdx.x < 64) {
        uint32_t o = 0;
        for (uint32_t j = threadIdx.x; j < blockIdx.x; j += 64) o += bsum[j];
        o = wave_sum(o);
        if (threadIdx.x == 0) s_off = o;
    }
    __syncthreads();
    const uint64_t i = ((uint64_t)blockIdx.x * 1024 + threadIdx.x) * 4;
    const uint4 v = count_quad(cnt, n, i);
    const uint32_t s = v.x + v.y + v.z + v.w;
    const uint32_t base = s_off + block_excl_scan(s, s_wave);
    if (i + 4 <= n) {
        *(uint4 *)(pre + i) = make_uint4(base, base + v.x, base + v.x + v.y, base + v.x + v.y + v.z);
    } else {
        if (i < n) pre[i] = base;
        if (i + 1 < n) pre[i + 1] = base + v.x;
        if (i + 2 < n) pre[i + 2] = base + v.x + v.y;
    }
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 1023) *total = base + s;
}

// ---------------------------------------------------------------------------
// 6. outputs
// ---------------------------------------------------------------------------
// The lane key of the group whose compressed key is at gk[s].  One 64-bit
// key column: the sort lane's encoding back to the raw key (special 2: the
// null key, 1: the key equal to the empty marker).
__device__ __forceinline__ void group_key(const SLaunch *__restrict__ L, uint64_t s, bool two, uint64_t &lo,
                                          uint64_t &hi, int &special) {
    decompress(L, gld<uint64_t>(L->gk_lo, s), two ? gld<uint64_t>(L->gk_hi, s) : 0, lo, hi);
    special = 0;
    if (L->q.key_mode == KM_RAW64) {
        if (lo & 1) {
            special = 2;
            lo = 0;
        } else {
            lo = (lo >> 1) | (hi << 63);
            special = lo == kEmpty ? 1 : 0;
        }
        hi = 0;
    }
}

__global__ __launch_bounds__(256) void so_emit(const SLaunch *__restrict__ L) {
    const QueryDesc *__restrict__ q = &L->q;
    const uint64_t ng = L->ctr[SC_GROUPS];
    const bool two = L->kw == 2;
    const GReader rd{L->arr};
    for (uint64_t gi = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; gi < ng; gi += (uint64_t)gridDim.x * blockDim.x) {
        uint64_t lo, hi;
        int special;
        group_key(L, gi, two, lo, hi, special);
        emit_group(q, L->out, L->n_out, rd, gi, lo, hi, special, gi);
    }
}

// Outputs of the bucket fold: bucket b's groups sit at sparse index
// bstart[b] + j and go to output row bpre[b] + j (bucket order is key order).
// One wave per bucket, so a block has four buckets in flight.
__global__ __launch_bounds__(256, 8) void so_emit_buckets(const SLaunch *__restrict__ L, const uint64_t *bstart,
                                                     const uint32_t *bg, const uint32_t *bpre, uint32_t nb) {
    const QueryDesc *__restrict__ q = &L->q;
    const bool two = L->kw == 2;
    const GReader rd{L->arr};
    const uint32_t lane = threadIdx.x & 63, nw = blockDim.x / 64;
    for (uint32_t b = blockIdx.x * nw + threadIdx.x / 64; b < nb; b += gridDim.x * nw) {
        const uint32_t cnt = bg[b];
        const uint64_t s0 = bstart[b], o0 = bpre[b];
        for (uint32_t j = lane; j < cnt; j += 64) {
            const uint64_t s = s0 + j;
            uint64_t lo, hi;
            int special;
            group_key(L, s, two, lo, hi, special);
            emit_group(q, L->out, L->n_out, rd, s, lo, hi, special, o0 + j);
        }
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
    Runtime &R = rt();
    const int cus = R.prop.multiProcessorCount > 0 ? R.prop.multiProcessorCount : 256;
    auto Lh = std::make_unique<SLaunch>();
    memset(Lh.get(), 0, sizeof(SLaunch));
    SLaunch &L = *Lh;
    L.q = P.q;
    L.n_out = P.n_out;
    L.ctr = (uint64_t *)slab("so:ctr", SC_N * 8);
    L.nblk = (uint32_t)(cus * 8);
    L.bpart = (uint64_t *)slab("so:bpart", (uint64_t)L.nblk * 5 * 8 + 64);
    L.boff = (uint64_t *)slab("so:boff", (uint64_t)L.nblk * 8 + 64);
    {
        bool aligned = true;
        for (int c = 0; c < L.q.n_cols; ++c) {
            const int w = tag_size(L.q.cols[c].tag);
            if (w > 1 && ((uintptr_t)L.q.cols[c].data % (uintptr_t)w) != 0) aligned = false;
        }
        L.fast = aligned && L.q.n_keys <= 4 && plan.opts.fast_scan != 0 && lower_fast_scan(L.q, L.fs) ? 1 : 0;
    }

    SLaunch *dL = (SLaunch *)slab("so:launch", sizeof(SLaunch));
    PooledEvent pe0, pe1;
    const hipEvent_t ev0 = pe0.ev, ev1 = pe1.ev;
    PMH_HIP(hipEventRecord(ev0, R.stream));

    // ---- 1. varying lane bits and the selected-row count: exactly, by a
    // pass over every row; or, for a large query without a predicate (every
    // row is a record), from an evenly spaced sample, which the encode then
    // checks every record against (a record outside it redoes the exact way)
    uint64_t c[SC_N];
    auto bits_pass = [&](bool sampled) {
        L.sampled = sampled ? 1 : 0;
        PMH_HIP(hipMemcpyAsync(dL, &L, sizeof(SLaunch), hipMemcpyHostToDevice, R.stream));
        hipLaunchKernelGGL(so_reset, dim3(1), dim3(64), 0, R.stream, dL);
        if (sampled) {
            hipLaunchKernelGGL(so_bits_sample, dim3(kBitsSample / 1024), dim3(1024), 0, R.stream, dL);
            launch_check();
        } else {
            if (L.fast) {
                KernelTimer kt("so_bits_fast");
                hipLaunchKernelGGL(so_scan_fast<false>, dim3(cus * 8), dim3(512), 0, R.stream, dL);
                launch_check();
            } else {
                KernelTimer kt("so_bits");
                hipLaunchKernelGGL(so_bits, dim3(cus * 8), dim3(256), 0, R.stream, dL);
                launch_check();
            }
            hipLaunchKernelGGL(so_bits_final, dim3(1), dim3(1024), 0, R.stream, dL);
            launch_check();
        }
        PMH_HIP(hipMemcpyAsync(c, L.ctr, sizeof(c), hipMemcpyDeviceToHost, R.stream));
        PMH_HIP(hipStreamSynchronize(R.stream));
    };
    uint64_t n_sel = 0;
    uint32_t width = 0;
    int full_passes = 0, tbits = 0;
    uint64_t bmean = 0;
    // the top-bit passes for tbits, or the whole key's
    std::vector<int> shifts;
    auto plan_shifts = [&]() {
        shifts.clear();
        if (tbits) {
            const int np = (tbits + kDigitBits - 1) / kDigitBits;
            for (int p = 0; p < np; ++p)
                shifts.push_back(std::min<int>((int)width - tbits + p * kDigitBits, (int)width - kDigitBits));
        } else {
            for (int p = 0; p < full_passes; ++p) shifts.push_back(p * kDigitBits);
        }
    };
    // the compressed record layout, the passes and the buffers, from the bits
    auto plan_layout = [&]() {
        n_sel = c[SC_SEL];
        L.n_sel = n_sel;
        plan_runs(L, c[SC_OR_LO] ^ c[SC_AND_LO], c[SC_OR_HI] ^ c[SC_AND_HI]);
        uint64_t vlo = 0, vhi = 0;  // bits covered by the runs
        for (int i = 0; i < L.n_runs; ++i)
            for (uint32_t b = L.runs[i].pos; b < L.runs[i].pos + L.runs[i].len; ++b) {
                if (b < 64) vlo |= 1ull << b;
                else vhi |= 1ull << (b - 64);
            }
        L.vmask_lo = vlo;
        L.vmask_hi = vhi;
        L.const_lo = c[SC_AND_LO] & ~vlo;
        L.const_hi = c[SC_AND_HI] & ~vhi;
        width = 0;
        for (int i = 0; i < L.n_runs; ++i) width += L.runs[i].len;
        full_passes = (int)((width + kDigitBits - 1) / kDigitBits);
        // the records carry the values where the fast scan encodes them and
        // the fast fold reads them (one or two value columns), and where the
        // scatter's staging of them fits its LDS (not two-word keys with two
        // value columns)
        L.carry = L.fast && P.q.n_vals >= 1 && P.q.n_vals <= 2 && !(L.kw == 2 && P.q.n_vals == 2) &&
                          plan.opts.sort_carry != 0
                      ? 1
                      : 0;
        // MSD finish: when the records split into buckets of about kBucketMean
        // by their top tbits bits with fewer passes than the whole key needs,
        // the passes order only those bits and LDS sorts finish each bucket
        tbits = 0;
        bmean = L.kw == 2 ? kBucketMean / 2 : kBucketMean;
        if (n_sel >= 2 * bmean && plan.opts.sort_buckets != 0) {
            int t = 1;
            while (t < 40 && (n_sel >> t) > bmean) ++t;
            t = std::max(t, kDigitBits);
            if ((t + kDigitBits - 1) / kDigitBits < full_passes) tbits = t;
        }
        plan_shifts();
        // ---- buffers
        const uint64_t cap = std::max<uint64_t>(n_sel, 1);
        L.ntiles = (uint32_t)std::max<uint64_t>(1, ceil_div(n_sel, kSortTile));
        L.ka_lo = (uint64_t *)slab("so:ka_lo", cap * 8 + 64);
        L.kb_lo = (uint64_t *)slab("so:kb_lo", cap * 8 + 64);
        L.ka_hi = L.kw == 2 ? (uint64_t *)slab("so:ka_hi", cap * 8 + 64) : nullptr;
        L.kb_hi = L.kw == 2 ? (uint64_t *)slab("so:kb_hi", cap * 8 + 64) : nullptr;
        if (L.carry) {
            for (int j = 0; j < P.q.n_vals; ++j) {
                L.va[j] = (uint64_t *)slab("so:va" + std::to_string(j), cap * 8 + 64);
                L.vb[j] = (uint64_t *)slab("so:vb" + std::to_string(j), cap * 8 + 64);
            }
            L.fa = (uint8_t *)slab("so:fa", cap + 64);
            L.fb = (uint8_t *)slab("so:fb", cap + 64);
        } else {
            L.ra = (uint32_t *)slab("so:ra", cap * 4 + 64);
            L.rb = (uint32_t *)slab("so:rb", cap * 4 + 64);
        }
        L.tile_hist = (uint32_t *)slab("so:tile_hist", (uint64_t)kDigits * L.ntiles * 4 + 64);
        L.digit_tot = (uint32_t *)slab("so:digit_tot", kDigits * 4 + 64);
        L.tile_cnt = (uint32_t *)slab("so:tile_cnt", ((uint64_t)L.ntiles + 1) * 4 + 64);
        PMH_HIP(hipMemcpyAsync(dL, &L, sizeof(SLaunch), hipMemcpyHostToDevice, R.stream));
    };
    bits_pass(L.fast && L.q.n_pred == 0 && L.q.n_rows >= (uint64_t(1) << 20) && plan.opts.sort_sample_bits != 0);
    plan_layout();
    int passes = (int)shifts.size();

    // group state and compressed keys for slots group slots
    auto group_state = [&](uint64_t slots) {
        L.gcap = std::max<uint64_t>(slots, 1);
        L.gk_lo = (uint64_t *)slab("so:gk_lo", L.gcap * 8 + 64);
        L.gk_hi = (uint64_t *)slab("so:gk_hi", L.gcap * 8 + 64);
        uint8_t *hb = (uint8_t *)slab("so:state", state_bytes(P, L.gcap) + 4096);
        size_t o = 0;
        carve_state(P, hb, o, L.gcap, L.arr);
    };
    bool exact = false;  // some aggregated column keeps exact f64 sums (limbs and residual)
    for (int j = 0; j < P.q.n_vals; ++j)
        if (P.q.vals[j].kind == VK_F && (P.q.vals[j].need & NEED_SUM)) exact = true;
    bool fused = false;  // the bucket fold ran: group state at sparse indices, emitted by bucket
    // exclusive prefix of n counts and their total (into SC_GROUPS)
    auto scan_counts = [&](const uint32_t *cnt, uint32_t *pre, uint32_t n) {
        const uint32_t blocks = (uint32_t)std::max<uint64_t>(1, ceil_div((uint64_t)n, 4096));
        uint32_t *bsum = (uint32_t *)slab("so:csum", (uint64_t)blocks * 4 + 64);
        hipLaunchKernelGGL(so_count_sum, dim3(blocks), dim3(1024), 0, R.stream, cnt, n, bsum);
        launch_check();
        hipLaunchKernelGGL(so_count_scan, dim3(blocks), dim3(1024), 0, R.stream, cnt, pre, n, (const uint32_t *)bsum,
                           L.ctr + SC_GROUPS);
        launch_check();
    };
    uint64_t *bstart = nullptr;
    uint32_t *bg = nullptr, *bpre = nullptr;
    uint32_t nb = 0;

    // ---- 2. records, 3. radix passes
    const int pgrid = cus * 2;
    uint64_t *slo = nullptr, *shi = nullptr, *dlo = nullptr, *dhi = nullptr;
    Payload spl{}, dpl{};
    if (n_sel) {
        for (;;) {
            slo = L.ka_lo;
            shi = L.ka_hi;
            dlo = L.kb_lo;
            dhi = L.kb_hi;
            spl = Payload{L.ra, {L.va[0], L.va[1]}, L.fa};
            dpl = Payload{L.rb, {L.vb[0], L.vb[1]}, L.fb};
            if (L.fast) {
                KernelTimer kt("so_encode_fast");
                hipLaunchKernelGGL(so_scan_fast<true>, dim3(cus * 8), dim3(512), 0, R.stream, dL);
                launch_check();
            } else {
                KernelTimer kt("so_encode");
                hipLaunchKernelGGL(so_encode, dim3(cus * 8), dim3(256), 0, R.stream, dL);
                launch_check();
            }
            // skewed keys would overflow the buckets: probe a sample first
            if (tbits) {
                hipLaunchKernelGGL(so_bucket_probe, dim3(1), dim3(1024), 0, R.stream, dL, slo, shi, (int)width - tbits);
                launch_check();
            }
            if (!tbits && !L.sampled) break;
            uint64_t pc[SC_N];
            PMH_HIP(hipMemcpyAsync(pc, L.ctr, sizeof(pc), hipMemcpyDeviceToHost, R.stream));
            PMH_HIP(hipStreamSynchronize(R.stream));
            if (L.sampled && pc[SC_MISS]) {
                // a record with a lane bit the sample did not see vary
                bits_pass(false);
                plan_layout();
                passes = (int)shifts.size();
                continue;
            }
            if (tbits) {
                const uint64_t S = std::min<uint64_t>(n_sel, kProbeSample);
                const double expect = (double)S * bucket_cap(L.kw == 2) / (double)n_sel;  // sampled rows of a full bucket
                if ((double)pc[SC_BIG] > std::max(16.0, 4.0 * expect)) {
                    tbits = 0;
                } else if (S >= 4 * (uint64_t(1) << tbits) && pc[SC_PROBE_D]) {
                    // a sample this dense meets every populated bucket: keys that
                    // fill only part of the top bits' range populate fewer
                    // buckets, each larger than the mean.  More top bits at the
                    // same number of passes split them (about in half per bit).
                    const int np = (tbits + kDigitBits - 1) / kDigitBits;
                    double per = (double)n_sel / (double)pc[SC_PROBE_D];
                    while (per > (double)bmean && tbits + 1 <= (int)width &&
                           (tbits + kDigitBits) / kDigitBits == np) {
                        ++tbits;
                        per /= 2;
                    }
                }
                plan_shifts();
                passes = (int)shifts.size();
                hipLaunchKernelGGL(so_reset_big, dim3(1), dim3(64), 0, R.stream, dL);
                launch_check();
            }
            break;
        }
        const size_t dyn = (L.kw == 2 ? (size_t)kSortTile * 8 : 0) +
                           (L.carry ? (size_t)kSortTile * (8 * (size_t)P.q.n_vals + 1) : (size_t)kSortTile * 4);
        const void *scatter_fn = L.carry ? (const void *)so_scatter<true> : (const void *)so_scatter<false>;
        PMH_HIP(hipFuncSetAttribute(scatter_fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)dyn));
        auto run_passes = [&](const std::vector<int> &sh) {
            for (const int shift : sh) {
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
                    int sh = shift;
                    void *args[] = {(void *)&dL, (void *)&slo, (void *)&shi, (void *)&spl,
                                    (void *)&dlo, (void *)&dhi, (void *)&dpl, (void *)&sh};
                    PMH_HIP(hipLaunchKernel(scatter_fn, dim3(pgrid), dim3(kSortThreads), args, dyn, R.stream));
                    launch_check();
                }
                std::swap(slo, dlo);
                std::swap(shi, dhi);
                std::swap(spl, dpl);
            }
        };
        // MSD top-bit passes (tbits <= 16): the first over the top tbits - 8
        // bits, the second over the next 8 inside each first-pass bucket, whose
        // starts are then the buckets' starts
        const bool msd = tbits && tbits <= 2 * kDigitBits && plan.opts.sort_msd != 0;
        auto run_msd = [&]() {
            nb = 1u << tbits;
            bstart = (uint64_t *)slab("so:bstart", ((uint64_t)nb + 1) * 8 + 64);
            const bool one = tbits <= kDigitBits;
            const uint32_t nd1 = one ? kDigits : 1u << (tbits - kDigitBits);
            uint32_t *xtot = (uint32_t *)slab("so:xtot", (uint64_t)nb * 8 * 4 + 64);
            uint64_t *xcur = (uint64_t *)slab("so:xcur", (uint64_t)nb * 8 * 8 + 64);
            uint32_t *bsum = (uint32_t *)slab("so:bsum", ceil_div((uint64_t)nb, 1024) * 4 + 64);
            uint64_t *st1 = one ? bstart : (uint64_t *)slab("so:st1", ((uint64_t)nd1 + 1) * 8 + 64);
            uint32_t *toff = (uint32_t *)slab("so:toff", ((uint64_t)nd1 + 1) * 4 + 64);
            uint32_t *treg = (uint32_t *)slab("so:treg", ((uint64_t)L.ntiles + nd1 + 1) * 4 + 64);
            static const void *const kScatter[2][2] = {
                {(const void *)so_msd_scatter<false, false>, (const void *)so_msd_scatter<false, true>},
                {(const void *)so_msd_scatter<true, false>, (const void *)so_msd_scatter<true, true>}};
            const void *fn = kScatter[L.carry ? 1 : 0][L.kw == 2 ? 1 : 0];
            PMH_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)dyn));
            auto pass = [&](MsdPass M, uint32_t nbk, uint64_t *starts) {
                const uint32_t np = nbk ? nbk * M.nd : M.nd;
                const uint32_t sblocks = (uint32_t)ceil_div((uint64_t)np, 1024);
                {
                    KernelTimer kt("so_msd_hist");
                    hipLaunchKernelGGL(so_msd_hist, dim3(pgrid), dim3(kSortThreads), 0, R.stream, dL, slo, shi, M,
                                       nbk);
                    launch_check();
                }
                hipLaunchKernelGGL(so_msd_sum, dim3(sblocks), dim3(1024), 0, R.stream, M.xtot, np, bsum);
                launch_check();
                hipLaunchKernelGGL(so_msd_scan, dim3(sblocks), dim3(1024), 0, R.stream, M.xtot, M.xcur, np, bsum,
                                   starts);
                launch_check();
                {
                    KernelTimer kt("so_msd_scatter");
                    void *args[] = {(void *)&dL, (void *)&slo, (void *)&shi, (void *)&spl, (void *)&dlo,
                                    (void *)&dhi, (void *)&dpl, (void *)&M,   (void *)&nbk};
                    PMH_HIP(hipLaunchKernel(fn, dim3(pgrid), dim3(kSortThreads), args, dyn, R.stream));
                    launch_check();
                }
                std::swap(slo, dlo);
                std::swap(shi, dhi);
                std::swap(spl, dpl);
            };
            // pass 1: every tile of the records (the counts cleared first)
            hipLaunchKernelGGL(so_msd_tiles, dim3(1), dim3(1024), 0, R.stream, (const uint64_t *)nullptr, 0u,
                               (uint32_t *)nullptr, (uint32_t *)nullptr, xtot, nd1 * 8);
            launch_check();
            MsdPass M1{};
            M1.shift = (int)width - (one ? kDigitBits : tbits - kDigitBits);
            M1.nd = nd1;
            M1.xtot = xtot;
            M1.xcur = xcur;
            M1.ntiles = L.ntiles;
            pass(M1, 0, st1);
            if (one) return;
            // pass 2: tiles inside each first-pass bucket
            const uint32_t nclear = nb * 8;
            hipLaunchKernelGGL(so_msd_tiles, dim3((uint32_t)std::min<uint64_t>(1024, ceil_div(nclear, 1024))),
                               dim3(1024), 0, R.stream, (const uint64_t *)st1, nd1, toff, treg, xtot, nclear);
            launch_check();
            MsdPass M2{};
            M2.shift = (int)width - tbits;
            M2.nd = kDigits;
            M2.second = 1;
            M2.bst = st1;
            M2.toff = toff;
            M2.treg = treg;
            M2.xtot = xtot;
            M2.xcur = xcur;
            pass(M2, nd1, bstart);
        };
        if (msd) run_msd();
        else run_passes(shifts);
        if (tbits) {
            // ---- 3b. bucket finish
            if (!msd) {
                nb = 1u << tbits;
                bstart = (uint64_t *)slab("so:bstart", ((uint64_t)nb + 1) * 8 + 64);
                KernelTimer kt("so_bucket_bounds");
                hipLaunchKernelGGL(so_bucket_bounds, dim3((uint32_t)ceil_div((uint64_t)nb + 1, 256)), dim3(256), 0,
                                   R.stream, dL, slo, shi, bstart, nb, (int)width - tbits);
                launch_check();
            }
            const bool two = L.kw == 2;
            const size_t lds = bucket_lds_bytes(two);
            const uint32_t grid = (uint32_t)std::min<uint64_t>(nb, (uint64_t)cus * (lds * 2 <= 160 * 1024 ? 2 : 1));
            // the bins: the kBinBits key bits below the bucket's top bits
            // (fewer when the key has fewer bits left)
            int bin_bits = std::min<int>(kBinBits, (int)width - tbits);
            int bin_shift = (int)width - tbits - bin_bits;
            fused = L.fast && P.q.n_vals <= 2 && plan.opts.sort_fold != 0;
            if (fused) {
                // group state and keys at sparse indices: one slot per record
                group_state(n_sel);
                PMH_HIP(hipMemcpyAsync(dL, &L, sizeof(SLaunch), hipMemcpyHostToDevice, R.stream));
                bg = (uint32_t *)slab("so:bg", (uint64_t)nb * 4 + 64);
                bpre = (uint32_t *)slab("so:bpre", (uint64_t)nb * 4 + 64);
                {
                    KernelTimer kt("so_bucket_fold");
                    static const void *const kFold[2][2][2] = {
                        {{(const void *)so_bucket_fold<false, false, false>,
                          (const void *)so_bucket_fold<false, false, true>},
                         {(const void *)so_bucket_fold<false, true, false>,
                          (const void *)so_bucket_fold<false, true, true>}},
                        {{(const void *)so_bucket_fold<true, false, false>,
                          (const void *)so_bucket_fold<true, false, true>},
                         {(const void *)so_bucket_fold<true, true, false>,
                          (const void *)so_bucket_fold<true, true, true>}}};
                    const void *fn = kFold[two ? 1 : 0][L.carry ? 1 : 0][exact ? 1 : 0];
                    const size_t flds = lds;
                    PMH_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)flds));
                    void *args[] = {(void *)&dL, (void *)&slo, (void *)&shi, (void *)&spl, (void *)&bstart,
                                    (void *)&nb, (void *)&bin_shift, (void *)&bin_bits, (void *)&bg};
                    const uint32_t fgrid =
                        (uint32_t)std::min<uint64_t>(nb, (uint64_t)cus * (flds * 2 <= 160 * 1024 ? 2 : 1));
                    PMH_HIP(hipLaunchKernel(fn, dim3(fgrid), dim3(kBucketThreads), args, flds, R.stream));
                    launch_check();
                }
                scan_counts(bg, bpre, nb);
            } else {
                KernelTimer kt("so_bucket_sort");
                static const void *const kBucket[2][2] = {
                    {(const void *)so_bucket_sort<false, false>, (const void *)so_bucket_sort<false, true>},
                    {(const void *)so_bucket_sort<true, false>, (const void *)so_bucket_sort<true, true>}};
                const void *fn = kBucket[two ? 1 : 0][L.carry ? 1 : 0];
                PMH_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
                void *args[] = {(void *)&dL, (void *)&slo, (void *)&shi, (void *)&spl, (void *)&dlo, (void *)&dhi,
                                (void *)&dpl, (void *)&bstart, (void *)&nb, (void *)&bin_shift, (void *)&bin_bits};
                PMH_HIP(hipLaunchKernel(fn, dim3(grid), dim3(kBucketThreads), args, lds, R.stream));
                launch_check();
                // the sorted records are in the other buffer
                std::swap(slo, dlo);
                std::swap(shi, dhi);
                std::swap(spl, dpl);
            }
        }
        // ---- 4. group starts
        auto run_bounds = [&]() {
            {
                KernelTimer kt("so_bound");
                hipLaunchKernelGGL(so_bound, dim3(pgrid), dim3(256), 0, R.stream, dL, slo, shi);
                launch_check();
            }
            scan_counts(L.tile_cnt, L.tile_cnt, L.ntiles);
        };
        if (!fused) run_bounds();
        PMH_HIP(hipMemcpyAsync(c, L.ctr, sizeof(c), hipMemcpyDeviceToHost, R.stream));
        PMH_HIP(hipStreamSynchronize(R.stream));
        if (c[SC_BIG]) {
            // a bucket outgrew the LDS sort (keys concentrated in a few top
            // bit patterns): the whole key by LSD passes after all, from the
            // buffer the bucket kernel read (the bucket sort skipped the
            // large buckets, so only that buffer holds every record; the
            // fold writes no records)
            if (!fused) {
                std::swap(slo, dlo);
                std::swap(shi, dhi);
                std::swap(spl, dpl);
            }
            fused = false;
            std::vector<int> all;
            for (int p = 0; p < full_passes; ++p) all.push_back(p * kDigitBits);
            run_passes(all);
            passes += full_passes;
            tbits = 0;
            run_bounds();
        }
    }
    PMH_HIP(hipMemcpyAsync(c, L.ctr, sizeof(c), hipMemcpyDeviceToHost, R.stream));
    PMH_HIP(hipStreamSynchronize(R.stream));
    const uint64_t ng = n_sel ? c[SC_GROUPS] : 0;

    // ---- group state (unless the fold placed it) and outputs, sized for ng
    if (!fused) group_state(ng);
    const uint64_t ocap = std::max<uint64_t>(ng, 1);
    for (int oi = 0; oi < P.n_out; ++oi) {
        OutCol &oc = L.out[oi];
        oc = P.outs[oi];
        const size_t esz = std::max(1, tag_size(oc.tag));
        oc.data = slab("so:out" + std::to_string(oi), ocap * esz + 64);
        const bool need_valid = !(oc.src == OS_AGG && (oc.op == AGG_LEN || oc.op == AGG_COUNT));
        oc.vbytes = need_valid ? (uint8_t *)slab("so:outvb" + std::to_string(oi), ocap + 64) : nullptr;
        oc.valid = need_valid ? (uint64_t *)slab("so:outv" + std::to_string(oi), ocap / 8 + 64) : nullptr;
        oc.data_bits = (oc.src == OS_KEY && oc.tag == PMH_BOOL)
                           ? (uint8_t *)slab("so:outbits" + std::to_string(oi), ocap / 8 + 64)
                           : nullptr;
    }
    PMH_HIP(hipMemcpyAsync(dL, &L, sizeof(SLaunch), hipMemcpyHostToDevice, R.stream));
    if (ng && !fused) {
        {
            KernelTimer kt("so_init_state");
            hipLaunchKernelGGL(so_init_state, dim3((uint32_t)std::min<uint64_t>(ceil_div(ng, 256), cus * 16)),
                               dim3(256), 0, R.stream, dL, ng);
            launch_check();
        }
        if (L.fast && P.q.n_vals <= 2) {
            KernelTimer kt("so_reduce_fast");
            static const void *const kReduce[2][2] = {
                {(const void *)so_reduce_fast<false, false>, (const void *)so_reduce_fast<false, true>},
                {(const void *)so_reduce_fast<true, false>, (const void *)so_reduce_fast<true, true>}};
            const void *fn = kReduce[L.carry ? 1 : 0][exact ? 1 : 0];
            void *args[] = {(void *)&dL, (void *)&slo, (void *)&shi, (void *)&spl};
            PMH_HIP(hipLaunchKernel(fn, dim3((L.ntiles + 3) / 4), dim3(256), args, 0, R.stream));
            launch_check();
        } else {
            KernelTimer kt("so_reduce");
            hipLaunchKernelGGL(so_reduce, dim3(pgrid), dim3(kSortThreads), 0, R.stream, dL, slo, shi, spl.row);
            launch_check();
        }
    }
    if (ng) {
        if (fused) {
            KernelTimer kt("so_emit");
            hipLaunchKernelGGL(so_emit_buckets, dim3((uint32_t)std::min<uint64_t>(ceil_div(nb, 4), (uint64_t)cus * 16)), dim3(256),
                               0, R.stream, dL, bstart, bg, bpre, nb);
            launch_check();
        } else {
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

    pmh_result *res = new_result(P.n_out);
    res->n_rows = ng;
    res->stats.method = PMH_METHOD_SORT;
    res->stats.n_passes = passes;
    res->stats.initial_capacity = std::max<uint64_t>(n_sel, 1);
    res->stats.final_capacity = std::max<uint64_t>(n_sel, 1);
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
            if (nb_copy) download_host(col.data, dsrc, nb_copy);
            if (oc.valid) {
                if (packed) download_host(col.valid, oc.valid, packed);
            } else {
                memset(col.valid, 0xFF, vbytes);  // counts / lens are never null
            }
        }
    }
    // host outputs wait for their copies; device outputs were complete at ev1
    if (!(flags & PMH_OUT_DEVICE)) PMH_HIP(hipStreamSynchronize(R.stream));
    collect_kernel_times();
    return res;
}

}  // namespace pmh
