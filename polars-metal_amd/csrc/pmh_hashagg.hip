// pmh_hashagg.hip -- partitioned hash aggregation, the default hash path.
//
// Global atomics to scattered addresses run at a few tens of Gops/s on
// MI355X while LDS atomics run at Tops/s, so every per-row update happens in
// LDS and global memory only sees coalesced streams:
//
//   hs_sample_gather  keys of a strided row sample that pass the predicate
//   hs_sample_count   one block counts the sample in LDS and lays the most
//                     frequent keys out in a hot table shared by every block
//                     (two-choice buckets of four slots, the most sampled
//                     keys placed first); it also records the sample sketch
//   hs_scan_fast      one block per CU streams a contiguous chunk of rows:
//                     predicates lowered to at most two intervals on an
//                     order-preserving int64 image, columns read as aligned
//                     8-byte words, selected rows compacted through a per-wave
//                     LDS ring so probes and folds run on dense waves,
//                     hot-table probe (read only, no CAS).  Hot rows fold into
//                     LDS state; the other rows ("tail") are appended,
//                     wave-contiguously, to the block's spill region.  At the
//                     end the block adds its hot partials into the global hot
//                     state with coalesced atomics (slot s of every block maps
//                     to global slot s).  hs_scan is the general form (any
//                     predicate tree, a register row cache).
//   rp_hist1/rp_scan1/rp_scatter1
//                     up to 1024 partitions: one pass groups the tail by
//                     partition (the low bits of the key hash); tiles rank
//                     rows with LDS counters, and each partition is laid out
//                     as eight per-XCD sub-regions so the runs that meet in a
//                     sub-region merge in one XCD's L2.
//   rp_tiles2/rp_hist2/rp_sum2/rp_scan2/rp_scatter2
//                     up to 2^17 partitions: the unstable pass takes the low
//                     bits, then a second unstable pass takes the next 7
//                     inside each first-pass bucket (tiles never cross one),
//                     with the same per-XCD sub-regions; its counts give the
//                     partition starts.
//   rp_hist/rp_scan/rp_scatter
//                     the general multi-pass form: stable LSD passes (7 bits
//                     each) after the unstable one: histogram per tile, one
//                     scan per digit, then a scatter that ranks rows with
//                     wave multisplit ballots, stages the tile in LDS in
//                     digit order and writes each digit run contiguously;
//                     hs_part_bounds then finds the partition starts.
//   hs_part_agg       one block per partition folds the partition's tail rows
//                     into an LDS table and writes its groups straight to the
//                     outputs; a row that finds no LDS slot goes to the global
//                     fallback table (written by hs_fallback_emit).
//   hs_hot_emit       writes the hot groups.
//   hs_pack           turns the per-row validity bytes into Arrow bitmaps.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "pmh_mlp.h"
#include "pmh_mlp_dev.h"
#include "pmh_table.h"

namespace pmh {

constexpr int kRadixBits = 7;  // stable passes
constexpr int kRadix = 1 << kRadixBits;
constexpr int kPartThreads = 1024;
constexpr int kPartRows = 4;               // rows per thread per partition tile
constexpr int kPartTile = kPartThreads * kPartRows;
constexpr int kWideRadixBits = 10;  // one-pass partitioning: up to 1024 partitions
constexpr int kWideRadix = 1 << kWideRadixBits;
// At most 2^17 tail partitions: two passes (an unstable 10-bit one, then a
// stable 7-bit one) reach them, and their hash bits stay below bit 20, where
// the per-partition table's slot bits begin.
constexpr int kMaxPartBits = kWideRadixBits + kRadixBits;
static_assert(kMaxPartBits <= 20, "partition bits and slot bits of the hash are disjoint");
static_assert(kWideRadix == kPartThreads, "one thread per digit in the block scans");
constexpr uint16_t kFlagAllOnes = 1u << 8, kFlagNull = 1u << 9;
constexpr int kProbe = 16;

struct Spill {
    uint64_t *klo, *khi;  // khi: wide keys only
    void *val[kMaxVals];  // value j as raw bits of width vw[j]
    uint16_t *flags;      // bit j: value j valid; kFlagAllOnes / kFlagNull: side keys
};

enum StatIdx : int32_t {
    ST_SELECTED = 0, ST_TAIL = 1, ST_SAMPLE_SEL = 2, ST_SAMPLE_DISTINCT = 3, ST_SAMPLE_F1 = 4,
    ST_SAMPLE_F2 = 5, ST_HOT = 6, ST_HOT_SAMPLE_ROWS = 7, ST_TILES0 = 8, ST_SAMPLE_NULLKEY = 9,
    ST_SAMPLE_DROPPED = 10, ST_ERROR = 11, ST_NARROW_MISS = 12, ST_N = 16
};
// The sizer block follows the stats (words): the sketch's features, the
// installed model's two outputs and the plain distinct estimate (f64), so
// the host plans from one read-back of ST_N + SZ_N words.  The count
// histogram follows it.
enum SizerIdx : int32_t {
    SZ_FEAT = 0, SZ_PRED = PMH_MLP_IN, SZ_EST = PMH_MLP_IN + 2, SZ_HAVE = PMH_MLP_IN + 4, SZ_SCALE = PMH_MLP_IN + 5,
    SZ_N = 32
};
static_assert(SZ_SCALE < SZ_N && (ST_N + SZ_EST) % 2 == 0, "sizer block layout");

struct HLaunch {
    QueryDesc q;
    int32_t wide, n_out;
    int32_t vw[kMaxVals];
    int32_t col_val[kMaxCols];
    // sample and hot set
    uint32_t n_sample, hot_slots, hot_max, pad0;
    uint64_t *samp_lo, *samp_hi;
    uint8_t *samp_sp;               // side-key marker of each sampled key
    uint64_t *skey_lo, *skey_hi;    // key of each fingerprint slot (written by the row that claims it)
    uint32_t sample_slots;          // LDS fingerprint-table size of hs_sample_count
    uint32_t pad2;
    uint64_t *hot_lo, *hot_hi;
    void *hot_arr[kMaxArrs];
    uint32_t lds1_off[kMaxArrs];
    // scan
    uint64_t chunk;
    uint32_t nb;
    uint32_t *tail_n;      // [nb]
    uint32_t *region_off;  // [nb + 1]
    uint32_t *tile_off;    // [nb + 1] first-pass tiles per region, prefix
    Spill spill;
    // partition passes
    Spill ping, pong;
    uint32_t *tile_hist;   // [kRadix][max_tiles]
    uint32_t *digit_base;  // [kRadix]
    uint64_t *part_start;  // [partitions + 1]: first spill row of each partition after the passes
    uint32_t *tile_reg;    // [max_tiles]: scan region of each first-pass tile
    uint32_t *xtot;        // one-pass partitioning: [8][kWideRadix] rows per (XCD slot, digit)
    uint64_t *xcur;        // [8][kWideRadix]: next free row of each (XCD slot, digit) sub-region
    // two-pass layout, second pass (MSD inside the first pass's buckets)
    uint32_t *xtot2;       // [partitions][8]: rows per (partition, XCD slot)
    uint64_t *xcur2;       // [partitions][8]: next free row of each (partition, XCD slot) sub-region
    uint64_t *bstart2;     // [buckets + 1]: first-pass bucket starts
    uint32_t *tile_off2;   // [buckets + 1]: first tile of each bucket
    uint32_t *tile_reg2;   // [tiles]: bucket of each tile
    uint32_t *bsum2;       // [partitions / 1024]: rows per block of 1024 partitions
    uint32_t max_tiles, pbits, sortbits, passes;
    // per-partition aggregation
    uint32_t tail_slots;
    uint32_t live_off;  // byte offset in hs_part_agg's LDS of its dense live-slot list (0: none)
    uint32_t lds3_off[kMaxArrs];
    GTable t;
    // outputs
    OutCol out[kOutMax];
    uint32_t *out_cursor;  // [0] groups written by hs_hot_emit and hs_part_agg
    uint32_t *stats;       // StatIdx, then the sizer block (SizerIdx)
    uint32_t *sketch_hist; // [1024]: sampled keys seen c times (c clipped at 1023)
    const float *sizer;    // the installed sizer's parameters, or null: no forward in the sketch block
    int32_t lane_bits, back_words;  // back_words: stats words the sketch hands the host
    uint32_t *mbox;        // the host's pinned mailbox (device address): read-backs land here
    FastScan fs;
};

// Digit of a spilled key for partitioning: side keys go to partition 0.
__device__ __forceinline__ uint64_t spill_hash(uint64_t lo, uint64_t hi, uint16_t flags) {
    return (flags & (kFlagAllOnes | kFlagNull)) ? 0 : hash_key(lo, hi);
}

__device__ __forceinline__ uint64_t load_w(const void *p, int w, uint64_t i) {
    switch (w) {
    case 8: return ((const uint64_t *)p)[i];
    case 4: return ((const uint32_t *)p)[i];
    case 2: return ((const uint16_t *)p)[i];
    default: return ((const uint8_t *)p)[i];
    }
}
__device__ __forceinline__ void store_w(void *p, int w, uint64_t i, uint64_t x) {
    switch (w) {
    case 8: ((uint64_t *)p)[i] = x; break;
    case 4: ((uint32_t *)p)[i] = (uint32_t)x; break;
    case 2: ((uint16_t *)p)[i] = (uint16_t)x; break;
    default: ((uint8_t *)p)[i] = (uint8_t)x; break;
    }
}

// Spill arrays live in global memory.
__device__ __forceinline__ uint64_t gload_w(const void *p, int w, uint64_t i) {
    switch (w) {
    case 8: return gld<uint64_t>(p, i);
    case 4: return gld<uint32_t>(p, i);
    case 2: return gld<uint16_t>(p, i);
    default: return gld<uint8_t>(p, i);
    }
}
__device__ __forceinline__ uint64_t gload_w_nt(const void *p, int w, uint64_t i) {
    switch (w) {
    case 8: return gld_nt<uint64_t>(p, i);
    case 4: return gld_nt<uint32_t>(p, i);
    case 2: return gld_nt<uint16_t>(p, i);
    default: return gld_nt<uint8_t>(p, i);
    }
}
__device__ __forceinline__ void gstore_w(void *p, int w, uint64_t i, uint64_t x) {
    switch (w) {
    case 8: gst<uint64_t>(p, i, x); break;
    case 4: gst<uint32_t>(p, i, (uint32_t)x); break;
    case 2: gst<uint16_t>(p, i, (uint16_t)x); break;
    default: gst<uint8_t>(p, i, (uint8_t)x); break;
    }
}

// Row accessor over a spilled record: value columns only.
// A spill row whose first two value columns are already loaded.
struct SpillRowV {
    const HLaunch *__restrict__ L;
    const Spill *S;
    uint64_t i;
    uint16_t flags;
    uint64_t v0, v1;
    __device__ __forceinline__ bool valid(int c) const { return (flags >> L->col_val[c]) & 1; }
    __device__ __forceinline__ uint64_t bits(int c) const {
        const int j = L->col_val[c];
        return j == 0 ? v0 : j == 1 ? v1 : gload_w(S->val[j], L->vw[j], i);
    }
    __device__ __forceinline__ int64_t i64(int c) const { return raw_to_i64(L->q.cols[c].tag, bits(c)); }
    __device__ __forceinline__ double f64(int c) const { return raw_to_f64(L->q.cols[c].tag, bits(c)); }
};

struct SpillRow {
    const HLaunch *__restrict__ L;
    const Spill *S;
    uint64_t i;
    uint16_t flags;
    __device__ __forceinline__ bool valid(int c) const { return (flags >> L->col_val[c]) & 1; }
    __device__ __forceinline__ uint64_t bits(int c) const {
        const int j = L->col_val[c];
        return gload_w(S->val[j], L->vw[j], i);
    }
    __device__ __forceinline__ int64_t i64(int c) const { return raw_to_i64(L->q.cols[c].tag, bits(c)); }
    __device__ __forceinline__ double f64(int c) const { return raw_to_f64(L->q.cols[c].tag, bits(c)); }
};


// ---------------------------------------------------------------------------
// sample
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void hs_sample_gather(const HLaunch *__restrict__ L) {
    const QueryDesc *__restrict__ q = &L->q;
    const uint64_t n = q->n_rows;
    const uint32_t ns = L->n_sample;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    bool take = false;
    Key k{0, 0, 0};
    if (i < ns && n > 0) {
        // row = floor(i * n / ns) without 128-bit division (i, n % ns < 2^32)
        const uint64_t r = (uint64_t)i * (n / ns) + ((uint64_t)i * (n % ns)) / ns;
        MemRow row{q, r};
        if (eval_pred(q, row)) {
            k = load_key(q, row);
            take = true;
        }
    }
    // wave-aggregated append: one atomic per wave
    const unsigned long long m = __ballot(take);
    const int lane = threadIdx.x & 63;
    uint32_t base = 0;
    const int leader = m ? __ffsll((long long)m) - 1 : 0;
    if (lane == leader && m) base = atomicAdd(&L->stats[ST_SAMPLE_SEL], (uint32_t)__popcll(m));
    base = __shfl(base, leader);
    if (take) {
        const uint32_t pos = base + (uint32_t)__popcll(m & ((1ull << lane) - 1));
        L->samp_lo[pos] = k.lo;
        L->samp_hi[pos] = k.hi;
        L->samp_sp[pos] = (uint8_t)k.special;
    }
}

// One block: count the sampled keys by 32-bit fingerprint in LDS, pick the
// most frequent fingerprints (up to hot_max), recover their keys from the
// sample, lay them out in the hot table, and record the sketch (sampled rows,
// distinct fingerprints, f1, f2).  It also resets the global hot state.
// The hot table (CL slots, a power of two, CL >= 16) is read only during the
// scan, so it is laid out once for cheap, bounded lookups: buckets of four
// slots, the first CL/4 - 1 of them for hashed hot keys.  A key goes to the
// first free slot of its first bucket, or of its second when the first was
// full; a key that finds both full stays out of the hot set (its rows take
// the tail path).  Buckets only fill, so a lookup that sees a free slot in
// the key's first bucket stops there.  Slot CL - 2 is the group of the
// all-ones key and slot CL - 1 the group of the null key of a single 64-bit
// key column, the two keys the empty marker cannot stand for.
__device__ __forceinline__ uint32_t hot_buckets(uint32_t CL) { return CL / 4 - 1; }
__device__ __forceinline__ uint32_t hot_b1(uint64_t h, uint32_t nb) { return (uint32_t)(((h >> 32) * nb) >> 32); }
__device__ __forceinline__ uint32_t hot_b2(uint64_t h, uint32_t nb) {
    return (uint32_t)(((h & 0xFFFFFFFFull) * nb) >> 32);
}
// Slot of a (non-special) 64-bit key in an LDS copy of the hot keys, or -1.
// Both buckets are read up front (four 16-byte LDS loads behind one wait)
// and the slot is chosen with selects: the first match in the first bucket,
// else -1 when that bucket has a free slot, else the first match in the
// second.  lk must be 16-byte aligned (buckets start at multiples of four).
typedef unsigned long long hot_pair __attribute__((ext_vector_type(2)));
__device__ __forceinline__ int32_t hot_find(const uint64_t *lk, uint64_t key, uint64_t h, uint32_t CL) {
    const uint32_t nb = hot_buckets(CL);
    const uint32_t b1 = hot_b1(h, nb) * 4, b2 = hot_b2(h, nb) * 4;
    const hot_pair *p1 = (const hot_pair *)(lk + b1), *p2 = (const hot_pair *)(lk + b2);
    const hot_pair x0 = p1[0], x1 = p1[1], y0 = p2[0], y1 = p2[1];
    int32_t s = -1;
    s = y1.y == key ? (int32_t)b2 + 3 : s;
    s = y1.x == key ? (int32_t)b2 + 2 : s;
    s = y0.y == key ? (int32_t)b2 + 1 : s;
    s = y0.x == key ? (int32_t)b2 : s;
    const bool open = (x0.x == kEmpty) | (x0.y == kEmpty) | (x1.x == kEmpty) | (x1.y == kEmpty);
    s = open ? -1 : s;
    s = x1.y == key ? (int32_t)b1 + 3 : s;
    s = x1.x == key ? (int32_t)b1 + 2 : s;
    s = x0.y == key ? (int32_t)b1 + 1 : s;
    s = x0.x == key ? (int32_t)b1 : s;
    return s;
}
__device__ __forceinline__ uint32_t hot_side_slot(int special, uint32_t CL) { return CL - 3 + (uint32_t)special; }

// Probe limit of the sample fingerprint table: a key that finds no slot in
// this many probes is dropped (and counted), which keeps a sample with more
// distinct keys than slots from probing the whole table for every row.  A
// key placed within kSampleProbe probes is found again within as many.
constexpr int kSampleProbe = 8;

// The end of the sketch block: the Shlosser sums over the count histogram
// (thread t takes bucket t), then, on wave 0, the plain distinct estimate and
// the sizer's features (lane i evaluates feature i, so the wave pays for one
// logarithm and one division), then, with a model installed, its forward
// pass on parameters staged in LDS while the sums ran.  hist holds the
// counts divided by scale (8 after a hash-subsampled recount; buckets 1 and
// 2 included).  Writes the sizer block after the stats.
static_assert((PMH_MLP_IN & (PMH_MLP_IN - 1)) == 0 && PMH_MLP_IN <= 64, "one wave-0 lane per feature");
__device__ void sketch_sizer(const HLaunch *__restrict__ L, uint32_t s_sel, uint32_t d, uint32_t f1, uint32_t f2,
                             uint32_t dropped, uint32_t side_keys, const uint32_t *hist, uint32_t scale,
                             float *s_par) {
    __shared__ double s_red[3][16];
    const uint32_t tid = threadIdx.x, bs = blockDim.x, lane = tid & 63;
    if (L->sizer)  // uniform across the block
        for (uint32_t i = tid; i < PMH_MLP_PARAMS; i += bs) s_par[i] = L->sizer[i];
    Sketch k = sketch_head(s_sel, L->n_sample, L->q.n_rows, d, f1, f2, dropped);
    const double qf = sketch_qf(k);
    double F2 = 0, num = 0, den = 0;
    for (uint32_t c = tid; c < 1024; c += bs)
        if (c >= 1) sketch_bucket((int)c, (double)hist[c] * scale + (c == 1 ? (double)dropped : 0.0), qf, F2, num, den);
#pragma unroll
    for (int o = 32; o; o >>= 1) {
        F2 += __shfl_xor(F2, o);
        num += __shfl_xor(num, o);
        den += __shfl_xor(den, o);
    }
    if (lane == 0) {
        s_red[0][tid / 64] = F2;
        s_red[1][tid / 64] = num;
        s_red[2][tid / 64] = den;
    }
    __syncthreads();
    if (tid >= 64) return;
    const uint32_t nw = bs / 64;
    F2 = lane < nw ? s_red[0][lane] : 0.0;
    num = lane < nw ? s_red[1][lane] : 0.0;
    den = lane < nw ? s_red[2][lane] : 0.0;
#pragma unroll
    for (int o = 32; o; o >>= 1) {
        F2 += __shfl_xor(F2, o);
        num += __shfl_xor(num, o);
        den += __shfl_xor(den, o);
    }
    k.F2 = F2;
    k.shl_num = num;
    k.shl_den = den;
    const double gee = gee_estimate(k), shl = shlosser_estimate(k), plain = plain_from(k, gee, shl);
    double a, b;
    const int kind = sizer_feature_arg((int)(lane & (PMH_MLP_IN - 1)), k, side_keys, L->lane_bits, gee, shl, plain, a, b);
    const float f = sizer_feature_value(kind, a, b);
    uint32_t *blk = L->stats + ST_N;
    if (lane < PMH_MLP_IN) ((float *)blk)[SZ_FEAT + lane] = f;
    if (lane == 0) {
        *(double *)(blk + SZ_EST) = plain;
        blk[SZ_HAVE] = L->sizer ? 1u : 0u;
        blk[SZ_SCALE] = scale;
    }
    if (!L->sizer) return;
    float x[PMH_MLP_IN];
#pragma unroll
    for (int i = 0; i < PMH_MLP_IN; ++i) x[i] = __shfl(f, i);
    const mlpdev::Acts o = mlpdev::forward_lane(s_par, x, (int)lane);
    if (lane == 0) {
        ((float *)blk)[SZ_PRED + 0] = o.o0;
        ((float *)blk)[SZ_PRED + 1] = o.o1;
    }
}

__global__ __launch_bounds__(1024) void hs_sample_count(const HLaunch *__restrict__ L) {
    extern __shared__ __align__(16) uint8_t smem[];
    const QueryDesc *__restrict__ q = &L->q;
    const bool wide = L->wide;
    const uint32_t SS = L->sample_slots;  // power of two
    uint32_t *s_fp = (uint32_t *)smem;    // fingerprint | 1, 0 = free
    uint32_t *s_cnt = s_fp + SS;          // count; bit 31: picked; bit 30: key placed
    uint32_t *hist = s_cnt + SS;          // 1024 count buckets
    uint64_t *h_lo = (uint64_t *)(hist + 1024);
    uint64_t *h_hi = h_lo + L->hot_slots;
    __shared__ uint32_t s_thresh, s_take_eq, s_hot, s_f1, s_f2, s_distinct, s_drop, s_hotrows;
    __shared__ uint32_t s_wave[16];
    __shared__ uint32_t s_eq_used, s_npick, s_side;
    const uint32_t tid = threadIdx.x, bs = blockDim.x;
    const uint32_t CL = L->hot_slots;
    for (uint32_t i = tid; i < SS; i += bs) { s_fp[i] = 0; s_cnt[i] = 0; }
    for (uint32_t i = tid; i < 1024; i += bs) hist[i] = 0;
    for (uint32_t i = tid; i < CL; i += bs) { h_lo[i] = kEmpty; if (wide) h_hi[i] = kEmpty; }
    if (tid == 0) { s_hot = 0; s_f1 = 0; s_f2 = 0; s_distinct = 0; s_drop = 0; s_take_eq = 0; s_hotrows = 0; s_eq_used = 0; s_npick = 0; s_side = 0; }
    __syncthreads();
    const uint32_t nsel = min(L->stats[ST_SAMPLE_SEL], L->n_sample);
    uint32_t drops = 0, sidekeys = 0;
    // sampled keys in batches of kSampleU rows per thread; the next batch's
    // loads are issued before the current batch is probed, so the table
    // work hides the sample's memory latency
    constexpr int kSampleU = 8;
    uint64_t nlo[kSampleU], nhi[kSampleU];
    uint8_t nsp[kSampleU];
    auto load_batch = [&](uint32_t j0) __attribute__((always_inline)) {
#pragma unroll
        for (int u = 0; u < kSampleU; ++u) {
            const uint32_t i = j0 + u * bs;
            const uint32_t ii = i < nsel ? i : 0;
            nlo[u] = gld<uint64_t>(L->samp_lo, ii);
            nhi[u] = wide ? gld<uint64_t>(L->samp_hi, ii) : 0;
            nsp[u] = gld<uint8_t>(L->samp_sp, ii);
        }
    };
    if (tid < nsel) load_batch(tid);
    for (uint32_t i0 = tid; i0 < nsel; i0 += bs * kSampleU) {
        uint64_t lo[kSampleU], hi[kSampleU];
        uint8_t sp[kSampleU];
#pragma unroll
        for (int u = 0; u < kSampleU; ++u) { lo[u] = nlo[u]; hi[u] = nhi[u]; sp[u] = nsp[u]; }
        if (i0 + bs * kSampleU < nsel) load_batch(i0 + bs * kSampleU);
#pragma unroll
        for (int u = 0; u < kSampleU; ++u) {
            if (i0 + u * bs >= nsel) break;
            if (sp[u]) { ++sidekeys; continue; }  // side keys are never hot
            const uint64_t h = hash_key(lo[u], hi[u]);
            const uint32_t fp = (uint32_t)(h >> 32) | 1u;
            uint32_t s = (uint32_t)h & (SS - 1);
            bool done = false;
            for (int p = 0; p < kSampleProbe; ++p) {
                uint32_t cur = s_fp[s];
                if (cur == 0) {
                    const uint32_t old = atomicCAS(&s_fp[s], 0u, fp);
                    cur = old == 0u ? fp : old;
                    if (old == 0u) {
                        // the claimer records the slot's key for the placement
                        gst<uint64_t>(L->skey_lo, s, lo[u]);
                        if (wide) gst<uint64_t>(L->skey_hi, s, hi[u]);
                    }
                }
                if (cur == fp) { atomicAdd(&s_cnt[s], 1u); done = true; break; }
                s = (s + 1) & (SS - 1);
            }
            if (!done) ++drops;
        }
    }
    if (drops) atomicAdd(&s_drop, drops);
    if (sidekeys) atomicAdd(&s_side, sidekeys);
    __syncthreads();
    // counts 1 and 2 (most of the keys) are tallied in registers and give
    // their histogram buckets once: thousands of LDS atomics on hist[1] and
    // hist[2] would serialise on one address each
    uint32_t f1 = 0, f2 = 0, dist = 0;
    for (uint32_t i = tid; i < SS; i += bs) {
        const uint32_t c = s_cnt[i];
        if (!c) continue;
        ++dist;
        if (c == 1) ++f1;
        else if (c == 2) ++f2;
        else atomicAdd(&hist[min(c, 1023u)], 1u);
    }
    f1 = wave_sum(f1);
    f2 = wave_sum(f2);
    dist = wave_sum(dist);
    if ((tid & 63) == 0) {
        atomicAdd(&s_f1, f1);
        atomicAdd(&s_f2, f2);
        atomicAdd(&s_distinct, dist);
    }
    __syncthreads();
    if (tid == 0) { hist[1] = s_f1; hist[2] = s_f2; }
    __syncthreads();
    for (uint32_t i = tid; i < 1024; i += bs) L->sketch_hist[i] = hist[i];
    {
        // the largest count threshold T such that the keys with count > T
        // fit in hot_max: with S(c) the keys of count >= c, T is the c >= 1
        // where S(c) > hot_max >= S(c + 1) (none: T = 0, every key fits).
        // Thread t holds bucket c = 1023 - t, so a block scan gives S(c + 1).
        const uint32_t c = 1023 - tid;
        const uint32_t x = c >= 1 ? hist[c] : 0;
        const uint32_t above = block_excl_scan(x, s_wave);  // S(c + 1)
        const uint32_t kmax = L->hot_max;
        if (tid == 0) { s_thresh = 0; s_take_eq = 0; }
        __syncthreads();
        if (c >= 1 && above + x > kmax && above <= kmax) {
            s_thresh = c;
            s_take_eq = kmax - above;  // keys with count == T still allowed
        }
    }
    if (tid == 0) {
        L->stats[ST_SAMPLE_DISTINCT] = s_distinct;
        L->stats[ST_SAMPLE_F1] = s_f1;
        L->stats[ST_SAMPLE_F2] = s_f2;
        L->stats[ST_SAMPLE_DROPPED] = s_drop;
        L->stats[ST_SAMPLE_NULLKEY] = s_side;
    }
    __syncthreads();
    // Picks, one slot per thread: keys with count > T, and keys with count
    // == T in wave order up to the remaining places (one atomic per wave
    // ranks its lanes; SS is a multiple of the block size, so every lane runs
    // every iteration).  A picked slot joins a compact list (slot, sampled
    // count) in the fingerprint array, whose counting is done (CL <= 2048
    // entries of 8 B inside its 4 * SS >= 32 KB).
    const uint32_t T = s_thresh;
    const uint32_t lane = tid & 63;
    uint32_t *pk_slot = s_fp;
    uint32_t *pk_cnt = s_fp + CL;
    for (uint32_t i0 = 0; i0 < SS && CL; i0 += bs) {
        const uint32_t i = i0 + tid;
        const uint32_t cnt = s_cnt[i];
        const uint32_t c = min(cnt, 1023u);
        bool pick = c && c > T;
        const bool eq = c && c == T;
        const unsigned long long m = __ballot(eq);
        if (m) {
            const int leader = __ffsll((long long)m) - 1;
            uint32_t base = 0;
            if ((int)lane == leader) base = atomicAdd(&s_eq_used, (uint32_t)__popcll(m));
            base = __shfl(base, leader);
            if (eq) pick = base + (uint32_t)__popcll(m & ((1ull << lane) - 1)) < s_take_eq;
        }
        const unsigned long long pm = __ballot(pick);
        if (pm) {
            const int leader = __ffsll((long long)pm) - 1;
            uint32_t base = 0;
            if ((int)lane == leader) base = atomicAdd(&s_npick, (uint32_t)__popcll(pm));
            base = __shfl(base, leader);
            const uint32_t e = base + (uint32_t)__popcll(pm & ((1ull << lane) - 1));
            if (pick && e < CL) {
                pk_slot[e] = i;
                pk_cnt[e] = cnt;
            }
        }
    }
    __syncthreads();
    // the picked keys, recorded by the rows that claimed their slots, move
    // into the count array, whose counts are now in the list (16 B per
    // entry inside its 4 * SS bytes)
    const uint32_t npick = min(s_npick, CL);
    uint64_t *pk_lo = (uint64_t *)s_cnt;
    uint64_t *pk_hi = pk_lo + CL;
    for (uint32_t e = tid; e < npick; e += bs) {
        const uint32_t i = pk_slot[e];
        pk_lo[e] = gld<uint64_t>(L->skey_lo, i);
        pk_hi[e] = wide ? gld<uint64_t>(L->skey_hi, i) : 0;
    }
    __syncthreads();
    // Placement: a key goes to the first free slot of its first bucket, else
    // of its second; a key that finds both full stays out of the hot set, and
    // its rows all go to one tail partition, so the hottest keys are placed
    // first: sampled at least 8 times, then 2..7 times, then once, each tier
    // after a barrier.
    const uint32_t nbk = CL ? hot_buckets(CL) : 0;
    uint32_t hot_rows = 0;
    constexpr uint32_t kTierLo[3] = {8, 2, 1}, kTierHi[3] = {0xFFFFFFFFu, 7, 1};
    for (int tier = 0; tier < 3; ++tier) {
        for (uint32_t e = tid; e < npick; e += bs) {
            const uint32_t cnt = pk_cnt[e];
            if (cnt < kTierLo[tier] || cnt > kTierHi[tier]) continue;
            const uint64_t lo = pk_lo[e], hi = pk_hi[e];
            const uint64_t h = hash_key(lo, hi);
            bool placed = false;
            for (int pass = 0; pass < 2 && !placed; ++pass) {
                const uint32_t b = (pass ? hot_b2(h, nbk) : hot_b1(h, nbk)) * 4;
                for (int jj = 0; jj < 4; ++jj) {
                    const unsigned long long old = atomicCAS((unsigned long long *)&h_lo[b + jj],
                                                             (unsigned long long)kEmpty, (unsigned long long)lo);
                    if (old == kEmpty) {
                        if (wide) h_hi[b + jj] = hi;
                        atomicAdd(&s_hot, 1u);
                        hot_rows += cnt;
                        placed = true;
                        break;
                    }
                }
            }
        }
        __syncthreads();
    }
    if (hot_rows) atomicAdd(&s_hotrows, hot_rows);
    __syncthreads();
    for (uint32_t i = tid; i < CL; i += bs) {
        L->hot_lo[i] = h_lo[i];
        if (wide) L->hot_hi[i] = h_hi[i];
        const Layout &lay = q->layout;
        for (int a = 0; a < lay.n_arrs; ++a) {
            const int op = lay.init[a];
            if (mo_size(op) == 4) ((uint32_t *)L->hot_arr[a])[i] = (uint32_t)mo_identity(op);
            else ((uint64_t *)L->hot_arr[a])[i] = mo_identity(op);
        }
    }
    if (tid == 0) {
        L->stats[ST_HOT] = s_hot;
        L->stats[ST_HOT_SAMPLE_ROWS] = s_hotrows;
        L->out_cursor[0] = 0;
    }
    // A table that dropped more than an eighth of the sample (far more
    // distinct sampled keys than slots) leaves a frequency profile biased to
    // the keys that arrived first.  The sketch then comes from a hash
    // subsample instead (distinct sampling): only keys whose hash has its top
    // three bits clear, an eighth of the key space, with all of their sampled
    // rows, counted again in the cleared table.  Their counts times eight
    // estimate the whole sample's, and the estimators are linear in them.
    const bool recount = s_drop * 8 > nsel;  // uniform: s_drop is final since the count's barrier
    if (recount) {
        __syncthreads();
        for (uint32_t i = tid; i < SS; i += bs) { s_fp[i] = 0; s_cnt[i] = 0; }
        for (uint32_t i = tid; i < 1024; i += bs) hist[i] = 0;
        if (tid == 0) { s_f1 = 0; s_f2 = 0; s_distinct = 0; s_drop = 0; }
        __syncthreads();
        uint32_t drops2 = 0;
        for (uint32_t i0 = tid; i0 < nsel; i0 += bs * kSampleU) {
            uint64_t lo[kSampleU], hi[kSampleU];
            uint8_t sp[kSampleU];
#pragma unroll
            for (int u = 0; u < kSampleU; ++u) {
                const uint32_t i = i0 + u * bs;
                const uint32_t ii = i < nsel ? i : 0;
                lo[u] = gld<uint64_t>(L->samp_lo, ii);
                hi[u] = wide ? gld<uint64_t>(L->samp_hi, ii) : 0;
                sp[u] = gld<uint8_t>(L->samp_sp, ii);
            }
#pragma unroll
            for (int u = 0; u < kSampleU; ++u) {
                if (i0 + u * bs >= nsel) break;
                if (sp[u]) continue;
                const uint64_t h = hash_key(lo[u], hi[u]);
                if (h >> 61) continue;
                const uint32_t fp = (uint32_t)(h >> 32) | 1u;
                uint32_t s = (uint32_t)h & (SS - 1);
                bool done = false;
                for (int p = 0; p < kSampleProbe; ++p) {
                    uint32_t cur = s_fp[s];
                    if (cur == 0) {
                        const uint32_t old = atomicCAS(&s_fp[s], 0u, fp);
                        cur = old == 0u ? fp : old;
                    }
                    if (cur == fp) { atomicAdd(&s_cnt[s], 1u); done = true; break; }
                    s = (s + 1) & (SS - 1);
                }
                if (!done) ++drops2;
            }
        }
        if (drops2) atomicAdd(&s_drop, drops2);
        __syncthreads();
        uint32_t g1 = 0, g2 = 0, gd = 0;
        for (uint32_t i = tid; i < SS; i += bs) {
            const uint32_t c = s_cnt[i];
            if (!c) continue;
            ++gd;
            if (c == 1) ++g1;
            else if (c == 2) ++g2;
            else atomicAdd(&hist[min(c, 1023u)], 1u);
        }
        g1 = wave_sum(g1);
        g2 = wave_sum(g2);
        gd = wave_sum(gd);
        if ((tid & 63) == 0) {
            atomicAdd(&s_f1, g1);
            atomicAdd(&s_f2, g2);
            atomicAdd(&s_distinct, gd);
        }
        __syncthreads();
        if (tid == 0) { hist[1] = s_f1; hist[2] = s_f2; }
        __syncthreads();
        for (uint32_t i = tid; i < 1024; i += bs) L->sketch_hist[i] = hist[i] * 8;
        if (tid == 0) {
            L->stats[ST_SAMPLE_DISTINCT] = s_distinct * 8;
            L->stats[ST_SAMPLE_F1] = s_f1 * 8;
            L->stats[ST_SAMPLE_F2] = s_f2 * 8;
            L->stats[ST_SAMPLE_DROPPED] = s_drop * 8;
        }
    }
    __syncthreads();  // the LDS tables are free from here
    const uint32_t sc = recount ? 8 : 1;
    // the model's parameters (21 KB) are staged over the fingerprint array
    // (at least 32 KB), away from the histogram the sums read
    static_assert(PMH_MLP_PARAMS * 4 <= 8192 * 4, "sizer parameters fit the smallest fingerprint array");
    sketch_sizer(L, nsel, s_distinct * sc, s_f1 * sc, s_f2 * sc, s_drop * sc, s_side, hist, sc, (float *)smem);
    // the words the host plans from go straight into its pinned mailbox
    // (vector stores over the link; no copy on the stream)
    __syncthreads();
    for (uint32_t i = tid; i < (uint32_t)L->back_words; i += bs) L->mbox[i] = L->stats[i];
    __threadfence_system();
}

// ---------------------------------------------------------------------------
// scan
// ---------------------------------------------------------------------------
struct ScanCtx {
    uint64_t *lkeys, *lkeys_hi;
    uint32_t CL;
    LSink ls;
    uint32_t *s_cursor;
    uint64_t region;  // spill index of this block's region
};

// One row of the scan: returns true when the row is a tail row (key in k).
template <class Row>
__device__ __forceinline__ bool scan_row(const HLaunch *__restrict__ L, const QueryDesc *__restrict__ q,
                                         const Row &row, bool live, ScanCtx &c, uint32_t &sel, Key &k) {
    if (!live || !eval_pred(q, row)) return false;
    ++sel;
    k = load_key(q, row);
    if (k.special || c.CL == 0) return true;
    const uint64_t h = hash_key(k.lo, k.hi);
    const uint32_t nb = hot_buckets(c.CL);
    const uint32_t b1 = hot_b1(h, nb) * 4, b2 = hot_b2(h, nb) * 4;
    bool open = false;
    for (int p = 0; p < 8; ++p) {
        if (p == 4 && open) return true;
        const uint32_t s = p < 4 ? b1 + p : b2 + (p - 4);
        const uint64_t cur = c.lkeys[s];
        open |= p < 4 && cur == kEmpty;
        if (cur == k.lo && (!L->wide || c.lkeys_hi[s] == k.hi)) {
            add_row(c.ls, q, s, row);
            return false;
        }
    }
    return true;
}

// Wave-contiguous append of the tail rows of one row step.
template <class Row>
__device__ __forceinline__ void append_tail(const HLaunch *__restrict__ L, const QueryDesc *__restrict__ q,
                                            const Row &row, bool is_tail, const Key &k, ScanCtx &c) {
    const unsigned long long m = __ballot(is_tail);
    if (!m) return;
    const int lane = threadIdx.x & 63;
    const int leader = __ffsll((long long)m) - 1;
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(c.s_cursor, (uint32_t)__popcll(m));
    base = __shfl(base, leader);
    if (!is_tail) return;
    const uint64_t pos = c.region + base + (uint32_t)__popcll(m & ((1ull << lane) - 1));
    const Spill &S = L->spill;
    uint16_t fl = k.special == 1 ? kFlagAllOnes : (k.special == 2 ? kFlagNull : 0);
    for (int j = 0; j < q->n_vals; ++j) {
        const int col = q->vals[j].col;
        if (row.valid(col)) fl |= (uint16_t)(1u << j);
        gstore_w(S.val[j], L->vw[j], pos, row.bits(col));
    }
    gst<uint64_t>(S.klo, pos, k.lo);
    if (L->wide) gst<uint64_t>(S.khi, pos, k.hi);
    gst<uint16_t>(S.flags, pos, fl);
}

template <bool CACHED>
__global__ __launch_bounds__(1024) void hs_scan(const HLaunch *__restrict__ L) {
    extern __shared__ __align__(16) uint8_t smem[];
    __shared__ uint32_t s_cursor, s_sel;
    const QueryDesc *__restrict__ q = &L->q;
    const uint32_t CL = L->hot_slots;
    const uint32_t tid = threadIdx.x, bs = blockDim.x;
    const bool wide = L->wide;
    uint64_t *lkeys = (uint64_t *)smem;
    uint64_t *lkeys_hi = lkeys + CL;
    uint8_t *state = (uint8_t *)(wide ? lkeys_hi + CL : lkeys_hi);
    ScanCtx c{lkeys, lkeys_hi, CL, LSink{state, L->lds1_off, CL}, &s_cursor,
              (uint64_t)blockIdx.x * L->chunk};
    if (tid == 0) { s_cursor = 0; s_sel = 0; }
    for (uint32_t i = tid; i < CL; i += bs) {
        lkeys[i] = L->hot_lo[i];
        if (wide) lkeys_hi[i] = L->hot_hi[i];
    }
    for (int a = 0; a < q->layout.n_arrs; ++a) {
        const int op = q->layout.init[a];
        if (mo_size(op) == 4) {
            const uint32_t id = (uint32_t)mo_identity(op);
            for (uint32_t i = tid; i < CL; i += bs) *c.ls.p<uint32_t>(a, i) = id;
        } else {
            const uint64_t id = mo_identity(op);
            for (uint32_t i = tid; i < CL; i += bs) *c.ls.p<uint64_t>(a, i) = id;
        }
    }
    __syncthreads();
    const uint64_t n = q->n_rows;
    const uint64_t r_begin = (uint64_t)blockIdx.x * L->chunk;
    const uint64_t r_end = min(n, r_begin + L->chunk);
    const int ncols = q->n_cols;
    uint32_t sel = 0;
    if constexpr (CACHED) {
        auto fill = [&](CacheRow &row, uint64_t r) __attribute__((always_inline)) {
            row.q = q;
            row.vm = 0;
            row.v0 = row.v1 = row.v2 = row.v3 = 0;
            if (r < r_end) {
#pragma unroll
                for (int cc = 0; cc < 4; ++cc) {
                    if (cc < ncols) {
                        const DevCol &col = q->cols[cc];
                        row.set(cc, load_raw(col, r));
                        row.vm |= (uint32_t)dvalid(col.valid, r) << cc;
                    }
                }
            }
        };
        // every lane runs the same number of steps so the ballots in
        // append_tail always see whole waves
        const uint64_t span = r_end > r_begin ? r_end - r_begin : 0;
        const uint64_t steps = (span + 4ull * bs - 1) / (4ull * bs);
        for (uint64_t it = 0; it < steps; ++it) {
            const uint64_t r0 = r_begin + it * 4ull * bs + tid;
            CacheRow a0, a1, a2, a3;
            fill(a0, r0);
            fill(a1, r0 + bs);
            fill(a2, r0 + 2ull * bs);
            fill(a3, r0 + 3ull * bs);
            Key k;
            bool t;
            t = scan_row(L, q, a0, r0 < r_end, c, sel, k);
            append_tail(L, q, a0, t, k, c);
            t = scan_row(L, q, a1, r0 + bs < r_end, c, sel, k);
            append_tail(L, q, a1, t, k, c);
            t = scan_row(L, q, a2, r0 + 2ull * bs < r_end, c, sel, k);
            append_tail(L, q, a2, t, k, c);
            t = scan_row(L, q, a3, r0 + 3ull * bs < r_end, c, sel, k);
            append_tail(L, q, a3, t, k, c);
        }
    } else {
        // Every lane of a wave iterates the same number of times, so the
        // ballots in append_tail see whole waves.
        const uint64_t span = r_end > r_begin ? r_end - r_begin : 0;
        const uint64_t iters = (span + bs - 1) / bs;
        for (uint64_t it = 0; it < iters; ++it) {
            const uint64_t r = r_begin + it * bs + tid;
            const bool live = r < r_end;
            MemRow row{q, live ? r : r_begin};
            Key k;
            const bool t = scan_row(L, q, row, live, c, sel, k);
            append_tail(L, q, row, t, k, c);
        }
    }
    atomicAdd(&s_sel, sel);
    __syncthreads();
    if (tid == 0) {
        L->tail_n[blockIdx.x] = s_cursor;
        atomicAdd(&L->stats[ST_SELECTED], s_sel);
        atomicAdd(&L->stats[ST_TAIL], s_cursor);
    }
    // hot partials: block slot s -> global slot s (coalesced atomics)
    GSink g{L->hot_arr, CL};
    const int a_len = q->layout.a_len;
    for (uint32_t i = tid; i < CL; i += bs) {
        if (*c.ls.p<uint32_t>(a_len, i) == 0) continue;
        merge_lds_entry(g, i, c.ls, i, q->layout, q->layout.init);
    }
}

}  // namespace pmh
#include "pmh_scanfast.h"
namespace pmh {

// Prefix of tail counts per region and of first-pass tiles per region.
__global__ __launch_bounds__(1024) void hs_regions(const HLaunch *__restrict__ L) {
    __shared__ uint32_t s_a[1024], s_b[1024];
    __shared__ uint32_t carry_a, carry_b;
    for (uint32_t i = threadIdx.x; i < 8 * kWideRadix; i += blockDim.x) L->xtot[i] = 0;
    if (threadIdx.x == 0) { carry_a = 0; carry_b = 0; }
    __syncthreads();
    for (uint32_t b0 = 0; b0 < L->nb; b0 += 1024) {
        const uint32_t b = b0 + threadIdx.x;
        const uint32_t x = b < L->nb ? L->tail_n[b] : 0;
        const uint32_t y = (x + kPartTile - 1) / kPartTile;
        s_a[threadIdx.x] = x;
        s_b[threadIdx.x] = y;
        __syncthreads();
        for (int off = 1; off < 1024; off <<= 1) {
            uint32_t xa = threadIdx.x >= (uint32_t)off ? s_a[threadIdx.x - off] : 0;
            uint32_t xb = threadIdx.x >= (uint32_t)off ? s_b[threadIdx.x - off] : 0;
            __syncthreads();
            s_a[threadIdx.x] += xa;
            s_b[threadIdx.x] += xb;
            __syncthreads();
        }
        if (b < L->nb) {
            L->region_off[b] = carry_a + s_a[threadIdx.x] - x;
            const uint32_t t0 = carry_b + s_b[threadIdx.x] - y;
            L->tile_off[b] = t0;
            for (uint32_t j = 0; j < y; ++j) L->tile_reg[t0 + j] = b;
        }
        __syncthreads();
        if (threadIdx.x == 1023) { carry_a += s_a[1023]; carry_b += s_b[1023]; }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        L->region_off[L->nb] = carry_a;
        L->tile_off[L->nb] = carry_b;
        L->stats[ST_TILES0] = carry_b;
    }
}

// ---------------------------------------------------------------------------
// radix partition passes
// ---------------------------------------------------------------------------
// Tile t of a pass.  first: tiles live inside the scan regions of `spill`
// (tile_off maps tiles to regions); otherwise tiles are contiguous in src.
struct TileSpan {
    uint64_t src_begin;   // index of the tile's first row in the source arrays
    uint32_t rows;        // rows in the tile
    uint64_t order_begin; // rank of the first row in the pass's input order
};

__device__ __forceinline__ bool tile_span(const HLaunch *__restrict__ L, bool first, uint32_t t, uint32_t ntiles,
                                          uint64_t total, TileSpan &ts) {
    if (t >= ntiles) return false;
    if (!first) {
        ts.src_begin = (uint64_t)t * kPartTile;
        ts.order_begin = ts.src_begin;
        ts.rows = (uint32_t)min<uint64_t>(kPartTile, total - ts.src_begin);
        return true;
    }
    const uint32_t lo = L->tile_reg[t];  // the region holding tile t
    const uint32_t j = t - L->tile_off[lo];
    const uint32_t n_b = L->tail_n[lo];
    ts.src_begin = (uint64_t)lo * L->chunk + (uint64_t)j * kPartTile;
    ts.order_begin = (uint64_t)L->region_off[lo] + (uint64_t)j * kPartTile;
    ts.rows = min<uint32_t>(kPartTile, n_b - j * kPartTile);
    return true;
}

__device__ __forceinline__ uint32_t pass_digit(uint64_t lo, uint64_t hi, uint16_t fl, int shift) {
    return (uint32_t)(spill_hash(lo, hi, fl) >> shift) & (kRadix - 1);
}

__global__ __launch_bounds__(kPartThreads) void rp_hist(const HLaunch *__restrict__ L, int first, const Spill *src_p,
                                                       int shift) {
    __shared__ uint32_t hist[kRadix];
    const Spill &src = *src_p;
    const uint32_t ntiles = first ? L->tile_off[L->nb] : (uint32_t)((L->region_off[L->nb] + kPartTile - 1) / kPartTile);
    const uint64_t total = L->region_off[L->nb];
    for (uint32_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        if (threadIdx.x < kRadix) hist[threadIdx.x] = 0;
        __syncthreads();
        TileSpan ts;
        tile_span(L, first, t, ntiles, total, ts);
        for (uint32_t k = 0; k < kPartRows; ++k) {
            const uint32_t i = k * kPartThreads + threadIdx.x;
            if (i < ts.rows) {
                const uint64_t s = ts.src_begin + i;
                const uint64_t hi = L->wide ? gld<uint64_t>(src.khi, s) : 0;
                atomicAdd(&hist[pass_digit(gld<uint64_t>(src.klo, s), hi, gld<uint16_t>(src.flags, s), shift)], 1u);
            }
        }
        __syncthreads();
        if (threadIdx.x < kRadix) L->tile_hist[(uint64_t)threadIdx.x * L->max_tiles + t] = hist[threadIdx.x];
        __syncthreads();
    }
}

// One block per digit: exclusive scan of the digit's tile counts in place;
// digit totals go to digit_base (then exclusive-scanned by block 0).
__global__ __launch_bounds__(1024) void rp_scan(const HLaunch *__restrict__ L, int first) {
    __shared__ uint32_t s[1024];
    __shared__ uint32_t carry;
    __shared__ uint32_t totals[kRadix];
    const uint32_t ntiles = first ? L->tile_off[L->nb] : (uint32_t)((L->region_off[L->nb] + kPartTile - 1) / kPartTile);
    const uint32_t d = blockIdx.x;
    uint32_t *row = L->tile_hist + (uint64_t)d * L->max_tiles;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (uint32_t t0 = 0; t0 < ntiles; t0 += 1024) {
        const uint32_t t = t0 + threadIdx.x;
        const uint32_t x = t < ntiles ? row[t] : 0;
        s[threadIdx.x] = x;
        __syncthreads();
        for (int off = 1; off < 1024; off <<= 1) {
            uint32_t y = threadIdx.x >= (uint32_t)off ? s[threadIdx.x - off] : 0;
            __syncthreads();
            s[threadIdx.x] += y;
            __syncthreads();
        }
        if (t < ntiles) row[t] = carry + s[threadIdx.x] - x;
        __syncthreads();
        if (threadIdx.x == 1023) carry += s[1023];
        __syncthreads();
    }
    if (threadIdx.x == 0) L->digit_base[d] = carry;  // digit total for now
    (void)totals;
}

// Stable scatter of one pass.  Rows are ranked in input order: within a
// wave by multisplit ballots, across waves and the tile's four row steps by
// LDS counters; the tile is staged in LDS in digit order and written out so
// consecutive threads write consecutive positions of each digit run.
__global__ __launch_bounds__(kPartThreads) void rp_scatter(const HLaunch *__restrict__ L, int first,
                                                          const Spill *src_p, const Spill *dst_p, int shift) {
    constexpr int kW = kPartThreads / 64;
    extern __shared__ __align__(16) uint8_t smem[];
    __shared__ uint32_t run[kRadix];       // rows of each digit in this tile
    __shared__ uint32_t dstart[kRadix];    // LDS start of each digit run
    __shared__ uint32_t gbase[kRadix];     // global start of each digit run for this tile
    __shared__ uint32_t dtot[kRadix];
    __shared__ uint32_t s_wave[16];
    const Spill &src = *src_p;
    const Spill &dst = *dst_p;
    const QueryDesc *__restrict__ q = &L->q;
    const int nv = q->n_vals;
    const bool wide = L->wide;
    const uint32_t ntiles = first ? L->tile_off[L->nb] : (uint32_t)((L->region_off[L->nb] + kPartTile - 1) / kPartTile);
    const uint64_t total = L->region_off[L->nb];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const unsigned long long lt = (1ull << lane) - 1;
    // LDS staging: keys, [hi], flags, digits, then value columns
    uint64_t *st_lo = (uint64_t *)smem;
    uint64_t *st_hi = st_lo + kPartTile;
    uint8_t *p = (uint8_t *)(wide ? st_hi + kPartTile : st_hi);
    uint16_t *st_fl = (uint16_t *)p;
    uint8_t *st_d = (uint8_t *)(st_fl + kPartTile);
    uint8_t *st_v = st_d + kPartTile;
    // per (row step, wave, digit): the wave's rows of that digit in that
    // step, then (in place) the rows of that digit before them in the tile;
    // it lives in the staging keys, which are written only after the ranks
    uint16_t *wc = (uint16_t *)smem;
    static_assert(kPartRows * kW * kRadix * 2 <= kPartTile * 8, "the counts fit the staged keys");
    // digit totals over all tiles -> exclusive digit bases
    {
        const uint32_t x = threadIdx.x < kRadix ? L->digit_base[threadIdx.x] : 0;
        const uint32_t ex = block_excl_scan(x, s_wave);
        if (threadIdx.x < kRadix) dtot[threadIdx.x] = ex;
    }
    __syncthreads();
    for (uint32_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        TileSpan ts;
        tile_span(L, first, t, ntiles, total, ts);
        uint32_t my_d[kPartRows], my_pos[kPartRows], my_rank[kPartRows];
        // every row of the tile (keys, flags, the first two value columns)
        // loads before the first rank: one wait on memory per tile
        uint64_t klo[kPartRows], khi[kPartRows], v0[kPartRows], v1[kPartRows];
        uint16_t fl[kPartRows];
#pragma unroll
        for (int k = 0; k < kPartRows; ++k) {
            const uint32_t i = k * kPartThreads + threadIdx.x;
            const uint64_t s = ts.src_begin + (i < ts.rows ? i : 0);
            klo[k] = gld_nt<uint64_t>(src.klo, s);
            khi[k] = wide ? gld_nt<uint64_t>(src.khi, s) : 0;
            fl[k] = gld_nt<uint16_t>(src.flags, s);
            v0[k] = nv > 0 ? gload_w_nt(src.val[0], L->vw[0], s) : 0;
            v1[k] = nv > 1 ? gload_w_nt(src.val[1], L->vw[1], s) : 0;
        }
        for (uint32_t x = threadIdx.x; x < kPartRows * kW * kRadix / 2; x += kPartThreads) ((uint32_t *)wc)[x] = 0;
        // every step's multisplit first (lanes holding the same digit), then
        // one count per (step, wave, digit) and one scan per digit in (step,
        // wave) order, which is row order: the rank stays stable
#pragma unroll
        for (int k = 0; k < kPartRows; ++k) {
            const uint32_t i = k * kPartThreads + threadIdx.x;
            const bool live = i < ts.rows;
            const uint32_t d = live ? pass_digit(klo[k], khi[k], fl[k], shift) : 0;
            unsigned long long peers = __ballot(live);
#pragma unroll
            for (int b = 0; b < kRadixBits; ++b) {
                const unsigned long long m = __ballot((d >> b) & 1);
                peers &= ((d >> b) & 1) ? m : ~m;
            }
            my_d[k] = d;
            my_rank[k] = live ? (uint32_t)__popcll(peers & lt) : 0xFFFFFFFFu;
            my_pos[k] = live ? (uint32_t)__popcll(peers) : 0;  // the peers' count, for the leader
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < kPartRows; ++k)
            if (my_rank[k] == 0) wc[(k * kW + wave) * kRadix + my_d[k]] = (uint16_t)my_pos[k];
        __syncthreads();
        if (threadIdx.x < kRadix) {
            const uint32_t dd = threadIdx.x;
            uint32_t acc = 0;
#pragma unroll
            for (int k = 0; k < kPartRows; ++k)
                for (int w = 0; w < kW; ++w) {
                    const uint32_t x = wc[(k * kW + w) * kRadix + dd];
                    wc[(k * kW + w) * kRadix + dd] = (uint16_t)acc;
                    acc += x;
                }
            run[dd] = acc;
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < kPartRows; ++k)
            my_pos[k] = my_rank[k] != 0xFFFFFFFFu ? wc[(k * kW + wave) * kRadix + my_d[k]] + my_rank[k] : 0xFFFFFFFFu;
        // LDS start of each digit run within the tile
        {
            const uint32_t c = threadIdx.x < kRadix ? run[threadIdx.x] : 0;
            const uint32_t st = block_excl_scan(c, s_wave);
            if (threadIdx.x < kRadix) {
                dstart[threadIdx.x] = st;
                gbase[threadIdx.x] = dtot[threadIdx.x] + L->tile_hist[(uint64_t)threadIdx.x * L->max_tiles + t];
            }
        }
        __syncthreads();
        // stage in digit order
#pragma unroll
        for (int k = 0; k < kPartRows; ++k) {
            if (my_pos[k] == 0xFFFFFFFFu) continue;
            const uint32_t i = k * kPartThreads + threadIdx.x;
            const uint64_t s = ts.src_begin + i;
            const uint32_t sp = dstart[my_d[k]] + my_pos[k];
            st_lo[sp] = klo[k];
            if (wide) st_hi[sp] = khi[k];
            st_fl[sp] = fl[k];
            st_d[sp] = (uint8_t)my_d[k];
            uint8_t *vp = st_v;
            for (int j = 0; j < nv; ++j) {
                const int w = L->vw[j];
                store_w(vp, w, sp, j == 0 ? v0[k] : j == 1 ? v1[k] : gload_w(src.val[j], w, s));
                vp += (size_t)w * kPartTile;
            }
        }
        __syncthreads();
        // write runs: consecutive staged rows -> consecutive positions
        for (uint32_t sp = threadIdx.x; sp < ts.rows; sp += kPartThreads) {
            const uint32_t d = st_d[sp];
            const uint64_t g = (uint64_t)gbase[d] + (sp - dstart[d]);
            gst<uint64_t>(dst.klo, g, st_lo[sp]);
            if (wide) gst<uint64_t>(dst.khi, g, st_hi[sp]);
            gst<uint16_t>(dst.flags, g, st_fl[sp]);
            const uint8_t *vp = st_v;
            for (int j = 0; j < nv; ++j) {
                const int w = L->vw[j];
                gstore_w(dst.val[j], w, g, load_w(vp, w, sp));
                vp += (size_t)w * kPartTile;
            }
        }
        __syncthreads();
    }
}

// One pass, up to 1024 partitions.  Rows only need to be grouped by
// partition, not kept in a stable order, so a tile ranks its rows with LDS
// counters; the digit is the partition, the low pbits of the spill hash.
//
// The grid is a multiple of 8 and block b takes tiles b, b + grid, ..., so
// all tiles of one "XCD slot" x = b mod 8 run on one XCD (workgroups are
// dispatched round-robin over the eight XCDs).  Each partition is laid out
// as eight sub-regions, one per slot, and a tile claims its runs in its
// slot's sub-region with one atomic per digit: runs written at about the
// same time by one XCD's tiles land next to each other and are written
// through that XCD's L2, which completes their lines before they reach HBM.
// (Exact per-block offsets, with no atomic, measured slower: a block's runs
// then fill their lines only tile after tile, and the partial lines of every
// block outgrow the L2.)  Placement decides only speed: both kernels use the
// same tile -> block -> slot mapping.
__global__ __launch_bounds__(kPartThreads) void rp_hist1(const HLaunch *__restrict__ L, const Spill *src_p,
                                                        uint32_t nd) {
    __shared__ uint32_t hist[kWideRadix];
    const Spill &src = *src_p;
    const uint32_t ntiles = L->tile_off[L->nb];
    const uint64_t total = L->region_off[L->nb];
    const bool wide = L->wide;
    const uint32_t tid = threadIdx.x;
    hist[tid] = 0;
    __syncthreads();
    for (uint32_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        TileSpan ts;
        tile_span(L, true, t, ntiles, total, ts);
        uint32_t d[kPartRows];
#pragma unroll
        for (int k = 0; k < kPartRows; ++k) {
            const uint32_t i = k * kPartThreads + tid;
            const uint64_t s = ts.src_begin + (i < ts.rows ? i : 0);
            d[k] = (uint32_t)spill_hash(gld_nt<uint64_t>(src.klo, s), wide ? gld_nt<uint64_t>(src.khi, s) : 0,
                                        gld_nt<uint16_t>(src.flags, s)) &
                   (nd - 1);
        }
#pragma unroll
        for (int k = 0; k < kPartRows; ++k)
            if (k * kPartThreads + tid < ts.rows) atomicAdd(&hist[d[k]], 1u);
    }
    __syncthreads();
    const uint32_t x = blockIdx.x & 7;
    if (tid < nd && hist[tid]) atomicAdd(&L->xtot[x * kWideRadix + tid], hist[tid]);
}

// One block: the start of every (digit, slot) sub-region, digit-major, as
// the slots' cursors; the partition starts.
__global__ __launch_bounds__(kPartThreads) void rp_scan1(const HLaunch *__restrict__ L, uint32_t nd) {
    __shared__ uint32_t s_wave[16];
    const uint32_t d = threadIdx.x;
    uint32_t c[8], sum = 0;
#pragma unroll
    for (int x = 0; x < 8; ++x) {
        c[x] = d < nd ? L->xtot[x * kWideRadix + d] : 0;
        sum += c[x];
    }
    const uint32_t base = block_excl_scan(sum, s_wave);
    if (d < nd) {
        uint64_t o = base;
#pragma unroll
        for (int x = 0; x < 8; ++x) {
            L->xcur[x * kWideRadix + d] = o;
            o += c[x];
        }
        L->part_start[d] = base;
    }
    if (d == 0) L->part_start[nd] = L->region_off[L->nb];
}

__global__ __launch_bounds__(kPartThreads) void rp_scatter1(const HLaunch *__restrict__ L, const Spill *src_p,
                                                           const Spill *dst_p, uint32_t nd) {
    extern __shared__ __align__(16) uint8_t smem[];
    __shared__ uint32_t run[kWideRadix];     // rows of each digit in this tile
    __shared__ uint32_t dstart[kWideRadix];  // LDS start of each digit run
    __shared__ uint64_t gbase[kWideRadix];   // global start of each digit run for this tile
    __shared__ uint32_t s_wave[16];
    const Spill &src = *src_p;
    const Spill &dst = *dst_p;
    const QueryDesc *__restrict__ q = &L->q;
    const int nv = q->n_vals;
    const bool wide = L->wide;
    const uint32_t tid = threadIdx.x;
    const uint32_t ntiles = L->tile_off[L->nb];
    const uint64_t total = L->region_off[L->nb];
    uint64_t *cur = L->xcur + (blockIdx.x & 7) * kWideRadix;
    // LDS staging: keys, [hi], flags, digits (16-bit), then value columns
    uint64_t *st_lo = (uint64_t *)smem;
    uint64_t *st_hi = st_lo + kPartTile;
    uint8_t *p = (uint8_t *)(wide ? st_hi + kPartTile : st_hi);
    uint16_t *st_fl = (uint16_t *)p;
    uint16_t *st_d = st_fl + kPartTile;
    uint8_t *st_v = (uint8_t *)(st_d + kPartTile);
    for (uint32_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        TileSpan tsc;
        tile_span(L, true, t, ntiles, total, tsc);
        run[tid] = 0;
        __syncthreads();
        // the first two value columns load with the keys, so the tile waits
        // on memory once before ranking instead of again while staging
        uint64_t klo[kPartRows], khi[kPartRows], v0[kPartRows], v1[kPartRows];
        uint16_t fl[kPartRows];
        uint32_t dg[kPartRows], rk[kPartRows];
#pragma unroll
        for (int k = 0; k < kPartRows; ++k) {
            const uint32_t i = k * kPartThreads + tid;
            const uint64_t s = tsc.src_begin + (i < tsc.rows ? i : 0);
            klo[k] = gld_nt<uint64_t>(src.klo, s);
            khi[k] = wide ? gld_nt<uint64_t>(src.khi, s) : 0;
            fl[k] = gld_nt<uint16_t>(src.flags, s);
            v0[k] = nv > 0 ? gload_w_nt(src.val[0], L->vw[0], s) : 0;
            v1[k] = nv > 1 ? gload_w_nt(src.val[1], L->vw[1], s) : 0;
        }
#pragma unroll
        for (int k = 0; k < kPartRows; ++k) {
            dg[k] = (uint32_t)spill_hash(klo[k], khi[k], fl[k]) & (nd - 1);
            rk[k] = k * kPartThreads + tid < tsc.rows ? atomicAdd(&run[dg[k]], 1u) : 0xFFFFFFFFu;
        }
        __syncthreads();
        {
            const uint32_t c = run[tid];
            dstart[tid] = block_excl_scan(c, s_wave);
            gbase[tid] = c ? atomicAdd((unsigned long long *)&cur[tid], (unsigned long long)c) : 0;
        }
        __syncthreads();
        // stage in digit order
#pragma unroll
        for (int k = 0; k < kPartRows; ++k) {
            if (rk[k] == 0xFFFFFFFFu) continue;
            const uint64_t s = tsc.src_begin + k * kPartThreads + tid;
            const uint32_t sp = dstart[dg[k]] + rk[k];
            st_lo[sp] = klo[k];
            if (wide) st_hi[sp] = khi[k];
            st_fl[sp] = fl[k];
            st_d[sp] = (uint16_t)dg[k];
            uint8_t *vp = st_v;
            for (int j = 0; j < nv; ++j) {
                const int w = L->vw[j];
                store_w(vp, w, sp, j == 0 ? v0[k] : j == 1 ? v1[k] : gload_w(src.val[j], w, s));
                vp += (size_t)w * kPartTile;
            }
        }
        __syncthreads();
        // write runs: consecutive staged rows -> consecutive positions
        for (uint32_t sp = tid; sp < tsc.rows; sp += kPartThreads) {
            const uint32_t d = st_d[sp];
            const uint64_t g = gbase[d] + (sp - dstart[d]);
            gst<uint64_t>(dst.klo, g, st_lo[sp]);
            if (wide) gst<uint64_t>(dst.khi, g, st_hi[sp]);
            gst<uint16_t>(dst.flags, g, st_fl[sp]);
            const uint8_t *vp = st_v;
            for (int j = 0; j < nv; ++j) {
                const int w = L->vw[j];
                gstore_w(dst.val[j], w, g, load_w(vp, w, sp));
                vp += (size_t)w * kPartTile;
            }
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Second pass of the two-pass layout, MSD: the first pass grouped the rows
// by the low pbits - 7 hash bits (bucket b); this one groups each bucket by
// the next 7 bits (d) inside it, unstably, into partition b * 128 + d.  A
// tile lies inside one bucket (bucket-aligned tiles), so it ranks 128
// digits, and its runs claim their place with one atomic per digit on the
// cursor of (b, d, XCD slot) as the first pass does: runs that meet in one
// sub-region are written through one L2.  Partitions end up contiguous in
// (b, d) order, so their starts come from the counts (no search).
// ---------------------------------------------------------------------------
// Bucket starts (from the first pass), bucket-aligned tiles, and the
// (partition, slot) counts cleared; block 0 maps tiles, every block clears.
__global__ __launch_bounds__(1024) void rp_tiles2(const HLaunch *__restrict__ L, uint32_t nd) {
    const uint32_t P = nd * kRadix;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < P * 8; i += gridDim.x * blockDim.x) L->xtot2[i] = 0;
    if (blockIdx.x != 0) return;
    __shared__ uint32_t s_wave[16];
    const uint32_t b = threadIdx.x;  // nd <= 1024
    const uint64_t s0 = b < nd ? L->part_start[b] : 0, s1 = b < nd ? L->part_start[b + 1] : 0;
    const uint32_t nt = b < nd ? (uint32_t)((s1 - s0 + kPartTile - 1) / kPartTile) : 0;
    const uint32_t t0 = block_excl_scan(nt, s_wave);
    if (b < nd) {
        L->bstart2[b] = s0;
        L->tile_off2[b] = t0;
        for (uint32_t j = 0; j < nt; ++j) L->tile_reg2[t0 + j] = b;
    }
    if (b + 1 == nd) {
        L->bstart2[nd] = s1;
        L->tile_off2[nd] = t0 + nt;
    }
}

// Tile t of the second pass: its bucket, first row and rows.
__device__ __forceinline__ void tile2(const HLaunch *__restrict__ L, uint32_t t, uint32_t &b, uint64_t &begin,
                                      uint32_t &rows) {
    b = L->tile_reg2[t];
    begin = L->bstart2[b] + (uint64_t)(t - L->tile_off2[b]) * kPartTile;
    rows = (uint32_t)min<uint64_t>(kPartTile, L->bstart2[b + 1] - begin);
}

// Rows per (partition, XCD slot), slot-minor: xtot2[(b * 128 + d) * 8 + slot].
__global__ __launch_bounds__(kPartThreads) void rp_hist2(const HLaunch *__restrict__ L, const Spill *src_p, uint32_t nd,
                                                        int shift) {
    __shared__ uint32_t hist[kRadix];
    const Spill &src = *src_p;
    const bool wide = L->wide;
    const uint32_t tid = threadIdx.x, slot = blockIdx.x & 7;
    const uint32_t ntiles = L->tile_off2[nd];
    for (uint32_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        uint32_t b, rows;
        uint64_t begin;
        tile2(L, t, b, begin, rows);
        if (tid < kRadix) hist[tid] = 0;
        __syncthreads();
        uint32_t d[kPartRows];
#pragma unroll
        for (int k = 0; k < kPartRows; ++k) {
            const uint32_t i = k * kPartThreads + tid;
            const uint64_t s = begin + (i < rows ? i : 0);
            d[k] = pass_digit(gld_nt<uint64_t>(src.klo, s), wide ? gld_nt<uint64_t>(src.khi, s) : 0,
                              gld_nt<uint16_t>(src.flags, s), shift);
        }
#pragma unroll
        for (int k = 0; k < kPartRows; ++k)
            if (k * kPartThreads + tid < rows) atomicAdd(&hist[d[k]], 1u);
        __syncthreads();
        if (tid < kRadix && hist[tid]) atomicAdd(&L->xtot2[((uint64_t)b * kRadix + tid) * 8 + slot], hist[tid]);
        __syncthreads();
    }
}

// Per 1024 partitions (one thread each, its eight slot counts) the rows
// they hold, into bsum2.
__global__ __launch_bounds__(1024) void rp_sum2(const HLaunch *__restrict__ L, uint32_t P2) {
    __shared__ uint32_t s_wave[16];
    const uint32_t c = blockIdx.x * 1024 + threadIdx.x;
    uint32_t s = 0;
    if (c < P2) {
        const uint4 a = *(const uint4 *)(L->xtot2 + (uint64_t)c * 8);
        const uint4 b = *(const uint4 *)(L->xtot2 + (uint64_t)c * 8 + 4);
        s = a.x + a.y + a.z + a.w + b.x + b.y + b.z + b.w;
    }
    const uint32_t e = block_excl_scan(s, s_wave);
    if (threadIdx.x == 1023) L->bsum2[blockIdx.x] = e + s;
}

// Each block of 1024 partitions: its first row (the blocks before it, by
// one wave), then each partition's start and its slots' cursors (the
// exclusive prefix of the (partition, slot) counts in order).
__global__ __launch_bounds__(1024) void rp_scan2(const HLaunch *__restrict__ L, uint32_t P2) {
    __shared__ uint32_t s_wave[16];
    __shared__ uint64_t s_off;
    if (threadIdx.x < 64) {
        uint64_t o = 0;
        for (uint32_t j = threadIdx.x; j < blockIdx.x; j += 64) o += L->bsum2[j];
#pragma unroll
        for (int k = 32; k; k >>= 1) o += __shfl_xor(o, k);
        if (threadIdx.x == 0) s_off = o;
    }
    __syncthreads();
    const uint32_t c = blockIdx.x * 1024 + threadIdx.x;
    uint32_t cnt[8], s = 0;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        cnt[k] = c < P2 ? L->xtot2[(uint64_t)c * 8 + k] : 0;
        s += cnt[k];
    }
    const uint64_t base = s_off + block_excl_scan(s, s_wave);
    if (c < P2) {
        L->part_start[c] = base;
        uint64_t o = base;
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            L->xcur2[(uint64_t)c * 8 + k] = o;
            o += cnt[k];
        }
    }
    if (c + 1 == P2) L->part_start[P2] = base + s;
}

__global__ __launch_bounds__(kPartThreads, 8) void rp_scatter2(const HLaunch *__restrict__ L, const Spill *src_p,
                                                           const Spill *dst_p, uint32_t nd, int shift) {
    extern __shared__ __align__(16) uint8_t smem[];
    __shared__ uint32_t run[kRadix];     // rows of each digit in this tile
    __shared__ uint32_t dstart[kRadix];  // LDS start of each digit run
    __shared__ uint64_t gbase[kRadix];   // global start of each digit run for this tile
    __shared__ uint32_t s_wave[16];
    const Spill &src = *src_p;
    const Spill &dst = *dst_p;
    const QueryDesc *__restrict__ q = &L->q;
    const int nv = q->n_vals;
    const bool wide = L->wide;
    const uint32_t tid = threadIdx.x, slot = blockIdx.x & 7;
    const uint32_t ntiles = L->tile_off2[nd];
    // LDS staging: keys, [hi], flags, digits, then value columns
    uint64_t *st_lo = (uint64_t *)smem;
    uint64_t *st_hi = st_lo + kPartTile;
    uint8_t *p = (uint8_t *)(wide ? st_hi + kPartTile : st_hi);
    uint16_t *st_fl = (uint16_t *)p;
    uint8_t *st_d = (uint8_t *)(st_fl + kPartTile);
    uint8_t *st_v = st_d + kPartTile;
    for (uint32_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        uint32_t b, rows;
        uint64_t begin;
        tile2(L, t, b, begin, rows);
        if (tid < kRadix) run[tid] = 0;
        __syncthreads();
        // the first two value columns load with the keys: one wait on
        // memory per tile before ranking
        uint64_t klo[kPartRows], khi[kPartRows], v0[kPartRows], v1[kPartRows];
        uint16_t fl[kPartRows];
        uint32_t dg[kPartRows], rk[kPartRows];
#pragma unroll
        for (int k = 0; k < kPartRows; ++k) {
            const uint32_t i = k * kPartThreads + tid;
            const uint64_t s = begin + (i < rows ? i : 0);
            klo[k] = gld_nt<uint64_t>(src.klo, s);
            khi[k] = wide ? gld_nt<uint64_t>(src.khi, s) : 0;
            fl[k] = gld_nt<uint16_t>(src.flags, s);
            v0[k] = nv > 0 ? gload_w_nt(src.val[0], L->vw[0], s) : 0;
            v1[k] = nv > 1 ? gload_w_nt(src.val[1], L->vw[1], s) : 0;
        }
#pragma unroll
        for (int k = 0; k < kPartRows; ++k) {
            dg[k] = pass_digit(klo[k], khi[k], fl[k], shift);
            rk[k] = k * kPartThreads + tid < rows ? atomicAdd(&run[dg[k]], 1u) : 0xFFFFFFFFu;
        }
        __syncthreads();
        {
            const uint32_t c = tid < kRadix ? run[tid] : 0;
            const uint32_t e = block_excl_scan(c, s_wave);
            if (tid < kRadix) {
                dstart[tid] = e;
                gbase[tid] = c ? atomicAdd((unsigned long long *)&L->xcur2[((uint64_t)b * kRadix + tid) * 8 + slot],
                                           (unsigned long long)c)
                               : 0;
            }
        }
        __syncthreads();
        // stage in digit order
#pragma unroll
        for (int k = 0; k < kPartRows; ++k) {
            if (rk[k] == 0xFFFFFFFFu) continue;
            const uint64_t s = begin + k * kPartThreads + tid;
            const uint32_t sp = dstart[dg[k]] + rk[k];
            st_lo[sp] = klo[k];
            if (wide) st_hi[sp] = khi[k];
            st_fl[sp] = fl[k];
            st_d[sp] = (uint8_t)dg[k];
            uint8_t *vp = st_v;
            for (int j = 0; j < nv; ++j) {
                const int w = L->vw[j];
                store_w(vp, w, sp, j == 0 ? v0[k] : j == 1 ? v1[k] : gload_w(src.val[j], w, s));
                vp += (size_t)w * kPartTile;
            }
        }
        __syncthreads();
        // write runs: consecutive staged rows -> consecutive positions
        for (uint32_t sp = tid; sp < rows; sp += kPartThreads) {
            const uint32_t d = st_d[sp];
            const uint64_t g = gbase[d] + (sp - dstart[d]);
            gst<uint64_t>(dst.klo, g, st_lo[sp]);
            if (wide) gst<uint64_t>(dst.khi, g, st_hi[sp]);
            gst<uint16_t>(dst.flags, g, st_fl[sp]);
            const uint8_t *vp = st_v;
            for (int j = 0; j < nv; ++j) {
                const int w = L->vw[j];
                gstore_w(dst.val[j], w, g, load_w(vp, w, sp));
                vp += (size_t)w * kPartTile;
            }
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// per-partition aggregation
// ---------------------------------------------------------------------------
// The passes order rows by the low `sortbits` bits of the hash; a partition
// is the top pbits of that field, so partitions are contiguous runs.
__device__ __forceinline__ uint32_t part_of(const HLaunch *__restrict__ L, const Spill &S, bool wide, uint64_t i) {
    const uint64_t v = spill_hash(gld<uint64_t>(S.klo, i), wide ? gld<uint64_t>(S.khi, i) : 0, gld<uint16_t>(S.flags, i)) &
                       ((1ull << L->sortbits) - 1);
    return (uint32_t)(v >> (L->sortbits - L->pbits));
}

// First index in [0, n) whose partition is >= p (the array is ordered by
// partition after the passes).
__device__ __forceinline__ uint64_t part_lower_bound(const HLaunch *__restrict__ L, const Spill &S, bool wide,
                                                     uint64_t n, uint32_t p) {
    uint64_t lo = 0, hi = n;
    while (lo < hi) {
        const uint64_t mid = (lo + hi) / 2;
        if (part_of(L, S, wide, mid) < p) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// Partition starts after stable multi-digit passes: one binary search per
// partition, all in parallel.
__global__ __launch_bounds__(256) void hs_part_bounds(const HLaunch *__restrict__ L, const Spill *src_p) {
    const Spill &S = *src_p;
    const uint32_t P = 1u << L->pbits;
    const uint64_t n = L->region_off[L->nb];
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p <= P; p += gridDim.x * blockDim.x)
        L->part_start[p] = p == 0 ? 0 : p == P ? n : part_lower_bound(L, S, L->wide, n, p);
}

// A tail row whose key found no slot in its partition's LDS table folds
// into the global fallback table: rare, so kept out of line.
// Scalar arguments only: structures passed by value to an out-of-line call
// go through the stack, which gave every instantiation of the caller 64
// bytes of scratch per lane.
__device__ __noinline__ void part_fallback(const HLaunch *__restrict__ L, const Spill *S, uint64_t i, uint32_t fl,
                                           uint64_t v0, uint64_t v1, uint64_t klo, uint64_t khi, int special,
                                           uint64_t h) {
    GSink gs{L->t.arr, L->t.mask + 3};
    const Key k{klo, khi, special};
    const SpillRowV row{L, S, i, (uint16_t)fl, v0, v1};
    const int64_t g = L->wide ? g_find128(L->t, k, h) : g_find64(L->t, k, h);
    if (g >= 0) add_row(gs, &L->q, (uint64_t)g, row);
}

// WIDE: keys of more than 64 bits; SIG: the aggregate signature (-1: the
// general fold, for more than two value columns).
template <bool WIDE, int SIG>
__global__ __launch_bounds__(1024) void hs_part_agg(const HLaunch *__restrict__ L, const Spill *src_p) {
    extern __shared__ __align__(16) uint8_t smem[];
    __shared__ uint64_t s_range[2];
    __shared__ uint32_t s_base, s_ntot, s_pwave[16];
    const Spill &S = *src_p;
    const QueryDesc *__restrict__ q = as_const(&L->q);
    constexpr bool wide = WIDE;
    const uint32_t CL = L->tail_slots;
    const uint32_t nls = CL + 2;
    const uint32_t tid = threadIdx.x, bs = blockDim.x;
    const uint32_t pmask = (1u << L->pbits) - 1;
    uint64_t *lk = (uint64_t *)smem;
    uint64_t *lh = lk + nls;
    uint32_t *lst = (uint32_t *)(lh + nls);  // wide: 0 free, 1 writing, 2 published
    uint8_t *state = wide ? (uint8_t *)(lst + nls) : (uint8_t *)lh;
    LSink ls{state, L->lds3_off, nls};
    const LdsAgg2 A = lds_agg2(q, state, L->lds3_off, nls);
    for (uint32_t p = blockIdx.x; p <= pmask; p += gridDim.x) {
        if (tid < 2) s_range[tid] = L->part_start[p + tid];
        for (uint32_t i = tid; i < nls; i += bs) {
            lk[i] = kEmpty;
            if (wide) { lh[i] = kEmpty; lst[i] = 0; }
        }
        for (int a = 0; a < q->layout.n_arrs; ++a) {
            const int op = q->layout.init[a];
            if (mo_size(op) == 4) {
                const uint32_t id = (uint32_t)mo_identity(op);
                for (uint32_t i = tid; i < nls; i += bs) *ls.p<uint32_t>(a, i) = id;
            } else {
                const uint64_t id = mo_identity(op);
                for (uint32_t i = tid; i < nls; i += bs) *ls.p<uint64_t>(a, i) = id;
            }
        }
        __syncthreads();
        const uint64_t b = s_range[0], e = s_range[1];
        // four rows per thread per step, every load issued before the first
        // probe
        constexpr int U = 4;
        const int nv = q->n_vals;
        for (uint64_t i0 = b + tid; i0 < e; i0 += (uint64_t)U * bs) {
            uint16_t flb[U];
            uint64_t klb[U], khb[U], v0b[U], v1b[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const uint64_t i = i0 + (uint64_t)u * bs;
                const uint64_t ii = i < e ? i : b;
                flb[u] = gld_nt<uint16_t>(S.flags, ii);
                klb[u] = gld_nt<uint64_t>(S.klo, ii);
                khb[u] = wide ? gld_nt<uint64_t>(S.khi, ii) : 0;
                v0b[u] = nv > 0 ? gload_w_nt(S.val[0], L->vw[0], ii) : 0;
                v1b[u] = nv > 1 ? gload_w_nt(S.val[1], L->vw[1], ii) : 0;
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const uint64_t i = i0 + (uint64_t)u * bs;
                if (i >= e) break;
                const uint16_t fl = flb[u];
                SpillRowV row{L, &S, i, fl, v0b[u], v1b[u]};
                Key k{klb[u], khb[u], (fl & kFlagAllOnes) ? 1 : ((fl & kFlagNull) ? 2 : 0)};
                int32_t slot = -1;
                uint64_t h = 0;
                if (k.special) {
                    slot = (int32_t)(CL + k.special - 1);
                } else {
                    h = hash_key(k.lo, k.hi);
                    uint32_t s = (uint32_t)(h >> 20) & (CL - 1);
                    for (int pr = 0; pr < kProbe && slot < 0; ++pr) {
                        if (!wide) {
                            uint64_t cur = __hip_atomic_load(&lk[s], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                            if (cur == kEmpty) {
                                unsigned long long old = atomicCAS((unsigned long long *)&lk[s],
                                                                   (unsigned long long)kEmpty, (unsigned long long)k.lo);
                                if (old == kEmpty || old == k.lo) { slot = (int32_t)s; break; }
                            } else if (cur == k.lo) {
                                slot = (int32_t)s;
                                break;
                            }
                        } else {
                            // A claimer publishes within its own loop iteration,
                            // so the wait ends; the bound turns a broken
                            // invariant into a reported error instead of a hang.
                            for (uint32_t spin = 0;; ++spin) {
                                if (spin > (1u << 22)) { atomicOr(&L->stats[ST_ERROR], 1u); break; }
                                uint32_t st = __hip_atomic_load(&lst[s], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP);
                                if (st == 2u) {
                                    if (lk[s] == k.lo && lh[s] == k.hi) slot = (int32_t)s;
                                    break;
                                }
                                if (st == 0u) {
                                    if (atomicCAS(&lst[s], 0u, 1u) == 0u) {
                                        lk[s] = k.lo;
                                        lh[s] = k.hi;
                                        __hip_atomic_store(&lst[s], 2u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
                                        slot = (int32_t)s;
                                        break;
                                    }
                                }
                                // claimed by another lane: re-read
                            }
                            if (slot >= 0) break;
                        }
                        s = (s + 1) & (CL - 1);
                    }
                }
                if (slot >= 0) {
                    if constexpr (SIG >= 0) lds_agg2_add_sig<SIG>(A, (uint32_t)slot, fl, v0b[u], v1b[u]);
                    else add_row(ls, q, (uint32_t)slot, row);
                } else {
                    part_fallback(L, &S, i, fl, v0b[u], v1b[u], k.lo, k.hi, k.special, h);
                }
            }
        }
        __syncthreads();
        // emit: compact the used slots of this partition into the outputs
        LReader rd{state, L->lds3_off};
        const int a_len = q->layout.a_len;
        const uint32_t iters = (nls + bs - 1) / bs;
        uint32_t my = 0;
        for (uint32_t it = 0; it < iters; ++it) {
            const uint32_t i = it * bs + tid;
            if (i < nls && *ls.p<uint32_t>(a_len, i) != 0) ++my;
        }
        // thread-ordered offsets, so a wave's groups land side by side
        const uint32_t my_off = block_excl_scan(my, s_pwave);
        if (tid == bs - 1) {
            s_ntot = my_off + my;
            s_base = my_off + my ? atomicAdd(L->out_cursor, my_off + my) : 0;
        }
        if (L->live_off) {
            // the live slots listed densely, so the groups are emitted by as
            // few waves as they fill, in one pass
            uint16_t *live = (uint16_t *)(smem + L->live_off);
            uint32_t pos = my_off;
            for (uint32_t it = 0; it < iters; ++it) {
                const uint32_t i = it * bs + tid;
                if (i < nls && *ls.p<uint32_t>(a_len, i) != 0) live[pos++] = (uint16_t)i;
            }
            __syncthreads();
            const uint32_t n_live = s_ntot, base = s_base;
            for (uint32_t j = tid; j < n_live; j += bs) {
                const uint32_t i = live[j];
                const int special = i >= CL ? (int)(i - CL) + 1 : 0;
                emit_group(q, L->out, L->n_out, rd, i, special ? 0 : lk[i], (special || !wide) ? 0 : lh[i],
                           special, base + j);
            }
        } else {
            __syncthreads();
            uint32_t o = s_base + my_off;
            for (uint32_t it = 0; it < iters; ++it) {
                const uint32_t i = it * bs + tid;
                if (i < nls && *ls.p<uint32_t>(a_len, i) != 0) {
                    const int special = i >= CL ? (int)(i - CL) + 1 : 0;
                    emit_group(q, L->out, L->n_out, rd, i, special ? 0 : lk[i], (special || !wide) ? 0 : lh[i],
                               special, o);
                    ++o;
                }
            }
        }
        __syncthreads();
    }
}

// Hot groups: slots of the global hot state that received rows.
__global__ __launch_bounds__(256) void hs_hot_emit(const HLaunch *__restrict__ L) {
    __shared__ uint32_t s_count, s_base;
    const QueryDesc *__restrict__ q = &L->q;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t CL = L->hot_slots;
    if (threadIdx.x == 0) s_count = 0;
    __syncthreads();
    GReader rd{L->hot_arr};
    const bool used = i < CL && rd.get<uint32_t>(q->layout.a_len, i) != 0;
    uint32_t off = 0;
    if (used) off = atomicAdd(&s_count, 1u);
    __syncthreads();
    if (threadIdx.x == 0) s_base = s_count ? atomicAdd(L->out_cursor, s_count) : 0;
    __syncthreads();
    if (!used) return;
    if (i >= CL - 2) {  // side slots: the all-ones key, the null key
        const int special = (int)(i - (CL - 3));
        emit_group(q, L->out, L->n_out, rd, i, special == 1 ? kEmpty : 0, special == 1 && L->wide ? kEmpty : 0,
                   special, s_base + off);
    } else {
        emit_group(q, L->out, L->n_out, rd, i, L->hot_lo[i], L->wide ? L->hot_hi[i] : 0, 0, s_base + off);
    }
}

// Groups of the fallback table, written after the cursor's groups; resets
// the slots it reads.
__global__ __launch_bounds__(256) void hs_fallback_emit(const HLaunch *__restrict__ L) {
    const QueryDesc *__restrict__ q = &L->q;
    const GTable &t = L->t;
    const Layout &lay = q->layout;
    GReader rd{t.arr};
    uint64_t ng = *(volatile uint32_t *)&t.ctrl[0];
    if (ng > t.list_cap) ng = t.list_cap;
    const uint64_t base = *(volatile uint32_t *)L->out_cursor;
    const uint64_t cap = t.mask + 1;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < ng; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t slot = t.list[i];
        int special = slot == cap ? 1 : (slot == cap + 1 ? 2 : 0);
        const uint64_t lo = special ? 0 : t.klo[slot];
        const uint64_t hi = (special || !L->wide) ? 0 : t.khi[slot];
        emit_group(q, L->out, L->n_out, rd, slot, lo, hi, special, base + i);
        t.klo[slot] = kEmpty;
        if (L->wide) { t.khi[slot] = kEmpty; t.kstate[slot] = 0; }
        for (int a = 0; a < lay.n_arrs; ++a) {
            const int op = lay.init[a];
            if (mo_size(op) == 4) ((uint32_t *)t.arr[a])[slot] = (uint32_t)mo_identity(op);
            else ((uint64_t *)t.arr[a])[slot] = mo_identity(op);
        }
    }
}

// Validity bytes -> bitmaps, Bool key bytes -> bit-packed data.  One thread
// per output byte; the group count is the cursor plus the fallback groups.
__global__ void hs_pack(const HLaunch *__restrict__ L) {
    uint64_t fb = *(volatile uint32_t *)&L->t.ctrl[0];
    if (fb > L->t.list_cap) fb = L->t.list_cap;
    const uint64_t ng = (uint64_t)*(volatile uint32_t *)L->out_cursor + fb;
    const uint64_t nbytes = (ng + 7) / 8;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nbytes;
         i += (uint64_t)gridDim.x * blockDim.x)
    for (int o = 0; o < L->n_out; ++o) {
        const OutCol &oc = L->out[o];
        if (oc.vbytes && oc.valid) {
            uint8_t b = 0;
            for (int k = 0; k < 8; ++k) {
                const uint64_t r = i * 8 + k;
                if (r < ng && oc.vbytes[r]) b |= (uint8_t)(1u << k);
            }
            ((uint8_t *)oc.valid)[i] = b;
        }
        if (oc.src == OS_KEY && oc.tag == PMH_BOOL && oc.data_bits) {
            uint8_t b = 0;
            for (int k = 0; k < 8; ++k) {
                const uint64_t r = i * 8 + k;
                if (r < ng && ((const uint8_t *)oc.data)[r]) b |= (uint8_t)(1u << k);
            }
            oc.data_bits[i] = b;
        }
    }
}

// The query's control words, statistics and group count into the host's
// pinned mailbox, then the control words cleared for the next query (64
// threads).
__global__ void hs_clear_ctrl(const HLaunch *__restrict__ L) {
    const uint32_t i = threadIdx.x;
    if (i < 8) L->mbox[i] = L->t.ctrl[i];
    if (i < ST_N) L->mbox[8 + i] = L->stats[i];
    if (i == 0) L->mbox[8 + ST_N] = L->out_cursor[0];
    __threadfence_system();
    __syncthreads();
    if (i < 8) L->t.ctrl[i] = 0;
}


__global__ void hs_reset(const HLaunch *__restrict__ L) {
    if (threadIdx.x < ST_N) L->stats[threadIdx.x] = 0;
    if (threadIdx.x == 0) L->out_cursor[0] = 0;
}

__global__ void hs_init_table(const HLaunch *__restrict__ L, uint64_t nslots) {
    const GTable &t = L->t;
    const Layout &lay = L->q.layout;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nslots; i += stride) {
        t.klo[i] = kEmpty;
        if (L->wide) { t.khi[i] = kEmpty; t.kstate[i] = 0; }
        for (int a = 0; a < lay.n_arrs; ++a) {
            const int op = lay.init[a];
            if (mo_size(op) == 4) ((uint32_t *)t.arr[a])[i] = (uint32_t)mo_identity(op);
            else ((uint64_t *)t.arr[a])[i] = mo_identity(op);
        }
    }
    if (blockIdx.x == 0 && threadIdx.x < 8) t.ctrl[threadIdx.x] = 0;
}

// ---------------------------------------------------------------------------
// host
// ---------------------------------------------------------------------------
namespace {

struct FallbackState {
    uint64_t cap = 0;
    std::string sig;
    bool clean = false;
};
FallbackState g_fb;

// g_fb describes the fallback table in the "hs:fb_table" slab.  An
// allocation-free call builds its table in the scratch array instead: it
// starts from an unknown (so re-initialised) table and leaves g_fb as it
// found it.
struct FallbackScratchGuard {
    FallbackState saved;
    bool on;
    explicit FallbackScratchGuard(bool scratch) : saved(g_fb), on(scratch) {
        if (on) g_fb.clean = false;
    }
    ~FallbackScratchGuard() {
        if (on) g_fb = saved;
    }
};

uint64_t pow2_floor(uint64_t x) {
    uint64_t p = 1;
    while (p * 2 <= x) p *= 2;
    return p;
}
uint64_t pow2_ceil(uint64_t x) {
    uint64_t p = 1;
    while (p < x) p *= 2;
    return p;
}
int log2u(uint64_t x) {
    int b = 0;
    while ((1ull << b) < x) ++b;
    return b;
}

// LDS offsets for nslots entries per array, arrays back to back.
size_t lds_layout(const Prep &P, uint32_t nslots, uint32_t *off) {
    const Layout &lay = P.q.layout;
    size_t o = 0;
    for (int a = 0; a < lay.n_arrs; ++a) {
        off[a] = (uint32_t)o;
        o += (size_t)nslots * lay.esz[a];
    }
    return o;
}

}  // namespace

// Carves nslots-entry state arrays from one slab; limb k > 0 of an exact
// accumulator sits nslots entries after limb k-1.
void carve_state(const Prep &P, uint8_t *base, size_t &o, uint64_t nslots, void **arr) {
    const Layout &lay = P.q.layout;
    for (int a = 0; a < lay.n_arrs; ++a) {
        if (P.limb_base[a] >= 0 && P.limb_k[a] > 0) {
            arr[a] = (uint8_t *)arr[P.limb_base[a]] + (size_t)P.limb_k[a] * nslots * 8;
        } else {
            const size_t b = P.limb_base[a] >= 0 ? nslots * 8 * kLimbs : nslots * lay.esz[a];
            arr[a] = base + o;
            o += (b + 255) & ~size_t(255);
        }
    }
}
size_t state_bytes(const Prep &P, uint64_t nslots) {
    const Layout &lay = P.q.layout;
    size_t o = 0;
    for (int a = 0; a < lay.n_arrs; ++a) {
        if (P.limb_base[a] >= 0 && P.limb_k[a] > 0) continue;
        const size_t b = P.limb_base[a] >= 0 ? nslots * 8 * kLimbs : nslots * lay.esz[a];
        o += (b + 255) & ~size_t(255);
    }
    return o;
}

Sketch sketch_from(uint32_t s_sel, uint32_t n_sample, uint64_t n, uint32_t d, uint32_t f1, uint32_t f2,
                   uint32_t dropped, const std::vector<uint32_t> &hist) {
    Sketch k = sketch_head(s_sel, n_sample, n, d, f1, f2, dropped);
    const double qf = sketch_qf(k);
    double F2 = 0, num = 0, den = 0;
    // hist[c]: sampled keys seen c times (c < 1024; a shorter histogram is
    // zero past its end)
    const int nc = (int)std::min<size_t>(hist.size(), 1024);
    for (int c = 1; c < nc; ++c)
        sketch_bucket(c, (double)hist[c] + (c == 1 ? (double)dropped : 0.0), qf, F2, num, den);
    k.F2 = F2;
    k.shl_num = num;
    k.shl_den = den;
    return k;
}

// Lower the predicate for hs_scan_fast: an AND tree whose leaves compare one
// column with a literal (any operator but Ne), test a Boolean column, or are
// the literal true.  Leaves on the same column and compare type intersect.
// Returns false when the predicate has another shape.
bool lower_fast_scan(const QueryDesc &q, FastScan &fs) {
    fs = FastScan{};
    if (q.n_pred == 0) return true;
    const int64_t MN = INT64_MIN, MX = INT64_MAX;
    auto add = [&](int col, int flt, int64_t lo, int64_t hi) {
        for (int r = 0; r < fs.nr; ++r)
            if (fs.rcol[r] == col && fs.rflt[r] == flt) {
                fs.rlo[r] = std::max(fs.rlo[r], lo);
                fs.rhi[r] = std::min(fs.rhi[r], hi);
                return true;
            }
        if (fs.nr == 2) return false;
        fs.rcol[fs.nr] = col;
        fs.rflt[fs.nr] = flt;
        fs.rlo[fs.nr] = lo;
        fs.rhi[fs.nr] = hi;
        ++fs.nr;
        return true;
    };
    std::vector<int> stack{q.n_pred - 1};
    while (!stack.empty()) {
        const PredNode &nd = q.pred[stack.back()];
        stack.pop_back();
        switch (nd.kind) {
        case PK_AND: stack.push_back(nd.lhs); stack.push_back(nd.rhs); break;
        case PK_CONST: if (!nd.cval) return false; break;
        case PK_COLBOOL:
            if (q.cols[nd.lhs].tag != PMH_BOOL || !add(nd.lhs, 0, 1, 1)) return false;
            break;
        case PK_CMP: {
            const bool lcol = nd.lkind == OK_COL, rcol = nd.rkind == OK_COL;
            if (lcol == rcol) return false;
            int op = nd.op;
            if (!lcol) op = op == PMH_LT ? PMH_GT : op == PMH_GT ? PMH_LT : op == PMH_LE ? PMH_GE : op == PMH_GE ? PMH_LE : op;
            const int col = lcol ? nd.lhs : nd.rhs;
            int64_t k;
            if (nd.is_f64) {
                const double x = lcol ? nd.rf : nd.lf;
                int64_t b;
                memcpy(&b, &x, 8);
                k = tot_key_bits(b, std::isnan(x));
            } else {
                k = lcol ? nd.ri : nd.li;
            }
            int64_t lo, hi;
            switch (op) {
            case PMH_EQ: lo = hi = k; break;
            case PMH_LT:
                if (k == MN) { lo = MX; hi = MN; }  // empty
                else { lo = MN; hi = k - 1; }
                break;
            case PMH_LE: lo = MN; hi = k; break;
            case PMH_GT:
                if (k == MX) { lo = MX; hi = MN; }  // empty
                else { lo = k + 1; hi = MX; }
                break;
            case PMH_GE: lo = k; hi = MX; break;
            default: return false;
            }
            if (!add(col, nd.is_f64 ? 1 : 0, lo, hi)) return false;
            break;
        }
        default: return false;
        }
    }
    return true;
}

// ---------------------------------------------------------------------------
// key narrowing: a lane wider than 64 bits whose columns' value ranges fit
// 64 bits runs as a narrow lane (8-byte keys through the whole engine)
// ---------------------------------------------------------------------------
// kr[3k]: smallest image of key k, kr[3k + 1]: largest, kr[3k + 2]: nonzero
// if a row of key k is null; over every row (the predicate is not applied,
// so the ranges cover any selection).  kr starts at (~0, 0, 0).
__global__ __launch_bounds__(256) void hs_key_range(const QueryDesc *__restrict__ q, unsigned long long *kr) {
    const uint64_t n = q->n_rows;
    for (int k = 0; k < q->n_keys; ++k) {
        const KeyField &f = q->keys[k];
        const DevCol &c = q->cols[f.col];
        const int xf = tag_signed(f.tag) ? KX_SIGNED : KX_UNSIGNED;
        uint64_t lo = ~0ull, hi = 0;
        uint32_t nul = 0;
        for (uint64_t r = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; r < n; r += (uint64_t)gridDim.x * blockDim.x) {
            if (!dvalid(c.valid, r)) {
                nul = 1;
                continue;
            }
            const uint64_t m = key_image(xf, f.tag, dload_bits(c, r));
            lo = m < lo ? m : lo;
            hi = m > hi ? m : hi;
        }
#pragma unroll
        for (int o = 32; o; o >>= 1) {
            const uint64_t a = __shfl_xor(lo, o), b = __shfl_xor(hi, o);
            lo = a < lo ? a : lo;
            hi = b > hi ? b : hi;
            nul |= __shfl_xor(nul, o);
        }
        if ((threadIdx.x & 63) == 0) {
            if (lo != ~0ull) atomicMin(&kr[3 * k], (unsigned long long)lo);
            if (hi) atomicMax(&kr[3 * k + 1], (unsigned long long)hi);
            if (nul) atomicOr(&kr[3 * k + 2], 1ull);
        }
    }
}

// hs_key_range over kNarrowSample evenly spaced rows, one per thread: the
// ranges the scan then checks every selected key against.
constexpr uint32_t kNarrowSample = 65536;
__global__ __launch_bounds__(1024) void hs_key_range_sample(const QueryDesc *__restrict__ q, unsigned long long *kr) {
    const uint64_t n = q->n_rows;
    const uint32_t S = (uint32_t)min<uint64_t>(n, kNarrowSample);
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t r = i < S ? (uint64_t)i * n / S : 0;
    for (int k = 0; k < q->n_keys; ++k) {
        const KeyField &f = q->keys[k];
        const DevCol &c = q->cols[f.col];
        const int xf = tag_signed(f.tag) ? KX_SIGNED : KX_UNSIGNED;
        uint64_t lo = ~0ull, hi = 0;
        uint32_t nul = 0;
        if (i < S) {
            if (!dvalid(c.valid, r)) {
                nul = 1;
            } else {
                lo = hi = key_image(xf, f.tag, dload_bits(c, r));
            }
        }
#pragma unroll
        for (int o = 32; o; o >>= 1) {
            const uint64_t a = __shfl_xor(lo, o), b = __shfl_xor(hi, o);
            lo = a < lo ? a : lo;
            hi = b > hi ? b : hi;
            nul |= __shfl_xor(nul, o);
        }
        if ((threadIdx.x & 63) == 0) {
            if (lo != ~0ull) atomicMin(&kr[3 * k], (unsigned long long)lo);
            if (hi) atomicMax(&kr[3 * k + 1], (unsigned long long)hi);
            if (nul) atomicOr(&kr[3 * k + 2], 1ull);
        }
    }
}

__global__ void hs_key_range_init(unsigned long long *kr, int n_keys) {
    const int i = threadIdx.x;
    if (i < 3 * n_keys) kr[i] = i % 3 == 0 ? ~0ull : 0ull;
}

namespace {

// Rows from which the narrowing pass (one read of the key columns and a
// host synchronisation) is worth trying.
constexpr uint64_t kNarrowMinRows = uint64_t(1) << 20;

// Builds in qn the lane of q narrowed to the key columns' value ranges:
// each column's null bit only if it holds a null, its data as the image
// minus the smallest image, in the bits the range needs.  False (qn
// unchanged) when the narrowed lane is still wider than 64 bits.  sampled:
// the ranges of an evenly spaced sample, which the specialised scan checks
// every selected key against (qn.narrow_check).
bool narrow_lane(const QueryDesc &q, QueryDesc &qn, bool sampled) {
    Runtime &R = rt();
    const int nk = q.n_keys;
    auto *dq = (QueryDesc *)slab("hs:narrow_q", sizeof(QueryDesc));
    auto *kr = (unsigned long long *)slab("hs:narrow_kr", 3 * kMaxKeys * 8);
    auto *pin = (unsigned long long *)pinned_host(3 * kMaxKeys * 8 + sizeof(QueryDesc));
    memcpy(pin + 3 * kMaxKeys, &q, sizeof(QueryDesc));
    PMH_HIP(hipMemcpyAsync(dq, pin + 3 * kMaxKeys, sizeof(QueryDesc), hipMemcpyHostToDevice, R.stream));
    hipLaunchKernelGGL(hs_key_range_init, dim3(1), dim3(64), 0, R.stream, kr, nk);
    PMH_HIP(hipGetLastError());
    const int cus = R.prop.multiProcessorCount > 0 ? R.prop.multiProcessorCount : 256;
    if (sampled) {
        KernelTimer kt("hs_key_range");
        hipLaunchKernelGGL(hs_key_range_sample, dim3(kNarrowSample / 1024), dim3(1024), 0, R.stream, dq, kr);
        PMH_HIP(hipGetLastError());
    } else {
        KernelTimer kt("hs_key_range");
        hipLaunchKernelGGL(hs_key_range, dim3(cus * 8), dim3(256), 0, R.stream, dq, kr);
        PMH_HIP(hipGetLastError());
    }
    PMH_HIP(hipMemcpyAsync(pin, kr, 3 * nk * 8, hipMemcpyDeviceToHost, R.stream));
    PMH_HIP(hipStreamSynchronize(R.stream));
    int off = 0;
    KeyField keys[kMaxKeys];
    for (int k = 0; k < nk; ++k) {
        KeyField f = q.keys[k];
        const uint64_t lo = pin[3 * k], hi = pin[3 * k + 1];
        // sampled: a column with a validity bitmap keeps a null bit even if
        // the sample saw no null
        const bool nul = pin[3 * k + 2] != 0 || (sampled && q.cols[f.col].valid != nullptr);
        const bool any = lo <= hi;  // some valid row
        int w = !any || hi == lo ? 0 : 64 - __builtin_clzll(hi - lo);
        uint64_t bias = any ? lo : 0;
        if (sampled && any) {
            // the sampled range doubled (about its middle, inside the type's
            // images), or the type's whole range where that is no wider:
            // values a little beyond the sample's extremes still fit
            const int W = tag_data_bits(f.tag);
            const bool sgn = tag_signed(f.tag);
            const uint64_t tmin = W >= 64 || !sgn ? 0 : (1ull << 63) - (1ull << (W - 1));
            const uint64_t tmax = W >= 64 ? ~0ull : (sgn ? (1ull << 63) + (1ull << (W - 1)) - 1 : (1ull << W) - 1);
            if (w + 1 >= W) {
                w = W;
                bias = tmin;
            } else {
                w += 1;
                const uint64_t span = (1ull << w) - 1;  // the field covers [bias, bias + span]
                bias = lo - std::min((span - (hi - lo)) / 2, lo - tmin);
                if (bias > tmax - span) bias = tmax - span;
            }
        }
        f.xform = tag_signed(f.tag) ? KX_SIGNED : KX_UNSIGNED;
        f.bias = bias;
        f.bits = w;
        f.null_off = nul ? off++ : -1;
        f.data_off = w ? off : 0;
        off += w;
        keys[k] = f;
    }
    // the margins do not fit: ranges from every row instead
    if (off > 64) return sampled ? narrow_lane(q, qn, false) : false;
    qn = q;
    for (int k = 0; k < nk; ++k) qn.keys[k] = keys[k];
    qn.total_bits = off;
    qn.key_mode = KM_LANE64;
    qn.allones_possible = off == 64;
    qn.narrow_check = sampled ? 1 : 0;
    return true;
}

void alloc_spill(const std::string &tag, Spill &S, uint64_t cap, bool wide, const int32_t *vw, int nv) {
    S.klo = (uint64_t *)slab(tag + ":klo", cap * 8 + 64);
    S.khi = wide ? (uint64_t *)slab(tag + ":khi", cap * 8 + 64) : nullptr;
    for (int j = 0; j < nv; ++j) S.val[j] = slab(tag + ":v" + std::to_string(j), cap * vw[j] + 64);
    S.flags = (uint16_t *)slab(tag + ":fl", cap * 2 + 64);
}

}  // namespace

pmh_result *run_hash_partitioned(Prep &P, const GroupByPlan &plan, uint32_t flags) {
    Runtime &R = rt();
    const FallbackScratchGuard fb_guard(R.scratch_mode);
    const uint64_t n = P.n_rows;
    const Options &opt0 = plan.opts;
    // a wide lane whose key columns' ranges fit 64 bits runs narrow here
    // (the other engines, and this one's fallbacks, keep P.q's lane)
    const int cus = R.prop.multiProcessorCount > 0 ? R.prop.multiProcessorCount : 256;
    // the specialised scan reads values as aligned 8-byte words, which needs
    // every column naturally aligned (Arrow buffers and numpy arrays are)
    bool aligned = true;
    for (int c = 0; c < P.q.n_cols; ++c) {
        const int w = tag_size(P.q.cols[c].tag);
        if (w > 1 && ((uintptr_t)P.q.cols[c].data % (uintptr_t)w) != 0) aligned = false;
    }
    // the specialised scan runs for the narrowed lane (the key mode aside,
    // the same conditions as below): its ranges may then come from a sample,
    // which the scan checks
    FastScan fs0{};
    const bool fast_narrow = opt0.fast_scan != 0 && aligned && P.q.n_vals <= 2 && P.q.n_keys <= 4 &&
                             ceil_div(n, cus) < (1ull << 25) && lower_fast_scan(P.q, fs0);
    QueryDesc qn;
    bool narrowed = false;
    if (P.wide && n >= kNarrowMinRows && !R.scratch_mode && opt0.narrow_keys != 0)
        narrowed = narrow_lane(P.q, qn, fast_narrow && opt0.narrow_keys == 1 && !P.narrow_exact);
    const QueryDesc &q0 = narrowed ? qn : P.q;
    const bool wide = narrowed ? false : P.wide;
    const int nv = q0.n_vals;
    const Options &opt = plan.opts;
    const size_t lds_max = std::min<size_t>(R.prop.sharedMemPerBlock ? R.prop.sharedMemPerBlock : 65536, 160 * 1024);
    const size_t lds_budget = lds_max - 1024;

    auto Lh = std::make_unique<HLaunch>();
    memset(Lh.get(), 0, sizeof(HLaunch));
    HLaunch &L = *Lh;
    L.q = q0;
    L.wide = wide;
    L.n_out = P.n_out;
    if (opt.fast_scan != 0 && aligned && !wide && nv <= 2 && q0.n_keys <= 4 && ceil_div(n, cus) < (1ull << 25) &&
        (q0.key_mode == KM_RAW64 || q0.key_mode == KM_LANE64 || q0.key_mode == KM_NONE) && lower_fast_scan(q0, L.fs))
        L.fs.enabled = 1;
    else
        L.fs = FastScan{};
    if (q0.narrow_check && !L.fs.enabled) {
        // sampled ranges that no scan would check: narrow exactly instead
        P.narrow_miss = true;
        return nullptr;
    }
    L.fs.rows = opt.scan_rows == 8 ? 8 : 4;
    size_t rec_bytes = 8 + (wide ? 8 : 0) + 2 + 1;
    for (int j = 0; j < nv; ++j) {
        L.vw[j] = tag_size(q0.vals[j].tag);
        L.col_val[q0.vals[j].col] = j;
        rec_bytes += L.vw[j];
    }
    // shapes that do not fit the LDS plans go to the global-table engine
    // the stable scatter stages a tile and keeps per-wave digit counts
    if ((size_t)kPartTile * rec_bytes + (size_t)kRadix * 4 * 4 + 128 > lds_budget) return nullptr;
    size_t state_per_slot = 0;
    for (int a = 0; a < q0.layout.n_arrs; ++a) state_per_slot += q0.layout.esz[a];
    const size_t key_bytes = wide ? 16 : 8;

    // ---- parameters
    // scan blocks: scan_blocks resident per CU, each with its share of LDS
    const uint32_t per_cu = (uint32_t)std::min<int64_t>(4, std::max<int64_t>(1, opt.scan_blocks));
    uint32_t nb = (uint32_t)std::min<uint64_t>((uint64_t)cus * per_cu, std::max<uint64_t>(1, ceil_div(n, 16384)));
    // blocks start on 64-row boundaries: no cache line or validity word of
    // a column is split between two blocks
    const uint64_t chunk = std::max<uint64_t>(64, ceil_div(ceil_div(n, nb), 64) * 64);
    nb = (uint32_t)std::max<uint64_t>(1, ceil_div(n, chunk));
    // the specialised scan keeps a row ring per wave next to the hot table
    // and has only 8 bytes of static LDS
    const size_t ring_bytes = (size_t)16 * kRing * 4;
    const size_t share = lds_max / per_cu;
    const size_t hot_budget = L.fs.enabled ? share - 64 - ring_bytes : share - 1024;
    uint32_t CL = (uint32_t)pow2_floor(hot_budget / (key_bytes + state_per_slot));
    CL = std::min<uint32_t>(CL, 2048);
    if (opt.lds_slots >= 0) CL = opt.lds_slots == 0 ? 0 : (uint32_t)std::min<uint64_t>(CL, pow2_floor(opt.lds_slots));
    if (CL < 16) CL = 0;
    // a 64K-row sample sketch (the sample is capped so its distinct keys fit
    // the fingerprint table)
    const uint32_t n_sample = (uint32_t)(opt.sample >= 0 ? std::min<int64_t>(opt.sample, 1 << 16)
                                                         : (int64_t)std::min<uint64_t>(n, 1 << 16));
    const uint32_t sample_slots = wide ? 8192 : 16384;
    // hs_sample_count lays its pick list (8 B per hot slot) and the picked
    // keys (16 B per hot slot) into its two 4 B-per-slot sample arrays
    if ((size_t)CL * 16 > (size_t)sample_slots * 4)
        fail(PMH_E_RUNTIME, "polars_metal: hot table larger than the sample table holds");
    const size_t tail_slot_bytes = key_bytes + (wide ? 4 : 0) + state_per_slot;
    uint32_t TS = (uint32_t)std::min<uint64_t>(4096, pow2_floor(lds_budget / tail_slot_bytes));
    while (TS > 16 && (size_t)(TS + 2) * tail_slot_bytes > lds_budget) TS /= 2;
    if (opt.tail_slots > 0) TS = (uint32_t)std::min<uint64_t>(TS, pow2_floor(opt.tail_slots));
    if (TS < 16) return nullptr;
    const uint32_t max_tiles = (uint32_t)(ceil_div(n, kPartTile) + nb + 1);

    L.n_sample = n_sample;
    L.hot_slots = CL;
    L.hot_max = CL ? (CL / 4 - 1) * 4 * 85 / 100 : 0;  // hot keys placed in the buckets' slots
    L.sample_slots = sample_slots;
    L.chunk = chunk;
    L.nb = nb;
    L.max_tiles = max_tiles;
    L.tail_slots = TS;

    // ---- buffers
    L.samp_lo = (uint64_t *)slab("hs:samp_lo", (uint64_t)n_sample * 8 + 64);
    L.samp_hi = (uint64_t *)slab("hs:samp_hi", (uint64_t)n_sample * 8 + 64);
    L.samp_sp = (uint8_t *)slab("hs:samp_sp", (uint64_t)n_sample + 64);
    L.skey_lo = (uint64_t *)slab("hs:skey_lo", (uint64_t)sample_slots * 8 + 64);
    L.skey_hi = (uint64_t *)slab("hs:skey_hi", (uint64_t)sample_slots * 8 + 64);
    L.hot_lo = (uint64_t *)slab("hs:hot_lo", (uint64_t)CL * 8 + 64);
    L.hot_hi = (uint64_t *)slab("hs:hot_hi", (uint64_t)CL * 8 + 64);
    {
        uint8_t *hb = (uint8_t *)slab("hs:hot_state", state_bytes(P, std::max<uint32_t>(CL, 1)) + 4096);
        size_t o = 0;
        carve_state(P, hb, o, CL, L.hot_arr);
    }
    size_t lds1 = (size_t)CL * key_bytes + lds_layout(P, CL, L.lds1_off);
    const size_t lds3 = (size_t)(TS + 2) * (key_bytes + (wide ? 4 : 0)) + lds_layout(P, TS + 2, L.lds3_off);
    if (L.fs.enabled) {
        L.fs.ring_off = (uint32_t)((lds1 + 15) & ~size_t(15));
        lds1 = L.fs.ring_off + ring_bytes;
        if (lds1 > hot_budget + ring_bytes) return nullptr;
    } else if (lds1 > lds_budget) {
        return nullptr;
    }
    if (lds3 > lds_budget) return nullptr;
    L.tail_n = (uint32_t *)slab("hs:tail_n", (nb + 1) * 4 + 64);
    L.region_off = (uint32_t *)slab("hs:region_off", (nb + 1) * 4 + 64);
    L.tile_off = (uint32_t *)slab("hs:tile_off", (nb + 1) * 4 + 64);
    const uint64_t cap = (uint64_t)nb * chunk;
    alloc_spill("hs:spill", L.spill, cap, wide, L.vw, nv);
    alloc_spill("hs:ping", L.ping, cap, wide, L.vw, nv);
    alloc_spill("hs:pong", L.pong, cap, wide, L.vw, nv);
    L.tile_hist = (uint32_t *)slab("hs:tile_hist", (uint64_t)std::max(kRadix, kWideRadix) * max_tiles * 4 + 64);
    L.digit_base = (uint32_t *)slab("hs:digit_base", std::max(kRadix, kWideRadix) * 4 + 64);
    L.part_start = (uint64_t *)slab("hs:part_start", ((1ull << kMaxPartBits) + 2) * 8 + 64);
    const uint32_t pgrid = (uint32_t)((cus * 2 + 7) / 8 * 8);  // partition kernels: a multiple of 8 (XCD slots)
    L.tile_reg = (uint32_t *)slab("hs:tile_reg", (uint64_t)max_tiles * 4 + 64);
    L.xtot = (uint32_t *)slab("hs:xtot", 8 * kWideRadix * 4 + 64);
    L.xcur = (uint64_t *)slab("hs:xcur", 8 * kWideRadix * 8 + 64);
    L.out_cursor = (uint32_t *)slab("hs:cursor", 64);
    // stats and the sketch histogram are adjacent: one read-back
    L.stats = (uint32_t *)slab("hs:stats", (ST_N + SZ_N + 1024) * 4 + 64);
    L.sketch_hist = L.stats + ST_N + SZ_N;
    // the learned sizer runs inside the sketch block when "auto" will use it
    const bool use_sizer = opt.method == "auto" && opt.distinct_hint <= 0 && P.d_retry <= 0 && !R.scratch_mode;
    L.sizer = sizer_installed() && (use_sizer || opt.sketch_only) ? sizer_device_params() : nullptr;
    L.lane_bits = P.q.total_bits;  // the sizer's features use the plan's lane width
    // fallback table
    {
        GTable &t = L.t;
        const uint64_t fcap = pow2_ceil(std::min<uint64_t>(std::max<uint64_t>(n / 64, 4096), 1ull << 21));
        const uint64_t nslots = fcap + 2;
        t.mask = fcap - 1;
        t.limit = (uint64_t)(fcap * 0.7);
        t.list_cap = nslots;
        size_t bytes = nslots * 8 + (wide ? nslots * 12 : 0) + nslots * 4 + 64 + state_bytes(P, nslots) + 8192;
        uint8_t *base = (uint8_t *)slab("hs:fb_table", bytes);
        size_t o = 0;
        auto carve = [&](size_t b) { uint8_t *p = base + o; o += (b + 255) & ~size_t(255); return p; };
        t.klo = (uint64_t *)carve(nslots * 8);
        if (wide) { t.khi = (uint64_t *)carve(nslots * 8); t.kstate = (uint32_t *)carve(nslots * 4); }
        t.list = (uint32_t *)carve(nslots * 4);
        t.ctrl = (uint32_t *)carve(64);
        carve_state(P, base, o, nslots, t.arr);
        const std::string fsig = P.sig + (wide ? "w" : "n");
        if (g_fb.cap != fcap || g_fb.sig != fsig) g_fb.clean = false;
        g_fb.cap = fcap;
        g_fb.sig = fsig;
    }
    // outputs: at most n groups
    const uint64_t maxg = std::max<uint64_t>(n, 1);
    for (int oi = 0; oi < P.n_out; ++oi) {
        OutCol &oc = L.out[oi];
        oc = P.outs[oi];
        const size_t esz = std::max(1, tag_size(oc.tag));
        oc.data = slab("hs:out" + std::to_string(oi), maxg * esz + 64);
        const bool need_valid = !(oc.src == OS_AGG && (oc.op == AGG_LEN || oc.op == AGG_COUNT));
        oc.vbytes = need_valid ? (uint8_t *)slab("hs:outvb" + std::to_string(oi), maxg + 64) : nullptr;
        oc.valid = need_valid ? (uint64_t *)slab("hs:outv" + std::to_string(oi), maxg / 8 + 64) : nullptr;
        oc.data_bits = (oc.src == OS_KEY && oc.tag == PMH_BOOL)
                           ? (uint8_t *)slab("hs:outbits" + std::to_string(oi), maxg / 8 + 64)
                           : nullptr;
    }

    // ---- launches
    HLaunch *dL = (HLaunch *)slab("hs:launch", sizeof(HLaunch));
    // Control traffic goes through pinned host memory: the descriptor
    // uploads from a staged copy, and read-backs land in a mailbox, so each
    // copy is a plain DMA on the stream instead of a staged, waited-for copy
    // of pageable memory.  The query ends synchronised, so a later query
    // never overwrites a staged copy still being read.
    const size_t up_bytes = (sizeof(HLaunch) + 255) & ~size_t(255);
    uint8_t *pin = (uint8_t *)pinned_host(up_bytes + (ST_N + SZ_N + 1024 + 64) * 4);
    HLaunch *up = (HLaunch *)pin;
    uint32_t *mbox = (uint32_t *)(pin + up_bytes);
    {
        void *dmbox = nullptr;
        PMH_HIP(hipHostGetDevicePointer(&dmbox, mbox, 0));
        L.mbox = (uint32_t *)dmbox;
    }
    L.back_words = (int32_t)(ST_N + SZ_N + (opt.sketch_only ? 1024 : 0));
    memcpy(up, &L, sizeof(HLaunch));
    PMH_HIP(hipMemcpyAsync(dL, up, sizeof(HLaunch), hipMemcpyHostToDevice, R.stream));
    PooledEvent pe0, pe1;
    const hipEvent_t ev0 = pe0.ev, ev1 = pe1.ev;
    PMH_HIP(hipEventRecord(ev0, R.stream));
    auto launch_check = [](void) { PMH_HIP(hipGetLastError()); };
    hipLaunchKernelGGL(hs_reset, dim3(1), dim3(64), 0, R.stream, dL);
    launch_check();
    if (!g_fb.clean) {
        KernelTimer kt("hs_init_table");
        const uint64_t nslots = L.t.mask + 3;
        hipLaunchKernelGGL(hs_init_table, dim3((uint32_t)std::min<uint64_t>(ceil_div(nslots, 256), (uint64_t)cus * 16)),
                           dim3(256), 0, R.stream, dL, nslots);
        launch_check();
        g_fb.clean = true;
    }
    if (n_sample > 0) {
        KernelTimer kt("hs_sample_gather");
        hipLaunchKernelGGL(hs_sample_gather, dim3((uint32_t)ceil_div(n_sample, 256)), dim3(256), 0, R.stream, dL);
        launch_check();
    }
    {
        KernelTimer kt("hs_sample_count");
        const size_t lds = (size_t)sample_slots * 8 + 4096 + (size_t)CL * key_bytes;
        PMH_HIP(hipFuncSetAttribute((const void *)hs_sample_count, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)lds));
        hipLaunchKernelGGL(hs_sample_count, dim3(1), dim3(1024), lds, R.stream, dL);
        launch_check();
    }
    // ---- partition count from the sketch (one small read-back)
    // The scan reads nothing the planning below sets (partition bits,
    // passes, per-partition table), so it is enqueued behind the sketch's
    // copy and the host plans while it runs; the planned descriptor uploads
    // behind it in stream order.  A query the learned sizer then routes to
    // the sort engine has run one scan for nothing (its results are unused).
    uint32_t sk[ST_N];
    PooledEvent pe_sketch;
    const hipEvent_t ev_sketch = pe_sketch.ev;
    PMH_HIP(hipEventRecord(ev_sketch, R.stream));
    auto launch_scan = [&]() {
        const bool cached = q0.n_cols <= 4;
        const size_t l1 = std::max<size_t>(lds1, 16);
        if (L.fs.enabled) {
            KernelTimer kt("hs_scan_fast");
            const int km = q0.key_mode == KM_RAW64 ? kScanRaw64 : q0.narrow_check ? kScanLaneChecked : kScanLane;
            const bool r8 = L.fs.rows == 8;
            // four rows per lane: one instantiation per key kind, aggregate
            // signature and predicate width; eight: the fold resolved at run time
#define PMH_SCAN4(K, N)                                                                                 \
    {(const void *)hs_scan_fast<K, 4, 0, N>, (const void *)hs_scan_fast<K, 4, 1, N>,                     \
     (const void *)hs_scan_fast<K, 4, 2, N>, (const void *)hs_scan_fast<K, 4, 3, N>,                     \
     (const void *)hs_scan_fast<K, 4, 4, N>}
            static const void *const kScan4[3][2][kNumAggSigs] = {
                {PMH_SCAN4(kScanLane, false), PMH_SCAN4(kScanLane, true)},
                {PMH_SCAN4(kScanRaw64, false), PMH_SCAN4(kScanRaw64, true)},
                {PMH_SCAN4(kScanLaneChecked, false), PMH_SCAN4(kScanLaneChecked, true)}};
#undef PMH_SCAN4
            static const void *const kScan8[3] = {(const void *)hs_scan_fast<kScanLane, 8, 0, true>,
                                                  (const void *)hs_scan_fast<kScanRaw64, 8, 0, true>,
                                                  (const void *)hs_scan_fast<kScanLaneChecked, 8, 0, true>};
            static_assert(kNumAggSigs == 5, "one instantiation per signature");
            const int sig = agg_signature(q0);  // the fast scan takes at most two value columns
            const bool nr2 = L.fs.nr > 1;
            const void *fn = r8 ? kScan8[km] : kScan4[km][nr2 ? 1 : 0][sig];
            PMH_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)l1));
            void *args[] = {(void *)&dL};
            PMH_HIP(hipLaunchKernel(fn, dim3(nb), dim3(1024), args, l1, R.stream));
        } else {
            KernelTimer kt("hs_scan");
            const void *fn = cached ? (const void *)hs_scan<true> : (const void *)hs_scan<false>;
            PMH_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)l1));
            if (cached) hipLaunchKernelGGL(hs_scan<true>, dim3(nb), dim3(1024), l1, R.stream, dL);
            else hipLaunchKernelGGL(hs_scan<false>, dim3(nb), dim3(1024), l1, R.stream, dL);
        }
        launch_check();
    };
    if (!opt.sketch_only) launch_scan();
    {
        PMH_HIP(hipEventSynchronize(ev_sketch));
    }
    memcpy(sk, mbox, sizeof(sk));
    const uint32_t *blk = mbox + ST_N;
    const double tail_cap = 0.5 * TS;  // target keys per partition
    double d_est;
    memcpy(&d_est, blk + SZ_EST, 8);
    memcpy(P.feat, blk + SZ_FEAT, sizeof(P.feat));
    float pred[2] = {NAN, NAN};
    if (blk[SZ_HAVE]) memcpy(pred, blk + SZ_PRED, sizeof(pred));
    if (opt.sketch_only) {
        if (pmh_sketch_info *o = P.sketch_out) {
            o->sample_size = n_sample;
            o->sample_rows = sk[ST_SAMPLE_SEL];
            o->distinct = sk[ST_SAMPLE_DISTINCT];
            o->f1 = sk[ST_SAMPLE_F1];
            o->f2 = sk[ST_SAMPLE_F2];
            o->dropped = sk[ST_SAMPLE_DROPPED];
            o->null_keys = sk[ST_SAMPLE_NULLKEY];
            o->hot_keys = sk[ST_HOT];
            o->lane_bits = (uint32_t)P.q.total_bits;
            o->scale = blk[SZ_SCALE];
            memcpy(o->hist, blk + SZ_N, sizeof(o->hist));
            memcpy(o->features, P.feat, sizeof(o->features));
            memcpy(o->prediction, pred, sizeof(o->prediction));
            o->distinct_estimate = d_est;
        }
        collect_kernel_times();
        pmh_result *res = new_result(0);
        res->stats.method = PMH_METHOD_HASH;
        res->stats.sample_rows = sk[ST_SAMPLE_SEL];
        res->stats.sample_distinct = sk[ST_SAMPLE_DISTINCT];
        res->stats.sample_f1 = sk[ST_SAMPLE_F1];
        res->stats.sample_f2 = sk[ST_SAMPLE_F2];
        res->stats.hot_keys = sk[ST_HOT];
        res->stats.distinct_estimate = d_est;
        return res;
    }
    if (L.sizer) {
        // the learned sizer (evaluated in the sketch block) picks the engine
        // and, for hash, the cardinality the partitions are sized for;
        // overflow still falls back exactly
        if (pred[1] > 0.f) {
            P.route = 1;
            P.sizer_choice = 2;
            collect_kernel_times();
            return nullptr;
        }
        P.sizer_choice = 1;
        const double g = std::exp2(std::min(40.0, (double)pred[0] * 32.0)) - 1.0;
        d_est = std::max<double>(g, sk[ST_SAMPLE_DISTINCT]);
    }
    if (opt.distinct_hint > 0) d_est = (double)opt.distinct_hint;
    if (P.d_retry > 0) d_est = P.d_retry;
    const double d_tail = std::max(0.0, d_est - sk[ST_HOT]);
    // a sample of near-unique keys (nine in ten seen once) estimates the
    // distinct count from a handful of repeats, and the power-of-two
    // partition count doubles on a 40% change of it: size the partitions for
    // half again as many keys, since too few partitions spill their full
    // tables into the fallback table (the slow side of the error)
    const bool near_unique = sk[ST_SAMPLE_SEL] && (double)sk[ST_SAMPLE_F1] >= 0.9 * (double)sk[ST_SAMPLE_SEL];
    const double d_part = near_unique && opt.distinct_hint <= 0 && P.d_retry <= 0 ? 1.5 * d_tail : d_tail;
    uint64_t P_parts;
    if (opt.partitions > 0) {
        P_parts = pow2_floor((uint64_t)opt.partitions);
    } else {
        // an estimate beyond what the largest partition count holds still
        // runs here: estimates of near-unique samples overshoot, and keys
        // that miss a full partition table go to the fallback table; only
        // its overflow sends the query to the global-table engine
        P_parts = pow2_ceil((uint64_t)std::max(1.0, std::ceil(d_part / tail_cap)));
        // and enough partitions for the aggregation to spread over the
        // device: about 8K tail rows (selected rows outside the hot keys,
        // projected from the sample) per partition, up to two radix passes
        const double hot_frac =
            sk[ST_SAMPLE_SEL] ? std::min(1.0, (double)sk[ST_HOT_SAMPLE_ROWS] / sk[ST_SAMPLE_SEL]) : 0.0;
        const double n_sel = n_sample ? (double)n * sk[ST_SAMPLE_SEL] / n_sample : 0.0;
        const double tail_rows = n_sel * (1.0 - hot_frac);
        P_parts = std::max<uint64_t>(P_parts, std::min<uint64_t>(1024, pow2_floor((uint64_t)(tail_rows / 8192) + 1)));
    }
    P_parts = std::min<uint64_t>(P_parts, 1u << kMaxPartBits);
    const int pbits = log2u(P_parts);
    // up to 1024 partitions: one unstable pass on the partition digit; more:
    // an unstable pass on the low pbits - 7 bits, then one stable 7-bit pass
    // (LSD order needs stability from the second pass on), when the unstable
    // pass's LDS staging (16-bit digits) fits
    const size_t stage_lds1 = (size_t)kPartTile * (rec_bytes + 1);
    const bool wide_fits = stage_lds1 + 5 * kWideRadix * 4 <= lds_budget;
    const bool one_pass = pbits <= kWideRadixBits && wide_fits;
    const bool two_pass = !one_pass && wide_fits && pbits - kRadixBits <= kWideRadixBits;
    const int passes = one_pass ? 1 : two_pass ? 2 : std::max(1, (pbits + kRadixBits - 1) / kRadixBits);
    const int sortbits = (one_pass || two_pass) ? pbits : passes * kRadixBits;
    L.pbits = (uint32_t)pbits;
    L.sortbits = (uint32_t)sortbits;
    L.passes = (uint32_t)passes;
    // per-partition table: about twice the expected keys per partition, so
    // the per-partition LDS reset stays proportional to the work
    if (opt.tail_slots <= 0) {
        const double per = d_part / (double)P_parts;
        uint32_t ts = (uint32_t)std::min<uint64_t>(TS, std::max<uint64_t>(256, pow2_ceil((uint64_t)(2.5 * per) + 1)));
        L.tail_slots = ts;
    }
    size_t lds3b = (size_t)(L.tail_slots + 2) * (key_bytes + (wide ? 4 : 0)) +
                   lds_layout(P, L.tail_slots + 2, L.lds3_off);
    {
        // the emit's dense list of live slots (2 bytes a slot), where it fits
        // and costs no block per CU
        const size_t off = (lds3b + 15) & ~size_t(15), with = off + (size_t)(L.tail_slots + 2) * 2;
        auto per_cu = [&](size_t b) { return std::min<size_t>(2, lds_max / b); };
        L.live_off = with <= lds_budget && per_cu(with) >= per_cu(lds3b) ? (uint32_t)off : 0;
        if (L.live_off) lds3b = with;
    }
    if (two_pass) {
        // the second pass's counts, cursors and bucket-aligned tiles
        const uint32_t nd = 1u << (pbits - kRadixBits);
        const uint64_t P2 = (uint64_t)nd * kRadix;
        const uint64_t max_t2 = ceil_div(n, kPartTile) + nd + 1;
        uint8_t *hb = (uint8_t *)slab("hs:pass2", P2 * 8 * 12 + (nd + 1) * 12 + max_t2 * 4 + P2 / 256 + 2048);
        size_t o = 0;
        auto carve2 = [&](size_t bytes) {
            void *c = hb + o;
            o += (bytes + 255) & ~size_t(255);
            return c;
        };
        L.xtot2 = (uint32_t *)carve2(P2 * 8 * 4);
        L.xcur2 = (uint64_t *)carve2(P2 * 8 * 8);
        L.bstart2 = (uint64_t *)carve2((size_t)(nd + 1) * 8);
        L.tile_off2 = (uint32_t *)carve2((size_t)(nd + 1) * 4);
        L.tile_reg2 = (uint32_t *)carve2(max_t2 * 4);
        L.bsum2 = (uint32_t *)carve2((size_t)ceil_div(P2, 1024) * 4);
    }
    memcpy(up, &L, sizeof(HLaunch));
    PMH_HIP(hipMemcpyAsync(dL, up, sizeof(HLaunch), hipMemcpyHostToDevice, R.stream));
    hipLaunchKernelGGL(hs_regions, dim3(1), dim3(1024), 0, R.stream, dL);
    launch_check();

    const size_t stage_lds = (size_t)kPartTile * rec_bytes;
    PMH_HIP(hipFuncSetAttribute((const void *)rp_scatter, hipFuncAttributeMaxDynamicSharedMemorySize, (int)stage_lds));
    const Spill *bufs[3] = {&dL->spill, &dL->ping, &dL->pong};
    int src = 0, dst = 1;
    if (one_pass || two_pass) {
        const uint32_t nd = 1u << (one_pass ? pbits : pbits - kRadixBits);
        PMH_HIP(hipFuncSetAttribute((const void *)rp_scatter1, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)stage_lds1));
        {
            KernelTimer kt("rp_hist");
            hipLaunchKernelGGL(rp_hist1, dim3(pgrid), dim3(kPartThreads), 0, R.stream, dL, bufs[0], nd);
            launch_check();
        }
        {
            KernelTimer kt("rp_scan");
            hipLaunchKernelGGL(rp_scan1, dim3(1), dim3(kPartThreads), 0, R.stream, dL, nd);
            launch_check();
        }
        {
            KernelTimer kt("rp_scatter");
            hipLaunchKernelGGL(rp_scatter1, dim3(pgrid), dim3(kPartThreads), stage_lds1, R.stream, dL, bufs[0], bufs[1],
                               nd);
            launch_check();
        }
        src = 1;
        dst = 2;
    }
    if (two_pass) {
        // MSD inside the first pass's buckets (nd of them)
        const uint32_t nd = 1u << (pbits - kRadixBits);
        const uint32_t P2 = nd * kRadix;
        const int shift = pbits - kRadixBits;
        hipLaunchKernelGGL(rp_tiles2, dim3((uint32_t)std::min<uint64_t>(ceil_div((uint64_t)P2 * 8, 1024), cus * 4)),
                           dim3(1024), 0, R.stream, dL, nd);
        launch_check();
        {
            KernelTimer kt("rp_hist");
            hipLaunchKernelGGL(rp_hist2, dim3(pgrid), dim3(kPartThreads), 0, R.stream, dL, bufs[src], nd, shift);
            launch_check();
        }
        {
            KernelTimer kt("rp_scan");
            const uint32_t sb = (uint32_t)ceil_div(P2, 1024);
            hipLaunchKernelGGL(rp_sum2, dim3(sb), dim3(1024), 0, R.stream, dL, P2);
            launch_check();
            hipLaunchKernelGGL(rp_scan2, dim3(sb), dim3(1024), 0, R.stream, dL, P2);
            launch_check();
        }
        {
            KernelTimer kt("rp_scatter");
            PMH_HIP(hipFuncSetAttribute((const void *)rp_scatter2, hipFuncAttributeMaxDynamicSharedMemorySize,
                                        (int)stage_lds));
            hipLaunchKernelGGL(rp_scatter2, dim3(pgrid), dim3(kPartThreads), stage_lds, R.stream, dL, bufs[src],
                               bufs[dst], nd, shift);
            launch_check();
        }
        src = dst;
    }
    for (int ps = 0; ps < passes && !one_pass && !two_pass; ++ps) {
        const int shift = ps * kRadixBits;
        {
            KernelTimer kt("rp_hist");
            hipLaunchKernelGGL(rp_hist, dim3(pgrid), dim3(kPartThreads), 0, R.stream, dL, ps == 0 ? 1 : 0, bufs[src],
                               shift);
            launch_check();
        }
        {
            KernelTimer kt("rp_scan");
            hipLaunchKernelGGL(rp_scan, dim3(kRadix), dim3(1024), 0, R.stream, dL, ps == 0 ? 1 : 0);
            launch_check();
        }
        {
            KernelTimer kt("rp_scatter");
            hipLaunchKernelGGL(rp_scatter, dim3(pgrid), dim3(kPartThreads), stage_lds, R.stream, dL, ps == 0 ? 1 : 0,
                               bufs[src], bufs[dst], shift);
            launch_check();
        }
        src = dst;
        dst = dst == 1 ? 2 : 1;
    }
    if (!one_pass && !two_pass) {
        KernelTimer kt("hs_part_bounds");
        hipLaunchKernelGGL(hs_part_bounds, dim3((uint32_t)ceil_div(P_parts + 1, 256)), dim3(256), 0, R.stream, dL,
                           bufs[src]);
        launch_check();
    }
    {
        KernelTimer kt("hs_part_agg");
        const int sig = q0.n_vals <= 2 ? agg_signature(q0) : -1;
        static const void *const kPartAgg[2][kNumAggSigs + 1] = {
            {(const void *)hs_part_agg<false, -1>, (const void *)hs_part_agg<false, 0>,
             (const void *)hs_part_agg<false, 1>, (const void *)hs_part_agg<false, 2>,
             (const void *)hs_part_agg<false, 3>, (const void *)hs_part_agg<false, 4>},
            {(const void *)hs_part_agg<true, -1>, (const void *)hs_part_agg<true, 0>, (const void *)hs_part_agg<true, 1>,
             (const void *)hs_part_agg<true, 2>, (const void *)hs_part_agg<true, 3>, (const void *)hs_part_agg<true, 4>}};
        static_assert(kNumAggSigs == 5, "one instantiation per signature");
        const void *fn = kPartAgg[wide ? 1 : 0][sig + 1];
        PMH_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds3b));
        const uint32_t g = (uint32_t)std::min<uint64_t>(P_parts, (uint64_t)cus * 2);
        const Spill *sp = bufs[src];
        void *args[] = {(void *)&dL, (void *)&sp};
        PMH_HIP(hipLaunchKernel(fn, dim3(g), dim3(1024), args, lds3b, R.stream));
        launch_check();
    }
    if (CL) {
        KernelTimer kt("hs_hot_emit");
        hipLaunchKernelGGL(hs_hot_emit, dim3((uint32_t)ceil_div(CL, 256)), dim3(256), 0, R.stream, dL);
        launch_check();
    }
    {
        KernelTimer kt("hs_fallback_emit");
        hipLaunchKernelGGL(hs_fallback_emit, dim3(cus), dim3(256), 0, R.stream, dL);
        launch_check();
    }
    {
        KernelTimer kt("hs_pack");
        hipLaunchKernelGGL(hs_pack, dim3(cus * 4), dim3(256), 0, R.stream, dL);
        launch_check();
    }
    uint32_t host_ctrl[8], host_stats[ST_N], cursor = 0;
    hipLaunchKernelGGL(hs_clear_ctrl, dim3(1), dim3(64), 0, R.stream, dL);
    launch_check();
    PMH_HIP(hipEventRecord(ev1, R.stream));
    PMH_HIP(hipEventSynchronize(ev1));
    memcpy(host_ctrl, mbox, sizeof(host_ctrl));
    memcpy(host_stats, mbox + 8, sizeof(host_stats));
    cursor = mbox[8 + ST_N];
    float dev_ms = 0.f;
    (void)hipEventElapsedTime(&dev_ms, ev0, ev1);
    if (host_stats[ST_ERROR]) {
        g_fb.clean = false;
        fail(PMH_E_COMPUTE, "polars_metal: partitioned group-by detected an inconsistent table state");
    }
    if (host_stats[ST_NARROW_MISS]) {
        // a selected key outside the sampled narrowing ranges: the caller
        // re-runs the query narrowed from every row
        g_fb.clean = false;
        P.narrow_miss = true;
        collect_kernel_times();
        return nullptr;
    }
    if (host_ctrl[1]) {
        // the fallback table overflowed and is dirty: the caller re-runs the
        // query with a larger estimate, or on the global-table engine
        g_fb.clean = false;
        P.overflowed = true;
        P.d_used = d_est;
        collect_kernel_times();
        return nullptr;
    }
    const uint64_t ng = (uint64_t)cursor + std::min<uint64_t>(host_ctrl[0], L.t.list_cap);

    pmh_result *res = new_result(P.n_out);
    res->n_rows = ng;
    res->stats.method = PMH_METHOD_HASH;
    res->stats.n_passes = 1;
    res->stats.initial_capacity = L.t.mask + 1;
    res->stats.final_capacity = L.t.mask + 1;
    res->stats.n_groups = ng;
    res->stats.n_selected = host_stats[ST_SELECTED];
    res->stats.device_ms = dev_ms;
    res->stats.sample_rows = host_stats[ST_SAMPLE_SEL];
    res->stats.sample_distinct = host_stats[ST_SAMPLE_DISTINCT];
    res->stats.sample_f1 = host_stats[ST_SAMPLE_F1];
    res->stats.sample_f2 = host_stats[ST_SAMPLE_F2];
    res->stats.hot_keys = host_stats[ST_HOT];
    res->stats.partitions = (uint32_t)P_parts;
    res->stats.tail_rows = host_stats[ST_TAIL];
    res->stats.distinct_estimate = d_est;
    res->stats.scan_kernel = L.fs.enabled;
    res->stats.sizer_choice = P.sizer_choice;
    const uint64_t vbytes = bitmap_bytes_padded(ng);
    const uint64_t packed = (ng + 7) / 8;
    for (int oi = 0; oi < P.n_out; ++oi) {
        const OutCol &oc = L.out[oi];
        pmh_out_column &c = res->columns[oi];
        set_out_name(c, P.names[oi]);
        c.tag = oc.tag;
        const bool is_bool = oc.src == OS_KEY && oc.tag == PMH_BOOL;
        c.data_len = is_bool ? vbytes : ng * tag_size(oc.tag);
        c.valid_len = vbytes;
        const void *dsrc = is_bool ? (const void *)oc.data_bits : oc.data;
        if (flags & PMH_OUT_DEVICE) {
            c.location = PMH_DEVICE;
            c.data = (void *)dsrc;
            c.valid = (uint8_t *)oc.valid;
        } else {
            c.location = PMH_HOST;
            c.data = malloc(c.data_len ? c.data_len : 1);
            c.valid = (uint8_t *)malloc(vbytes);
            memset(c.valid, 0, vbytes);
            if (is_bool) memset(c.data, 0, c.data_len);
            const uint64_t nb_copy = is_bool ? packed : c.data_len;
            if (nb_copy) download_host(c.data, dsrc, nb_copy);
            if (oc.valid) {
                if (packed) download_host(c.valid, oc.valid, packed);
            } else {
                memset(c.valid, 0xFF, vbytes);  // counts / lens are never null
            }
        }
    }
    // host outputs wait for their copies; device outputs were complete at ev1
    if (!(flags & PMH_OUT_DEVICE)) PMH_HIP(hipStreamSynchronize(R.stream));
    collect_kernel_times();
    return res;
}

}  // namespace pmh
