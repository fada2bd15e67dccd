// pmh_batch.hip -- many group-by queries of one shape in one device pass.
//
// No counterpart in the reference, which runs one query per call, each with
// its own dispatches and blocking waits (crates/polars-metal-core/src/
// command.rs:137-254).  A batch here is a list of queries that share one
// plan template (keys, aggregates, column names and dtypes) but each bring
// their own columns, row count and predicate.  Per query the path is the
// partitioned hash engine's, with every stage run for all queries at once:
//
//   bt_sketch     one block per query: a strided row sample's distinct keys
//                 (host: per-query distinct estimate -> partition count)
//   bt_hist       tiles of all queries: rows per (query, partition)
//   bt_scan*      exclusive scan over all (query, partition) counts
//   bt_scatter    tiles of all queries: rows written partition-contiguous
//   bt_agg        (query, partition) work items: LDS table fold, emit into
//                 the query's output rows; keys that miss a full table go
//                 to a fallback table keyed by (key, query)
//   bt_fb_emit    fallback groups out
//   bt_pack       per query, validity bytes -> bitmaps
//
// Every kernel after the sketch reads its work counts from device memory
// and walks it with a fixed grid, so the chain is captured once into a
// hipGraph and replayed for later batches (re-captured only when a buffer
// grows).  Queries whose output capacity (sized from the estimate) proves
// too small are re-run one by one on the single-query engines.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>

#include "pmh_mlp.h"
#include "pmh_table.h"
#include "pmh_words.h"

namespace pmh {

namespace {

constexpr int kBCols = 8;           // columns of the template a batch may reference
constexpr uint32_t kBMaxParts = 4096;  // partitions per query
constexpr int kBSample = 2048;      // sampled rows per query for the sketch
constexpr int kBSampleSlots = 4096; // its LDS table
constexpr uint32_t kBMaxSlots = 4096;  // LDS table slots of a (query, partition) work item, at most
constexpr int kBProbe = 32;         // LDS table probes (triangular steps) before a key goes to the fallback table
constexpr int kBKeys = 4;           // key columns of a batch whose key lane is packed (lane mode 0)

struct BQuery {
    uint64_t n_rows;
    const void *data[kBCols];
    const uint8_t *valid[kBCols];
    FastScan fs;         // predicate intervals (nr = 0: none)
    uint32_t P, pad;     // partitions
    uint64_t part_base;  // first partition in the batch arrays
    uint64_t out_base, out_cap;
    uint64_t vb_base;    // first byte of this query's validity bitmaps
    uint32_t tile_base, n_tiles;
};

enum BCtr : int32_t { BC_WORK = 0, BC_N = 8 };

struct BLaunch {
    QueryDesc q;  // the template: layout, keys, values, column tags
    uint32_t n_queries, n_tiles, tile_rows, part_slots;  // part_slots: LDS table size (power of two)
    uint64_t n_parts;
    uint32_t scat_parts, scat_pad;  // bt_tiles<scatter>: hist and base entries (the largest P, even)
    BQuery *qs;
    uint32_t *tile_q;      // [n_tiles]
    uint32_t *part_q;      // [n_parts]
    uint32_t *part_cnt;    // [n_parts]
    uint64_t *part_off;    // [n_parts + 1]
    uint64_t *part_cur;    // [n_parts]
    uint64_t *scan_blk;    // [ceil(n_parts / 4096) + 1]
    uint32_t *sk;          // [n_queries][5]: sampled selected rows, distinct, f1, f2, null keys
    uint64_t *rk;          // records: key lane
    uint64_t *rv[2];       // records: value bits
    uint16_t *rf;          // records: flags (bit j value valid; side-key bits)
    uint32_t *out_cursor;  // [n_queries]
    uint32_t *q_sel;       // [n_queries]
    uint32_t *q_fb;        // [n_queries]: 1 once a row of the query missed its partition table
    GTable t;              // fallback, keys (lane, query)
    uint64_t fb_slots;     // t.mask + 3
    OutCol out[kOutMax];   // big arrays; query q's groups at out_base
    int32_t n_out, lane_mode;  // lane_mode: 1 one 64-bit raw key, 0 a lane of at most 64 bits
    int32_t col_val[kMaxCols];
};

constexpr uint16_t kBAllOnes = 1u << 14, kBNull = 1u << 15;

// One column of one query: its data and validity pointers (validity null:
// no nulls).  Widths and tags come from the template.
struct BCol {
    const uint8_t *d, *v;
};

// The columns that one tile of a query reads, resolved once per tile into
// statically indexed locals.  (A copy of the whole BQuery would be indexed
// by column number at run time, which puts it in scratch memory.)
struct TileCols {
    uint64_t n_rows, part_base;
    uint32_t P, tile_base;
    int32_t nr, nk;
    int32_t rflt[2], rtag[2];
    int64_t rlo[2], rhi[2];
    BCol pc[2], kc[kBKeys], vc[2];
};

__device__ __forceinline__ BCol bq_col(const BQuery *__restrict__ Qp, int c) {
    return BCol{(const uint8_t *)Qp->data[c], Qp->valid[c]};
}

__device__ __forceinline__ void tile_cols(const BLaunch *__restrict__ L, uint32_t qi, TileCols &T) {
    const BQuery *__restrict__ Qp = L->qs + qi;
    T.n_rows = Qp->n_rows;
    T.part_base = Qp->part_base;
    T.P = Qp->P;
    T.tile_base = Qp->tile_base;
    T.nr = Qp->fs.nr;
#pragma unroll
    for (int i = 0; i < 2; ++i)
        if (i < T.nr) {
            const int c = Qp->fs.rcol[i];
            T.pc[i] = bq_col(Qp, c);
            T.rtag[i] = L->q.cols[c].tag;
            T.rflt[i] = Qp->fs.rflt[i];
            T.rlo[i] = Qp->fs.rlo[i];
            T.rhi[i] = Qp->fs.rhi[i];
        }
    T.nk = L->lane_mode == 1 ? 1 : L->q.n_keys;
#pragma unroll
    for (int i = 0; i < kBKeys; ++i)
        if (i < T.nk) T.kc[i] = bq_col(Qp, L->q.keys[i].col);
#pragma unroll
    for (int j = 0; j < 2; ++j)
        if (j < L->q.n_vals) T.vc[j] = bq_col(Qp, L->q.vals[j].col);
}

// Raw bits of U values of a naturally aligned column (zero-extended; Bool
// bit-packed).  The width switch is uniform and taken once for all U loads.
template <int U>
__device__ __forceinline__ void col_bits(int tag, const uint8_t *d, const uint64_t (&r)[U], uint64_t (&x)[U]) {
    if (tag == PMH_BOOL) {
#pragma unroll
        for (int u = 0; u < U; ++u) x[u] = (gld<uint8_t>(d, r[u] >> 3) >> (r[u] & 7)) & 1u;
        return;
    }
    switch (tag_size(tag)) {
    case 8:
#pragma unroll
        for (int u = 0; u < U; ++u) x[u] = gld<uint64_t>(d, r[u]);
        break;
    case 4:
#pragma unroll
        for (int u = 0; u < U; ++u) x[u] = gld<uint32_t>(d, r[u]);
        break;
    case 2:
#pragma unroll
        for (int u = 0; u < U; ++u) x[u] = gld<uint16_t>(d, r[u]);
        break;
    default:
#pragma unroll
        for (int u = 0; u < U; ++u) x[u] = gld<uint8_t>(d, r[u]);
        break;
    }
}

// Validity bits of U rows (a column without a bitmap: all 1, no loads).
template <int U>
__device__ __forceinline__ void col_valid(const uint8_t *v, const uint64_t (&r)[U], uint32_t (&vb)[U]) {
    if (!v) {
#pragma unroll
        for (int u = 0; u < U; ++u) vb[u] = 1u;
        return;
    }
#pragma unroll
    for (int u = 0; u < U; ++u) vb[u] = (gld<uint8_t>(v, r[u] >> 3) >> (r[u] & 7)) & 1u;
}

// Predicate of U rows (rows must be in range; sel[u] starts as the caller's
// in-range flag).
template <int U>
__device__ __forceinline__ void eval_pred(const TileCols &T, const uint64_t (&r)[U], bool (&sel)[U]) {
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        if (i >= T.nr) break;
        uint64_t raw[U];
        uint32_t vb[U];
        col_bits<U>(T.rtag[i], T.pc[i].d, r, raw);
        col_valid<U>(T.pc[i].v, r, vb);
        int64_t img[U];
        pred_images<U>(T.rtag[i], T.rflt[i] != 0, raw, img);
#pragma unroll
        for (int u = 0; u < U; ++u) sel[u] = sel[u] && vb[u] && img[u] >= T.rlo[i] && img[u] <= T.rhi[i];
    }
}

// Key lanes of U rows; special: 1 the all-ones lane, 2 the null key of a
// one-column 64-bit key.
template <int U>
__device__ __forceinline__ void eval_keys(const BLaunch *__restrict__ L, const TileCols &T, const uint64_t (&r)[U],
                                          uint64_t (&lane)[U], uint32_t (&special)[U]) {
    if (L->lane_mode == 1) {
        uint32_t vb[U];
#pragma unroll
        for (int u = 0; u < U; ++u) lane[u] = gld<uint64_t>(T.kc[0].d, r[u]);
        col_valid<U>(T.kc[0].v, r, vb);
#pragma unroll
        for (int u = 0; u < U; ++u) special[u] = !vb[u] ? 2u : (lane[u] == kEmpty ? 1u : 0u);
        return;
    }
#pragma unroll
    for (int u = 0; u < U; ++u) lane[u] = 0;
#pragma unroll
    for (int i = 0; i < kBKeys; ++i) {
        if (i >= T.nk) break;
        const KeyField &kf = L->q.keys[i];
        uint64_t x[U];
        uint32_t vb[U];
        col_bits<U>(L->q.cols[kf.col].tag, T.kc[i].d, r, x);
        col_valid<U>(T.kc[i].v, r, vb);
        const int doff = kf.data_off, noff = kf.null_off;
#pragma unroll
        for (int u = 0; u < U; ++u) lane[u] |= vb[u] ? x[u] << doff : 1ull << noff;
    }
    const bool ao = L->q.allones_possible;
#pragma unroll
    for (int u = 0; u < U; ++u) special[u] = (ao && lane[u] == kEmpty) ? 1u : 0u;
}

// Partition of a key: multiply-shift of the hash's high word onto [0, P)
// (P need not be a power of two); side keys live in partition 0.
__device__ __forceinline__ uint32_t bq_part(uint32_t P, uint64_t lane, uint32_t special) {
    if (special || P <= 1) return 0;
    return (uint32_t)(((hash_key(lane, 0) >> 32) * (uint64_t)P) >> 32);
}

// ---------------------------------------------------------------------------
// sketch: one block per query
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void bt_sketch(const BLaunch *__restrict__ L) {
    __shared__ uint64_t keys[kBSampleSlots];
    __shared__ uint32_t cnts[kBSampleSlots];
    __shared__ uint32_t s_sel, s_d, s_f1, s_f2, s_null;
    const uint32_t qi = blockIdx.x;
    if (qi >= L->n_queries) return;
    TileCols T;
    tile_cols(L, qi, T);
    for (uint32_t i = threadIdx.x; i < kBSampleSlots; i += blockDim.x) { keys[i] = kEmpty; cnts[i] = 0; }
    if (threadIdx.x == 0) { s_sel = s_d = s_f1 = s_f2 = s_null = 0; }
    __syncthreads();
    const uint64_t n = T.n_rows;
    const uint32_t ns = (uint32_t)min<uint64_t>(n, kBSample);
    constexpr int U = kBSample / 256;
    uint64_t r[U], lane[U];
    bool sel[U];
    uint32_t sp[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const uint32_t i = threadIdx.x + u * 256;
        sel[u] = i < ns;
        r[u] = sel[u] ? (uint64_t)i * (n / ns) + (uint64_t)i * (n % ns) / ns : 0;
    }
    eval_pred<U>(T, r, sel);
    eval_keys<U>(L, T, r, lane, sp);
#pragma unroll
    for (int u = 0; u < U; ++u) {
        if (!sel[u]) continue;
        atomicAdd(&s_sel, 1u);
        if (sp[u]) { atomicAdd(&s_null, 1u); continue; }
        uint32_t s = (uint32_t)(hash_key(lane[u], 0) >> 32) & (kBSampleSlots - 1);
        for (int p = 0; p < 16; ++p) {
            const unsigned long long old = atomicCAS((unsigned long long *)&keys[s], (unsigned long long)kEmpty,
                                                     (unsigned long long)lane[u]);
            if (old == kEmpty || old == lane[u]) { atomicAdd(&cnts[s], 1u); break; }
            s = (s + 1) & (kBSampleSlots - 1);
        }
    }
    __syncthreads();
    uint32_t d = 0, f1 = 0, f2 = 0;
    for (uint32_t i = threadIdx.x; i < kBSampleSlots; i += blockDim.x) {
        const uint32_t c = cnts[i];
        d += c > 0;
        f1 += c == 1;
        f2 += c == 2;
    }
    atomicAdd(&s_d, d);
    atomicAdd(&s_f1, f1);
    atomicAdd(&s_f2, f2);
    __syncthreads();
    if (threadIdx.x == 0) {
        L->sk[qi * 5 + 0] = s_sel;
        L->sk[qi * 5 + 1] = s_d;
        L->sk[qi * 5 + 2] = s_f1;
        L->sk[qi * 5 + 3] = s_f2;
        L->sk[qi * 5 + 4] = s_null;
    }
}

// ---------------------------------------------------------------------------
// the captured chain
// ---------------------------------------------------------------------------
__global__ void bt_fill(const BLaunch *__restrict__ L) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    const uint64_t i0 = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (uint64_t i = i0; i < L->n_parts; i += stride) L->part_cnt[i] = 0;
    for (uint64_t i = i0; i < L->n_queries; i += stride) { L->out_cursor[i] = 0; L->q_sel[i] = 0; L->q_fb[i] = 0; }
    const GTable &t = L->t;
    const Layout &lay = L->q.layout;
    for (uint64_t i = i0; i < L->fb_slots; i += stride) {
        t.klo[i] = kEmpty;
        t.khi[i] = kEmpty;
        t.kstate[i] = 0;
        for (int a = 0; a < lay.n_arrs; ++a) {
            const int op = lay.init[a];
            if (mo_size(op) == 4) ((uint32_t *)t.arr[a])[i] = (uint32_t)mo_identity(op);
            else ((uint64_t *)t.arr[a])[i] = mo_identity(op);
        }
    }
    if (i0 < 16) t.ctrl[i0] = 0;
}

// Tiles of every query: the histogram pass counts selected rows per (query,
// partition); the scatter pass writes them partition-contiguous as records
// (key lane, value bits, flags).  A thread takes 16 rows of its tile (at
// most 8192 rows, 512 threads) in passes of 8 (the count pass: of 4), each issuing all its
// loads before using any.  The scatter pass ranks the tile's rows within
// their partitions, then lists them in partition order in LDS, 4 bytes a row
// (value validity bits << 25 | partition << 13 | row in tile) beside their
// value bits, which are loaded in row order; it then walks the list,
// re-reading only each row's key (its lines are in L2 from the ranking), so
// consecutive lanes write consecutive records of a partition's run.  Its LDS
// is dynamic: hist and base (L->scat_parts entries each), the list, then the
// value bits (n_vals arrays of tile_rows).
template <bool SCATTER>
__global__ __launch_bounds__(512) void bt_tiles(const BLaunch *__restrict__ L) {
    extern __shared__ __align__(16) uint8_t smem[];
    __shared__ uint32_t s_hist[SCATTER ? 1 : kBMaxParts];
    __shared__ uint32_t s_sel, s_wave[16];
    // counts; then (scatter) each run's start in the list
    uint32_t *hist = SCATTER ? (uint32_t *)smem : s_hist;
    constexpr int U = SCATTER ? 8 : 4, kPass = 16 / U;  // 16 rows a thread; the count pass holds fewer words
    const int nv = L->q.n_vals;
    const int vtag[2] = {L->q.cols[L->q.vals[0].col].tag, nv > 1 ? L->q.cols[L->q.vals[1].col].tag : 0};
    const uint32_t bs = blockDim.x, tid = threadIdx.x;
    const uint32_t tile_rows = L->tile_rows;
    for (uint32_t t = blockIdx.x; t < L->n_tiles; t += gridDim.x) {
        const uint32_t qi = L->tile_q[t];
        TileCols T;
        tile_cols(L, qi, T);
        const uint64_t r0 = (uint64_t)(t - T.tile_base) * tile_rows;
        const uint64_t r1 = min<uint64_t>(T.n_rows, r0 + tile_rows);
        for (uint32_t p = tid; p < T.P; p += bs) hist[p] = 0;
        if (tid == 0) s_sel = 0;
        __syncthreads();
        // per row: special << 28 | partition << 16 | rank, or ~0u when the
        // row is not selected (partition < 4096, rank < 8192)
        uint32_t pr[kPass][U];
        uint32_t mysel = 0;
#pragma unroll
        for (int ps = 0; ps < kPass; ++ps) {
#pragma unroll
            for (int u = 0; u < U; ++u) pr[ps][u] = ~0u;
            if ((uint32_t)ps * U * bs >= tile_rows) continue;
            uint64_t r[U], lane[U];
            bool sel[U];
            uint32_t sp[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const uint64_t rr = r0 + (uint64_t)(ps * U + u) * bs + tid;
                sel[u] = rr < r1;
                r[u] = sel[u] ? rr : r0;
            }
            eval_pred<U>(T, r, sel);
            eval_keys<U>(L, T, r, lane, sp);
#pragma unroll
            for (int u = 0; u < U; ++u) {
                if (!sel[u]) continue;
                ++mysel;
                const uint32_t p = bq_part(T.P, lane[u], sp[u]);
                pr[ps][u] = (sp[u] << 28) | (p << 16) | atomicAdd(&hist[p], 1u);
            }
        }
        if (mysel) atomicAdd(&s_sel, mysel);
        __syncthreads();
        if (!SCATTER) {
            for (uint32_t p = tid; p < T.P; p += bs)
                if (hist[p]) atomicAdd(&L->part_cnt[T.part_base + p], hist[p]);
            if (tid == 0 && s_sel) atomicAdd(&L->q_sel[qi], s_sel);
        } else {
            uint32_t *base = hist + L->scat_parts;  // each run's start in the query's records
            uint32_t *list = base + L->scat_parts;
            uint64_t *vals = (uint64_t *)(list + tile_rows);
            // runs claimed in the query's records; list starts by a block
            // scan of the counts, each thread over a chunk of partitions
            const uint64_t qoff = L->part_off[T.part_base];
            const uint32_t per = (T.P + bs - 1) / bs;
            const uint32_t p0 = min(T.P, tid * per), p1 = min(T.P, p0 + per);
            uint32_t sum = 0;
            for (uint32_t p = p0; p < p1; ++p) {
                const uint32_t h = hist[p];
                if (h)
                    base[p] = (uint32_t)(atomicAdd((unsigned long long *)&L->part_cur[T.part_base + p],
                                                   (unsigned long long)h) -
                                         qoff);
                sum += h;
            }
            uint32_t acc = block_excl_scan(sum, s_wave);
            for (uint32_t p = p0; p < p1; ++p) {
                const uint32_t h = hist[p];
                hist[p] = acc;
                acc += h;
            }
            __syncthreads();
            // list entries, and value bits loaded in row order
#pragma unroll
            for (int ps = 0; ps < kPass; ++ps) {
                if ((uint32_t)ps * U * bs >= tile_rows) continue;
                uint64_t r[U], vw[2][U];
                uint32_t vb[2][U];
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const uint64_t rr = r0 + (uint64_t)(ps * U + u) * bs + tid;
                    r[u] = pr[ps][u] != ~0u ? rr : r0;
                }
#pragma unroll
                for (int j = 0; j < 2; ++j) {
                    if (j >= nv) break;
                    col_bits<U>(vtag[j], T.vc[j].d, r, vw[j]);
                    col_valid<U>(T.vc[j].v, r, vb[j]);
                }
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const uint32_t x = pr[ps][u];
                    if (x == ~0u) continue;
                    const uint32_t p = (x >> 16) & 0xFFFu, s = hist[p] + (x & 0xFFFFu);
                    uint32_t ent = (p << 13) | ((uint32_t)(ps * U + u) * bs + tid);
#pragma unroll
                    for (int j = 0; j < 2; ++j) {
                        if (j >= nv) break;
                        vals[(uint32_t)j * tile_rows + s] = vw[j][u];
                        ent |= (vb[j][u] & 1u) << (25 + j);
                    }
                    list[s] = ent;
                }
            }
            __syncthreads();
            const uint32_t nsel = s_sel;
            for (uint32_t s0 = 0; s0 < nsel; s0 += U * bs) {
                uint64_t r[U], lane[U];
                uint32_t sp[U], ent[U];
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const uint32_t s = s0 + (uint32_t)u * bs + tid;
                    ent[u] = s < nsel ? list[s] : ~0u;
                    r[u] = ent[u] != ~0u ? r0 + (ent[u] & 8191u) : r0;
                }
                eval_keys<U>(L, T, r, lane, sp);
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    if (ent[u] == ~0u) continue;
                    const uint32_t p = (ent[u] >> 13) & 0xFFFu, s = s0 + (uint32_t)u * bs + tid;
                    const uint64_t pos = qoff + base[p] + (s - hist[p]);
                    uint16_t fl = sp[u] == 1 ? kBAllOnes : (sp[u] == 2 ? kBNull : 0);
#pragma unroll
                    for (int j = 0; j < 2; ++j) {
                        if (j >= nv) break;
                        gst<uint64_t>(L->rv[j], pos, vals[(uint32_t)j * tile_rows + s]);
                        fl |= (uint16_t)(((ent[u] >> (25 + j)) & 1u) << j);
                    }
                    gst<uint64_t>(L->rk, pos, lane[u]);
                    gst<uint16_t>(L->rf, pos, fl);
                }
            }
        }
        __syncthreads();
    }
}

// Exclusive scan of part_cnt into part_off (and part_cur): blocks of 4096
// entries scan locally and publish their sums, one block scans the sums,
// then the sums are added back.
__global__ __launch_bounds__(1024) void bt_scan_local(const BLaunch *__restrict__ L) {
    __shared__ uint64_t s[1024];
    const uint64_t n = L->n_parts;
    const uint64_t nblk = (n + 4095) / 4096;
    for (uint64_t b = blockIdx.x; b < nblk; b += gridDim.x) {
        uint64_t x[4], tot = 0;
        for (int k = 0; k < 4; ++k) {
            const uint64_t i = b * 4096 + threadIdx.x * 4 + k;
            x[k] = i < n ? L->part_cnt[i] : 0;
            tot += x[k];
        }
        s[threadIdx.x] = tot;
        __syncthreads();
        for (int off = 1; off < 1024; off <<= 1) {
            const uint64_t y = threadIdx.x >= (uint32_t)off ? s[threadIdx.x - off] : 0;
            __syncthreads();
            s[threadIdx.x] += y;
            __syncthreads();
        }
        uint64_t acc = s[threadIdx.x] - tot;
        for (int k = 0; k < 4; ++k) {
            const uint64_t i = b * 4096 + threadIdx.x * 4 + k;
            if (i < n) L->part_off[i] = acc;
            acc += x[k];
        }
        if (threadIdx.x == 1023) L->scan_blk[b] = s[1023];
        __syncthreads();
    }
}

__global__ __launch_bounds__(1024) void bt_scan_blocks(const BLaunch *__restrict__ L) {
    __shared__ uint64_t s[1024];
    __shared__ uint64_t carry;
    const uint64_t nblk = (L->n_parts + 4095) / 4096;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (uint64_t b0 = 0; b0 < nblk; b0 += 1024) {
        const uint64_t b = b0 + threadIdx.x;
        const uint64_t x = b < nblk ? L->scan_blk[b] : 0;
        s[threadIdx.x] = x;
        __syncthreads();
        for (int off = 1; off < 1024; off <<= 1) {
            const uint64_t y = threadIdx.x >= (uint32_t)off ? s[threadIdx.x - off] : 0;
            __syncthreads();
            s[threadIdx.x] += y;
            __syncthreads();
        }
        if (b < nblk) L->scan_blk[b] = carry + s[threadIdx.x] - x;
        __syncthreads();
        if (threadIdx.x == 1023) carry += s[1023];
        __syncthreads();
    }
    if (threadIdx.x == 0) L->part_off[L->n_parts] = carry;
}

__global__ void bt_scan_add(const BLaunch *__restrict__ L) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < L->n_parts; i += stride) {
        const uint64_t v = L->part_off[i] + L->scan_blk[i / 4096];
        L->part_off[i] = v;
        L->part_cur[i] = v;
    }
}

// One record of a batch partition, with its value bits already loaded.
struct BRec {
    const BLaunch *__restrict__ L;
    uint16_t fl;
    uint64_t v0, v1;
    __device__ __forceinline__ bool valid(int c) const { return (fl >> L->col_val[c]) & 1; }
    __device__ __forceinline__ uint64_t bits(int c) const { return L->col_val[c] ? v1 : v0; }
    __device__ __forceinline__ int64_t i64(int c) const { return raw_to_i64(L->q.cols[c].tag, bits(c)); }
    __device__ __forceinline__ double f64(int c) const { return raw_to_f64(L->q.cols[c].tag, bits(c)); }
};

// (query, partition) work items: fold the partition's records into an LDS
// table, then emit its groups into the query's output rows.  The table has
// cl live slots, the smallest power of two holding twice the partition's
// records (at most part_slots), then the all-ones and null side slots.
// The whole table is set to its identities once per block; after that the
// emit puts back each slot it emits, so every partition starts from a clean
// table without a pass over all its slots.
// SIG: the plan's compile-time aggregate signature (0: resolved at run time).
template <int SIG>
__global__ __launch_bounds__(1024) void bt_agg(const BLaunch *__restrict__ L, const uint32_t *lds_off) {
    extern __shared__ __align__(16) uint8_t smem[];
    __shared__ uint32_t s_count, s_base, s_pwave[16];
    __shared__ uint16_t s_live[kBMaxSlots + 2];  // live slots of the partition, dense
    // per state array: LDS byte offset, identity, element bytes
    __shared__ uint32_t s_aoff[kMaxArrs], s_asz[kMaxArrs];
    __shared__ uint64_t s_aid[kMaxArrs];
    const QueryDesc *__restrict__ q = &L->q;
    const uint32_t CL = L->part_slots;
    const uint32_t nls_max = CL + 2;
    uint64_t *lk = (uint64_t *)smem;
    uint8_t *state = (uint8_t *)(lk + nls_max);
    LSink ls{state, lds_off, nls_max};
    GSink gs{L->t.arr, L->fb_slots};
    const LdsAgg2 A = lds_agg2(q, state, lds_off, nls_max);
    const uint32_t tid = threadIdx.x, bs = blockDim.x;
    const int nv = q->n_vals;
    const int n_arrs = q->layout.n_arrs;
    constexpr int U = 4;
    if (tid < (uint32_t)n_arrs) {
        const int op = q->layout.init[tid];
        s_aoff[tid] = lds_off[tid];
        s_asz[tid] = (uint32_t)mo_size(op);
        s_aid[tid] = mo_size(op) == 4 ? (uint64_t)(uint32_t)mo_identity(op) : mo_identity(op);
    }
    __syncthreads();
    // slot i of every state array back to its identity
    auto reset_slot = [&](uint32_t i) {
        for (int a = 0; a < n_arrs; ++a) {
            uint8_t *p = state + s_aoff[a];
            if (s_asz[a] == 4) ((uint32_t *)p)[i] = (uint32_t)s_aid[a];
            else ((uint64_t *)p)[i] = s_aid[a];
        }
    };
    for (uint32_t i = tid; i < nls_max; i += bs) {
        lk[i] = kEmpty;
        reset_slot(i);
    }
    __syncthreads();
    for (uint64_t pi = blockIdx.x; pi < L->n_parts; pi += gridDim.x) {
        const uint64_t b = L->part_off[pi], e = L->part_off[pi + 1];
        if (b == e) continue;  // uniform: the whole block skips
        const uint32_t qi = L->part_q[pi];
        uint32_t cl = 64;
        while (cl < CL && (uint64_t)cl < 2 * (e - b)) cl *= 2;
        const uint32_t nls = cl + 2;
        for (uint64_t i0 = b + tid; i0 < e; i0 += (uint64_t)U * bs) {
            uint64_t ln[U], v0[U], v1[U];
            uint16_t fl[U];
            bool in[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const uint64_t i = i0 + (uint64_t)u * bs;
                in[u] = i < e;
                const uint64_t ii = in[u] ? i : b;
                fl[u] = gld<uint16_t>(L->rf, ii);
                ln[u] = gld<uint64_t>(L->rk, ii);
                v0[u] = nv > 0 ? gld<uint64_t>(L->rv[0], ii) : 0;
                v1[u] = nv > 1 ? gld<uint64_t>(L->rv[1], ii) : 0;
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                if (!in[u]) continue;
                const BRec row{L, fl[u], v0[u], v1[u]};
                const uint64_t lane = ln[u];
                const int special = (fl[u] & kBAllOnes) ? 1 : ((fl[u] & kBNull) ? 2 : 0);
                int32_t slot = -1;
                uint64_t h = 0;
                if (special) {
                    slot = (int32_t)(cl + special - 1);
                } else {
                    h = hash_key(lane, 0);
                    uint32_t s = (uint32_t)(h >> 20) & (cl - 1);
                    for (int pr = 0; pr < kBProbe; ++pr) {
                        const uint64_t cur = __hip_atomic_load(&lk[s], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                        if (cur == lane) { slot = (int32_t)s; break; }
                        if (cur == kEmpty) {
                            const unsigned long long old = atomicCAS((unsigned long long *)&lk[s],
                                                                     (unsigned long long)kEmpty, (unsigned long long)lane);
                            if (old == kEmpty || old == lane) { slot = (int32_t)s; break; }
                        }
                        // triangular steps visit every slot of the power-of-two
                        // table and break up the runs linear steps pile into at
                        // the 0.8 load partitions are sized for
                        s = (s + (uint32_t)pr + 1) & (cl - 1);
                    }
                }
                if (slot >= 0) {
                    lds_agg2_add_sig<SIG>(A, (uint32_t)slot, fl[u], v0[u], v1[u]);
                } else {
                    const Key k{lane, (uint64_t)qi, 0};
                    if (!L->q_fb[qi]) L->q_fb[qi] = 1;
                    const int64_t g = g_find128(L->t, k, h ^ ((uint64_t)qi * 0x9E3779B97F4A7C15ull));
                    if (g >= 0) add_row(gs, q, (uint64_t)g, row);
                }
            }
        }
        __syncthreads();
        // emit this partition's groups into the query's output rows
        const BQuery *__restrict__ Qp = L->qs + qi;
        const uint64_t out_cap = Qp->out_cap, out_base = Qp->out_base;
        const int a_len = q->layout.a_len;
        // the live slots, compacted in thread order, so the groups are
        // emitted by as few waves as they fill and in one pass
        uint32_t my = 0;
        for (uint32_t i = tid; i < nls; i += bs) my += *ls.p<uint32_t>(a_len, i) != 0;
        uint32_t pos = block_excl_scan(my, s_pwave);
        if (tid == bs - 1) {
            s_count = pos + my;
            s_base = pos + my ? atomicAdd(&L->out_cursor[qi], pos + my) : 0;
        }
        for (uint32_t i = tid; i < nls; i += bs)
            if (*ls.p<uint32_t>(a_len, i) != 0) s_live[pos++] = (uint16_t)i;
        __syncthreads();
        const uint32_t n_live = s_count;
        const uint64_t base = s_base;
        LReader rd{state, lds_off};
        for (uint32_t j = tid; j < n_live; j += bs) {
            const uint32_t i = s_live[j];
            const int special = i >= cl ? (int)(i - cl) + 1 : 0;
            if (base + j < out_cap)
                emit_group(q, L->out, L->n_out, rd, i, special ? 0 : lk[i], 0, special, out_base + base + j);
            // the slot back to empty for the next partition (also when its
            // group did not fit: the query re-runs alone)
            lk[i] = kEmpty;
            reset_slot(i);
        }
        __syncthreads();
    }
}

__global__ __launch_bounds__(256) void bt_fb_emit(const BLaunch *__restrict__ L) {
    const GTable &t = L->t;
    const uint64_t n = min<uint64_t>(t.ctrl[0], t.list_cap);
    const GReader rd{t.arr};
    for (uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; j < n; j += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t s = t.list[j];
        const uint32_t qi = (uint32_t)t.khi[s];
        const BQuery &Q = L->qs[qi];
        const uint64_t o = atomicAdd(&L->out_cursor[qi], 1u);
        if (o < Q.out_cap) emit_group(&L->q, L->out, L->n_out, rd, s, t.klo[s], 0, 0, Q.out_base + o);
    }
}

// per query: validity bytes -> bitmaps at the query's own bit 0
// Byte flags of rows [0, ng) at src into the bitmap at dst ((ng + 7) / 8
// bytes): each wave takes 64-row chunks, one coalesced byte per lane, and
// eight lanes write the ballot's eight bytes; four chunks' loads are issued
// before their ballots.
__device__ __forceinline__ void pack_flags(const uint8_t *src, uint8_t *dst, uint64_t ng) {
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    const uint64_t nbytes = (ng + 7) / 8, nchunks = (ng + 63) / 64;
    constexpr int U = 4;
    for (uint64_t c0 = wave; c0 < nchunks; c0 += (uint64_t)U * nw) {
        uint8_t f[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const uint64_t r = (c0 + (uint64_t)u * nw) * 64 + lane;
            f[u] = r < ng ? src[r] : 0;
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const uint64_t c = c0 + (uint64_t)u * nw;
            const uint64_t m = __ballot(f[u] != 0);
            if (c < nchunks && lane < 8 && c * 8 + lane < nbytes) dst[c * 8 + lane] = (uint8_t)(m >> (8 * lane));
        }
    }
}

__global__ __launch_bounds__(1024) void bt_pack(const BLaunch *__restrict__ L) {
    for (uint32_t qi = blockIdx.x; qi < L->n_queries; qi += gridDim.x) {
        const BQuery &Q = L->qs[qi];
        const uint64_t ng = min<uint64_t>(L->out_cursor[qi], Q.out_cap);
        for (int o = 0; o < L->n_out; ++o) {
            const OutCol &oc = L->out[o];
            if (oc.vbytes && oc.valid) pack_flags(oc.vbytes + Q.out_base, (uint8_t *)oc.valid + Q.vb_base, ng);
            if (oc.src == OS_KEY && oc.tag == PMH_BOOL && oc.data_bits)
                pack_flags((const uint8_t *)oc.data + Q.out_base, oc.data_bits + Q.vb_base, ng);
        }
    }
}

void launch_check() { PMH_HIP(hipGetLastError()); }

uint64_t pow2_ceil(uint64_t x) {
    uint64_t p = 1;
    while (p < x) p *= 2;
    return p;
}

// The captured chain and the buffers it was captured against.
struct BatchGraph {
    hipGraphExec_t exec = nullptr;
    std::string key;  // buffer addresses and launch shapes
};
// The chains captured for recent batch shapes, most recent first: a stream
// that alternates plan templates replays each one's graph instead of
// capturing again.  A graph is replayed only when every buffer address and
// launch shape it was captured with is the current one.
constexpr size_t kGraphCache = 8;
std::vector<BatchGraph> g_graphs;

}  // namespace

void run_batch(const GroupByPlan &tmpl, const std::vector<BatchQueryIn> &qs, uint32_t flags,
               std::vector<pmh_result *> &results, pmh_batch_stats *stats) {
    Runtime &R = rt();
    const uint32_t nq = (uint32_t)qs.size();
    results.assign(nq, nullptr);
    if (!nq) return;
    const int cus = R.prop.multiProcessorCount > 0 ? R.prop.multiProcessorCount : 256;
    const size_t lds_max = std::min<size_t>(R.prop.sharedMemPerBlock ? R.prop.sharedMemPerBlock : 65536, 160 * 1024);
    // ---- the template, prepared against query 0's columns
    Prep P;
    prepare_query(P, tmpl, qs[0].n_rows, qs[0].cols, qs[0].n_cols);
    const QueryDesc &q0 = P.q;
    if (q0.n_cols > kBCols || q0.n_vals > 2 || q0.n_keys > kBKeys ||
        !(q0.key_mode == KM_RAW64 || q0.key_mode == KM_LANE64))
        fail(PMH_E_NOTIMPL, "polars_metal: batch needs one key lane of at most 64 bits from at most 4 key columns, "
                            "at most two aggregated columns and at most 8 referenced columns");
    auto Lh = std::make_unique<BLaunch>();
    memset(Lh.get(), 0, sizeof(BLaunch));
    BLaunch &L = *Lh;
    L.q = q0;
    L.n_queries = nq;
    L.n_out = P.n_out;
    L.lane_mode = q0.key_mode == KM_RAW64 ? 1 : 0;
    for (int j = 0; j < q0.n_vals; ++j) L.col_val[q0.vals[j].col] = j;
    const Options &opt = tmpl.opts;
    L.tile_rows = (uint32_t)std::min<int64_t>(8192, std::max<int64_t>(1024, opt.tile > 0 ? opt.tile : 8192));
    L.tile_rows = pow2_ceil(L.tile_rows);  // 16 rows per thread of a 64..512-thread block
    // the scatter's LDS (hist and base for up to kBMaxParts partitions, the
    // list, the value bits) must fit
    while (L.tile_rows > 1024 &&
           (size_t)8 * kBMaxParts + (size_t)L.tile_rows * (4 + 8 * q0.n_vals) > lds_max - 1024)
        L.tile_rows /= 2;
    // LDS table of a (query, partition) work item: the largest power of two
    // up to kBMaxSlots that fits beside bt_agg's static LDS (its live-slot
    // list and state-array table, under 10 KB), or the option's size;
    // partitions are sized for 0.8 of it
    size_t state_per_slot = 0;
    for (int a = 0; a < q0.layout.n_arrs; ++a) state_per_slot += q0.layout.esz[a];
    uint32_t CL = kBMaxSlots;
    while (CL > 256 && (size_t)(CL + 2) * (8 + state_per_slot) > lds_max - 12288) CL /= 2;
    if (opt.tail_slots > 0) {
        uint32_t c = 256;
        while (c * 2 <= (uint64_t)opt.tail_slots && c * 2 <= CL) c *= 2;
        CL = c;
    }
    const double part_keys = CL * 0.8;
    // ---- per query: columns by name, predicate intervals (cached by text)
    std::vector<BQuery> bq(nq);
    std::map<std::string, FastScan> pred_cache;
    std::vector<std::string> names(q0.n_cols);
    for (int c = 0; c < q0.n_cols; ++c) names[c] = P.dcols[c].name;
    for (uint32_t i = 0; i < nq; ++i) {
        BQuery &Q = bq[i];
        memset(&Q, 0, sizeof(Q));
        Q.n_rows = qs[i].n_rows;
        // the same row limit as a single query, which is also what a query
        // the batch cannot finish is re-run on
        if (Q.n_rows >= (1ull << 31))
            fail(PMH_E_VALUE, "polars_metal: batch query " + std::to_string(i) + " holds 2^31 rows or more");
        if (Q.n_rows && (!qs[i].cols || !qs[i].n_cols))
            fail(PMH_E_VALUE, "polars_metal: batch query " + std::to_string(i) + " has no columns");
        for (int c = 0; c < q0.n_cols; ++c) {
            const pmh_column *col = nullptr;
            for (uint32_t k = 0; k < qs[i].n_cols; ++k)
                if (qs[i].cols[k].name && names[c] == qs[i].cols[k].name) col = &qs[i].cols[k];
            if (!col) fail(PMH_E_KEY, "polars_metal: batch query " + std::to_string(i) + " has no column " + names[c]);
            // every query's buffers are checked against its own row count
            // before any launch (query 0's again here, cheaply)
            check_column(*col, Q.n_rows, names[c]);
            if (col->location != PMH_DEVICE)
                fail(PMH_E_VALUE, "polars_metal: batch columns must be device-resident");
            if (col->tag != q0.cols[c].tag)
                fail(PMH_E_VALUE, "polars_metal: batch query " + std::to_string(i) + " column " + names[c] +
                                      " differs in dtype from the template");
            const int w = tag_size(col->tag);
            if (w > 1 && ((uintptr_t)col->data % (uintptr_t)w) != 0)
                fail(PMH_E_VALUE, "polars_metal: batch columns must be naturally aligned");
            Q.data[c] = col->data;
            Q.valid[c] = (const uint8_t *)col->valid;
        }
        const std::string pj = qs[i].predicate ? qs[i].predicate : "";
        auto it = pred_cache.find(pj);
        if (it == pred_cache.end()) {
            FastScan fs{};
            if (!pj.empty()) {
                std::string js = "{\"keys\": [], \"aggs\": [], \"predicate\": " + pj + "}";
                GroupByPlan parsed = parse_groupby_plan(js, true);
                QueryDesc qd = q0;
                lower_predicate_into(qd, parsed, P);
                if (!lower_fast_scan(qd, fs))
                    fail(PMH_E_NOTIMPL, "polars_metal: batch predicates must be conjunctions of column-vs-literal "
                                        "comparisons");
            }
            it = pred_cache.emplace(pj, fs).first;
        }
        Q.fs = it->second;
    }
    // ---- buffers independent of the sketch
    L.qs = (BQuery *)slab("bt:qs", (uint64_t)nq * sizeof(BQuery) + 64);
    L.sk = (uint32_t *)slab("bt:sk", (uint64_t)nq * 5 * 4 + 64);
    BLaunch *dL = (BLaunch *)slab("bt:launch", sizeof(BLaunch));
    hipEvent_t ev0, ev1;
    PMH_HIP(hipEventCreate(&ev0));
    PMH_HIP(hipEventCreate(&ev1));
    PMH_HIP(hipEventRecord(ev0, R.stream));
    PMH_HIP(hipMemcpyAsync(L.qs, bq.data(), (uint64_t)nq * sizeof(BQuery), hipMemcpyHostToDevice, R.stream));
    PMH_HIP(hipMemcpyAsync(dL, &L, sizeof(BLaunch), hipMemcpyHostToDevice, R.stream));
    {
        KernelTimer kt("bt_sketch");
        hipLaunchKernelGGL(bt_sketch, dim3(nq), dim3(256), 0, R.stream, dL);
        launch_check();
    }
    std::vector<uint32_t> sk((size_t)nq * 5);
    PMH_HIP(hipMemcpyAsync(sk.data(), L.sk, sk.size() * 4, hipMemcpyDeviceToHost, R.stream));
    PMH_HIP(hipStreamSynchronize(R.stream));
    // ---- sizes from the sketch (the learned sizer when installed)
    std::vector<double> est(nq);
    std::vector<float> feats;
    const bool learned = sizer_installed() && opt.method == "auto";
    // the batch sketch holds f1 and f2 only: a three-entry count histogram
    std::vector<uint32_t> hist(3, 0);
    for (uint32_t i = 0; i < nq; ++i) {
        const uint32_t ns = (uint32_t)std::min<uint64_t>(bq[i].n_rows, kBSample);
        hist[1] = sk[i * 5 + 2];
        hist[2] = sk[i * 5 + 3];
        Sketch k = sketch_from(sk[i * 5 + 0], ns, bq[i].n_rows, sk[i * 5 + 1], sk[i * 5 + 2], sk[i * 5 + 3], 0, hist);
        est[i] = plain_distinct_estimate(k) + (sk[i * 5 + 4] ? 2 : 0);
        if (learned) {
            float f[PMH_MLP_IN];
            sizer_features(k, sk[i * 5 + 4], q0.total_bits, f);
            feats.insert(feats.end(), f, f + PMH_MLP_IN);
        }
    }
    if (learned) {
        std::vector<float> pred((size_t)nq * 2);
        sizer_predict_many(feats.data(), nq, pred.data());
        for (uint32_t i = 0; i < nq; ++i)
            est[i] = std::max<double>(std::exp2(std::min(40.0, (double)pred[2 * i] * 32.0)) - 1.0, sk[i * 5 + 1]);
    }
    uint64_t n_parts = 0, n_tiles = 0, out_rows = 0, vb_bytes = 0, max_rows = 0;
    for (uint32_t i = 0; i < nq; ++i) {
        BQuery &Q = bq[i];
        const double sel = std::min<uint64_t>(Q.n_rows, kBSample)
                               ? (double)sk[i * 5 + 0] / (double)std::min<uint64_t>(Q.n_rows, kBSample)
                               : 0.0;
        const uint64_t n_sel_est = (uint64_t)std::ceil(sel * Q.n_rows) + 64;
        Q.P = (uint32_t)std::min<uint64_t>(kBMaxParts, (uint64_t)std::max(1.0, std::ceil(est[i] / part_keys)));
        Q.part_base = n_parts;
        n_parts += Q.P;
        Q.n_tiles = (uint32_t)ceil_div(Q.n_rows, L.tile_rows);
        Q.tile_base = (uint32_t)n_tiles;
        n_tiles += Q.n_tiles;
        // output rows: twice the estimate (never more than the selected
        // rows could make); a query that overflows re-runs alone
        Q.out_cap = std::max<uint64_t>(1, std::min<uint64_t>(std::min<uint64_t>(Q.n_rows, n_sel_est),
                                                            (uint64_t)(2.0 * est[i]) + 1024));
        Q.out_base = out_rows;
        out_rows += Q.out_cap;
        Q.vb_base = vb_bytes;
        vb_bytes += (Q.out_cap + 7) / 8 + 8;
        max_rows += Q.n_rows;
    }
    if (n_tiles >= (1ull << 32)) fail(PMH_E_VALUE, "polars_metal: batch too large");
    L.n_tiles = (uint32_t)n_tiles;
    L.n_parts = n_parts;
    uint32_t max_p = 1;
    for (uint32_t i = 0; i < nq; ++i) max_p = std::max(max_p, bq[i].P);
    L.scat_parts = (max_p + 1) & ~1u;  // keeps the value bits 8-byte aligned
    const size_t scat_lds = (size_t)8 * L.scat_parts + (size_t)L.tile_rows * (4 + 8 * q0.n_vals);
    std::vector<uint32_t> tile_q(n_tiles), part_q(n_parts);
    for (uint32_t i = 0; i < nq; ++i) {
        for (uint32_t t = 0; t < bq[i].n_tiles; ++t) tile_q[bq[i].tile_base + t] = i;
        for (uint32_t p = 0; p < bq[i].P; ++p) part_q[bq[i].part_base + p] = i;
    }
    L.part_slots = CL;
    uint32_t lds_off_h[kMaxArrs];
    size_t lds_state = 0;
    for (int a = 0; a < q0.layout.n_arrs; ++a) {
        lds_off_h[a] = (uint32_t)lds_state;
        lds_state += (size_t)(CL + 2) * q0.layout.esz[a];
    }
    const size_t lds_agg = (size_t)(CL + 2) * 8 + lds_state;
    // ---- buffers
    L.tile_q = (uint32_t *)slab("bt:tile_q", n_tiles * 4 + 64);
    L.part_q = (uint32_t *)slab("bt:part_q", n_parts * 4 + 64);
    L.part_cnt = (uint32_t *)slab("bt:part_cnt", n_parts * 4 + 64);
    L.part_off = (uint64_t *)slab("bt:part_off", (n_parts + 1) * 8 + 64);
    L.part_cur = (uint64_t *)slab("bt:part_cur", n_parts * 8 + 64);
    L.scan_blk = (uint64_t *)slab("bt:scan_blk", ((n_parts + 4095) / 4096 + 1) * 8 + 64);
    L.rk = (uint64_t *)slab("bt:rk", max_rows * 8 + 64);
    for (int j = 0; j < q0.n_vals; ++j) L.rv[j] = (uint64_t *)slab("bt:rv" + std::to_string(j), max_rows * 8 + 64);
    L.rf = (uint16_t *)slab("bt:rf", max_rows * 2 + 64);
    L.out_cursor = (uint32_t *)slab("bt:out_cursor", (uint64_t)nq * 4 + 64);
    L.q_sel = (uint32_t *)slab("bt:q_sel", (uint64_t)nq * 4 + 64);
    L.q_fb = (uint32_t *)slab("bt:q_fb", (uint64_t)nq * 4 + 64);
    uint32_t *d_lds_off = (uint32_t *)slab("bt:lds_off", sizeof(lds_off_h));
    {
        GTable &t = L.t;
        // keys that miss a full partition table only: a modest table, and
        // if it overflows every query re-runs alone
        uint64_t fcap = 1u << 16;
        while (fcap < max_rows / 16 && fcap < (1ull << 20)) fcap *= 2;
        const uint64_t nslots = fcap + 2;
        t.mask = fcap - 1;
        t.limit = (uint64_t)(fcap * 0.7);
        t.list_cap = nslots;
        L.fb_slots = nslots;
        const size_t bytes = nslots * (8 + 8 + 4 + 4) + 4096 + state_bytes(P, nslots) + 8192;
        uint8_t *base = (uint8_t *)slab("bt:fb", bytes);
        size_t o = 0;
        auto carve = [&](size_t b) { uint8_t *p = base + o; o += (b + 255) & ~size_t(255); return p; };
        t.klo = (uint64_t *)carve(nslots * 8);
        t.khi = (uint64_t *)carve(nslots * 8);
        t.kstate = (uint32_t *)carve(nslots * 4);
        t.list = (uint32_t *)carve(nslots * 4);
        t.ctrl = (uint32_t *)carve(64);
        carve_state(P, base, o, nslots, t.arr);
    }
    for (int oi = 0; oi < P.n_out; ++oi) {
        OutCol &oc = L.out[oi];
        oc = P.outs[oi];
        const size_t esz = std::max(1, tag_size(oc.tag));
        oc.data = slab("bt:out" + std::to_string(oi), out_rows * esz + 64);
        const bool need_valid = !(oc.src == OS_AGG && (oc.op == AGG_LEN || oc.op == AGG_COUNT));
        oc.vbytes = need_valid ? (uint8_t *)slab("bt:outvb" + std::to_string(oi), out_rows + 64) : nullptr;
        oc.valid = need_valid ? (uint64_t *)slab("bt:outv" + std::to_string(oi), vb_bytes + 64) : nullptr;
        oc.data_bits = (oc.src == OS_KEY && oc.tag == PMH_BOOL) ? (uint8_t *)slab("bt:outbits" + std::to_string(oi),
                                                                                  vb_bytes + 64)
                                                               : nullptr;
    }
    PMH_HIP(hipMemcpyAsync(L.qs, bq.data(), (uint64_t)nq * sizeof(BQuery), hipMemcpyHostToDevice, R.stream));
    PMH_HIP(hipMemcpyAsync(L.tile_q, tile_q.data(), n_tiles * 4, hipMemcpyHostToDevice, R.stream));
    PMH_HIP(hipMemcpyAsync(L.part_q, part_q.data(), n_parts * 4, hipMemcpyHostToDevice, R.stream));
    PMH_HIP(hipMemcpyAsync(d_lds_off, lds_off_h, sizeof(lds_off_h), hipMemcpyHostToDevice, R.stream));
    PMH_HIP(hipMemcpyAsync(dL, &L, sizeof(BLaunch), hipMemcpyHostToDevice, R.stream));
    // the fold is specialised by the plan's aggregate signature
    static const void *const kBtAgg[] = {(const void *)bt_agg<0>, (const void *)bt_agg<1>, (const void *)bt_agg<2>,
                                         (const void *)bt_agg<3>, (const void *)bt_agg<4>};
    static_assert(sizeof(kBtAgg) / sizeof(kBtAgg[0]) == kNumAggSigs, "one instantiation per signature");
    const int sig = agg_signature(q0);
    const void *agg_fn = kBtAgg[sig];
    PMH_HIP(hipFuncSetAttribute(agg_fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_agg));
    PMH_HIP(hipFuncSetAttribute((const void *)bt_tiles<true>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)scat_lds));
    // ---- the chain: captured once per buffer layout, replayed after
    const uint32_t grid = (uint32_t)cus * 4;
    const uint32_t tile_block = L.tile_rows / 16;
    // a table of at most 512 slots leaves room for several blocks per CU
    const uint32_t agg_block = CL >= 1024 ? 1024 : 512;
    char keybuf[512];
    snprintf(keybuf, sizeof(keybuf), "%p %p %p %p %p %p %p %zu %u %u %u %u %d %zu", (void *)dL, (void *)L.rk,
             (void *)L.part_cnt, (void *)L.part_off, (void *)L.t.klo, (void *)L.out[0].data, (void *)d_lds_off, lds_agg,
             grid, (unsigned)P.n_out, tile_block, agg_block, sig, scat_lds);
    size_t gi = 0;
    while (gi < g_graphs.size() && g_graphs[gi].key != keybuf) ++gi;
    const bool reused = gi < g_graphs.size();
    if (reused) {
        std::rotate(g_graphs.begin(), g_graphs.begin() + gi, g_graphs.begin() + gi + 1);
    } else {
        if (g_graphs.size() >= kGraphCache) {
            PMH_HIP(hipGraphExecDestroy(g_graphs.back().exec));
            g_graphs.pop_back();
        }
        hipGraph_t graph;
        PMH_HIP(hipStreamBeginCapture(R.stream, hipStreamCaptureModeThreadLocal));
        hipLaunchKernelGGL(bt_fill, dim3(grid), dim3(256), 0, R.stream, dL);
        // 16 KB of LDS a block: twice the blocks of the other kernels, for
        // more loads in flight
        hipLaunchKernelGGL(bt_tiles<false>, dim3(grid * 2), dim3(tile_block), 0, R.stream, dL);
        hipLaunchKernelGGL(bt_scan_local, dim3(grid), dim3(1024), 0, R.stream, dL);
        hipLaunchKernelGGL(bt_scan_blocks, dim3(1), dim3(1024), 0, R.stream, dL);
        hipLaunchKernelGGL(bt_scan_add, dim3(grid), dim3(256), 0, R.stream, dL);
        hipLaunchKernelGGL(bt_tiles<true>, dim3(grid), dim3(tile_block), scat_lds, R.stream, dL);
        {
            const uint32_t *lo = d_lds_off;
            void *args[] = {(void *)&dL, (void *)&lo};
            PMH_HIP(hipLaunchKernel(agg_fn, dim3(grid), dim3(agg_block), args, lds_agg, R.stream));
        }
        hipLaunchKernelGGL(bt_fb_emit, dim3(grid), dim3(256), 0, R.stream, dL);
        hipLaunchKernelGGL(bt_pack, dim3(grid), dim3(1024), 0, R.stream, dL);
        PMH_HIP(hipStreamEndCapture(R.stream, &graph));
        BatchGraph bg;
        bg.key = keybuf;
        const hipError_t ie = hipGraphInstantiate(&bg.exec, graph, nullptr, nullptr, 0);
        PMH_HIP(hipGraphDestroy(graph));
        PMH_HIP(ie);
        g_graphs.insert(g_graphs.begin(), bg);
    }
    {
        KernelTimer kt("bt_graph");
        PMH_HIP(hipGraphLaunch(g_graphs.front().exec, R.stream));
    }
    std::vector<uint32_t> cursor(nq), qsel(nq), qfb(nq);
    uint32_t fb_ctrl[2];
    PMH_HIP(hipMemcpyAsync(cursor.data(), L.out_cursor, (uint64_t)nq * 4, hipMemcpyDeviceToHost, R.stream));
    PMH_HIP(hipMemcpyAsync(qsel.data(), L.q_sel, (uint64_t)nq * 4, hipMemcpyDeviceToHost, R.stream));
    PMH_HIP(hipMemcpyAsync(qfb.data(), L.q_fb, (uint64_t)nq * 4, hipMemcpyDeviceToHost, R.stream));
    PMH_HIP(hipMemcpyAsync(fb_ctrl, L.t.ctrl, 8, hipMemcpyDeviceToHost, R.stream));
    PMH_HIP(hipEventRecord(ev1, R.stream));
    PMH_HIP(hipEventSynchronize(ev1));
    float dev_ms = 0;
    PMH_HIP(hipEventElapsedTime(&dev_ms, ev0, ev1));
    PMH_HIP(hipEventDestroy(ev0));
    PMH_HIP(hipEventDestroy(ev1));
    collect_kernel_times();
    const bool fb_overflow = fb_ctrl[1] != 0 || fb_ctrl[0] > L.t.list_cap;
    // ---- results.  A query whose groups outgrew its output rows, or that
    // needed the fallback table when it overflowed, re-runs alone; its
    // columns come back in host memory, since the single-query engines'
    // device outputs last only until their next call.
    uint32_t reruns = 0;
    for (uint32_t i = 0; i < nq; ++i) {
        const BQuery &Q = bq[i];
        if ((fb_overflow && qfb[i]) || cursor[i] > Q.out_cap) {
            GroupByPlan one;
            one.keys = tmpl.keys;
            for (const auto &a : tmpl.aggs) one.aggs.push_back(a);
            one.opts = tmpl.opts;
            one.opts.method = "auto";
            if (qs[i].predicate) {
                std::string js = "{\"keys\": [], \"aggs\": [], \"predicate\": " + std::string(qs[i].predicate) + "}";
                one.predicate = std::move(parse_groupby_plan(js, true).predicate);
            }
            results[i] = run_filter_groupby(one, qs[i].n_rows, qs[i].cols, qs[i].n_cols, PMH_OUT_HOST);
            ++reruns;
            continue;
        }
        const uint64_t ng = cursor[i];
        pmh_result *res = new_result(P.n_out);
        res->n_rows = ng;
        res->stats.method = PMH_METHOD_HASH;
        res->stats.n_passes = 1;
        res->stats.n_groups = ng;
        res->stats.n_selected = qsel[i];
        res->stats.partitions = Q.P;
        res->stats.distinct_estimate = est[i];
        res->stats.sizer_choice = learned ? 1 : 0;
        const uint64_t vbytes = bitmap_bytes_padded(ng);
        const uint64_t packed = (ng + 7) / 8;
        for (int oi = 0; oi < P.n_out; ++oi) {
            const OutCol &oc = L.out[oi];
            pmh_out_column &col = res->columns[oi];
            set_out_name(col, P.names[oi]);
            col.tag = oc.tag;
            const bool is_bool = oc.src == OS_KEY && oc.tag == PMH_BOOL;
            const size_t esz = std::max(1, tag_size(oc.tag));
            col.data_len = is_bool ? vbytes : ng * tag_size(oc.tag);
            col.valid_len = vbytes;
            const uint8_t *dsrc = is_bool ? oc.data_bits + Q.vb_base : (const uint8_t *)oc.data + Q.out_base * esz;
            const uint8_t *vsrc = oc.valid ? (const uint8_t *)oc.valid + Q.vb_base : nullptr;
            if (flags & PMH_OUT_DEVICE) {
                col.location = PMH_DEVICE;
                col.data = (void *)dsrc;
                col.valid = (uint8_t *)vsrc;
            } else {
                col.location = PMH_HOST;
                col.data = malloc(col.data_len ? col.data_len : 1);
                col.valid = (uint8_t *)malloc(vbytes);
                memset(col.valid, 0, vbytes);
                if (is_bool) memset(col.data, 0, col.data_len);
                const uint64_t nb_copy = is_bool ? packed : col.data_len;
                if (nb_copy) PMH_HIP(hipMemcpyAsync(col.data, dsrc, nb_copy, hipMemcpyDeviceToHost, R.stream));
                if (vsrc) {
                    if (packed) PMH_HIP(hipMemcpyAsync(col.valid, vsrc, packed, hipMemcpyDeviceToHost, R.stream));
                } else {
                    memset(col.valid, 0xFF, vbytes);
                }
            }
        }
        results[i] = res;
    }
    PMH_HIP(hipStreamSynchronize(R.stream));
    if (stats) {
        stats->device_ms = dev_ms;
        stats->n_queries = nq;
        stats->n_tiles = (uint32_t)n_tiles;
        stats->n_partitions = n_parts;
        stats->n_rerun = reruns;
        stats->tile_rows = L.tile_rows;
        stats->part_slots = CL;
        stats->graph_reused = reused ? 1 : 0;
        stats->learned = learned ? 1 : 0;
        stats->n_fallback_keys = (uint32_t)std::min<uint64_t>(fb_ctrl[0], L.t.list_cap);
    }
}

}  // namespace pmh
