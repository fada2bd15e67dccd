// pmh_groupby.hip -- fused filter + hash group-by + aggregation for gfx950.
//
// Replaces the reference's GroupBy(Filter(Scan)) path: the Polars-CPU
// compaction (python/polars_metal/_udf.py:550-597), key encoding
// (crates/polars-metal-kernels/src/groupby.rs:155-302), the build
// (groupby.rs:450-538, groupby_build_partitioned/, groupby_global_hash/) and
// the per-aggregation dispatches (groupby.rs:2189-2496).
//
// Design (one pass over the input, no materialised intermediates):
//   fgb_main      every block streams a slice of rows, evaluates the
//                 predicate in registers, and folds each surviving row into
//                 a block-local open-addressing table in LDS; rows whose key
//                 finds no LDS slot go straight to the global table.  At the
//                 end the block merges its LDS table into the global one.
//   global table  open addressing on 64-bit key words claimed by one CAS
//                 (lane keys wider than 64 bits use a claim word with
//                 release/acquire publication); one state array per
//                 aggregate (SoA), merged with native atomics; exact f64
//                 sums through five 32-bit-chunk int64 limbs.
//   fgb_finalize  walks the dense list of claimed slots, decodes keys,
//                 writes every output column (validity through wave
//                 ballots) and resets the slots it read, so the table is
//                 clean for the next call without a full clear.
//   plain-method resize: a pass whose distinct-key count exceeds 70% of
//   capacity raises a flag, the remaining blocks stop early, and the query
//   re-runs with 4x the capacity.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <functional>

#include "pmh_table.h"

namespace pmh {

// ---------------------------------------------------------------------------
// launch descriptor (device memory, read with scalar loads)
// ---------------------------------------------------------------------------
struct Launch {
    QueryDesc q;
    GTable t;
    uint32_t lds_slots;  // block-local table size (power of two), 0: none
    uint32_t lds_off[kMaxArrs];
    uint64_t tile;
    int32_t n_out, pad;
    OutCol out[kOutMax];
};


// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------
__global__ void fgb_init_table(const Launch *__restrict__ L, uint64_t nslots, int wide) {
    const GTable &t = L->t;
    const Layout &lay = L->q.layout;
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nslots; i += stride) {
        t.klo[i] = kEmpty;
        if (wide) { t.khi[i] = kEmpty; t.kstate[i] = 0; }
        for (int a = 0; a < lay.n_arrs; ++a) {
            int op = lay.init[a];
            if (mo_size(op) == 4) ((uint32_t *)t.arr[a])[i] = (uint32_t)mo_identity(op);
            else ((uint64_t *)t.arr[a])[i] = mo_identity(op);
        }
    }
    if (blockIdx.x == 0 && threadIdx.x < 8) t.ctrl[threadIdx.x] = 0;
}

// One selected-or-not row of the narrow-key path: predicate, key, then the
// LDS table (probe / claim) or, when the row finds no LDS slot, the global
// table.
template <class Row>
__device__ __forceinline__ void narrow_row(const Launch *__restrict__ L, const QueryDesc *__restrict__ q,
                                           const Row &row, uint64_t *lkeys, uint32_t CL, LSink &ls, GSink &gs,
                                           uint32_t &my_sel) {
    if (!eval_pred(q, row)) return;
    ++my_sel;
    Key k = load_key(q, row);
    int32_t lslot = -1;
    uint64_t h = 0;
    if (k.special) {
        if (CL) lslot = (int32_t)(CL + k.special - 1);
    } else {
        h = hash_key(k.lo, 0);
        if (CL) {
            uint32_t s = (uint32_t)(h >> 40) & (CL - 1);
            for (int p = 0; p < kLdsProbe; ++p) {
                uint64_t cur = __hip_atomic_load(&lkeys[s], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                if (cur == k.lo) { lslot = (int32_t)s; break; }
                if (cur == kEmpty) {
                    unsigned long long old = atomicCAS((unsigned long long *)&lkeys[s], (unsigned long long)kEmpty,
                                                       (unsigned long long)k.lo);
                    if (old == kEmpty || old == k.lo) { lslot = (int32_t)s; break; }
                }
                s = (s + 1) & (CL - 1);
            }
        }
    }
    if (lslot >= 0) {
        add_row(ls, q, (uint32_t)lslot, row);
    } else {
        int64_t g = g_find64(L->t, k, h);
        if (g >= 0) add_row(gs, q, (uint64_t)g, row);
    }
}

// Narrow keys (KM_NONE / KM_LANE64 / KM_RAW64): LDS table + global table.
// CACHED: the query touches at most kCacheCols columns; each thread loads U
// rows of all of them into registers before evaluating any, so every wave
// keeps U x (columns) independent loads in flight.
constexpr int kCacheCols = 4;

template <int U, bool CACHED>
__global__ __launch_bounds__(1024) void fgb_main_narrow(const Launch *__restrict__ L) {
    extern __shared__ __align__(16) uint8_t smem[];
    __shared__ int s_stop;
    __shared__ uint32_t s_sel;
    const QueryDesc *__restrict__ q = &L->q;
    const uint32_t CL = L->lds_slots;
    const uint32_t nls = CL + 2;
    uint64_t *lkeys = (uint64_t *)smem;
    LSink ls{smem + (size_t)nls * 8, L->lds_off, nls};
    GSink gs{L->t.arr, L->t.mask + 3};
    const int32_t *ops = q->layout.init;
    const uint32_t tid = threadIdx.x, bs = blockDim.x;

    if (tid == 0) s_sel = 0;
    if (CL) {
        for (uint32_t i = tid; i < nls; i += bs) lkeys[i] = kEmpty;
        for (int a = 0; a < q->layout.n_arrs; ++a) {
            int op = ops[a];
            if (mo_size(op) == 4) {
                uint32_t id = (uint32_t)mo_identity(op);
                for (uint32_t i = tid; i < nls; i += bs) *ls.p<uint32_t>(a, i) = id;
            } else {
                uint64_t id = mo_identity(op);
                for (uint32_t i = tid; i < nls; i += bs) *ls.p<uint64_t>(a, i) = id;
            }
        }
    }
    __syncthreads();

    const uint64_t n = q->n_rows;
    const uint64_t tile = L->tile;
    const int ncols = q->n_cols;
    uint32_t my_sel = 0;
    for (uint64_t t0 = (uint64_t)blockIdx.x * tile; t0 < n; t0 += (uint64_t)gridDim.x * tile) {
        if (tid == 0) s_stop = (int)__hip_atomic_load(&L->t.ctrl[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __syncthreads();
        const int stop = s_stop;
        __syncthreads();
        if (stop) break;
        const uint64_t tend = min(t0 + tile, n);
        if constexpr (CACHED) {
            static_assert(U == 4, "the cached path loads four rows per thread");
            auto fill = [&](CacheRow &row, uint64_t r) __attribute__((always_inline)) {
                row.q = q;
                row.vm = 0;
                row.v0 = row.v1 = row.v2 = row.v3 = 0;
                if (r < tend) {
#pragma unroll
                    for (int c = 0; c < kCacheCols; ++c) {
                        if (c < ncols) {
                            const DevCol &col = q->cols[c];
                            row.set(c, load_raw(col, r));
                            row.vm |= (uint32_t)dvalid(col.valid, r) << c;
                        }
                    }
                }
            };
            for (uint64_t r0 = t0 + tid; r0 < tend; r0 += 4ull * bs) {
                CacheRow a0, a1, a2, a3;
                fill(a0, r0);
                fill(a1, r0 + bs);
                fill(a2, r0 + 2ull * bs);
                fill(a3, r0 + 3ull * bs);
                narrow_row(L, q, a0, lkeys, CL, ls, gs, my_sel);
                if (r0 + bs < tend) narrow_row(L, q, a1, lkeys, CL, ls, gs, my_sel);
                if (r0 + 2ull * bs < tend) narrow_row(L, q, a2, lkeys, CL, ls, gs, my_sel);
                if (r0 + 3ull * bs < tend) narrow_row(L, q, a3, lkeys, CL, ls, gs, my_sel);
            }
        } else {
            for (uint64_t r = t0 + tid; r < tend; r += bs) narrow_row(L, q, MemRow{q, r}, lkeys, CL, ls, gs, my_sel);
        }
    }
    atomicAdd(&s_sel, my_sel);
    __syncthreads();
    if (tid == 0) atomicAdd(&L->t.ctrl[4], s_sel);
    if (CL) {
        const int a_len = q->layout.a_len;
        for (uint32_t i = tid; i < nls; i += bs) {
            if (*ls.p<uint32_t>(a_len, i) == 0) continue;
            Key k{0, 0, 0};
            uint64_t h = 0;
            if (i >= CL) k.special = (int)(i - CL) + 1;
            else { k.lo = lkeys[i]; h = hash_key(k.lo, 0); }
            int64_t g = g_find64(L->t, k, h);
            if (g >= 0) merge_lds_entry(gs, (uint64_t)g, ls, i, q->layout, ops);
        }
    }
}

// Wide keys (KM_LANE128): global table only.
__global__ __launch_bounds__(256) void fgb_main_wide(const Launch *__restrict__ L) {
    __shared__ int s_stop;
    __shared__ uint32_t s_sel;
    const QueryDesc *__restrict__ q = &L->q;
    GSink gs{L->t.arr, L->t.mask + 3};
    const uint32_t tid = threadIdx.x, bs = blockDim.x;
    if (tid == 0) s_sel = 0;
    const uint64_t n = q->n_rows;
    const uint64_t tile = L->tile;
    uint32_t my_sel = 0;
    for (uint64_t t0 = (uint64_t)blockIdx.x * tile; t0 < n; t0 += (uint64_t)gridDim.x * tile) {
        if (tid == 0) s_stop = (int)__hip_atomic_load(&L->t.ctrl[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __syncthreads();
        const int stop = s_stop;
        __syncthreads();
        if (stop) break;
        const uint64_t tend = min(t0 + tile, n);
        for (uint64_t r = t0 + tid; r < tend; r += bs) {
            MemRow row{q, r};
            if (!eval_pred(q, row)) continue;
            ++my_sel;
            Key k = load_key(q, row);
            uint64_t h = k.special ? 0 : hash_key(k.lo, k.hi);
            int64_t g = g_find128(L->t, k, h);
            if (g >= 0) add_row(gs, q, (uint64_t)g, row);
        }
    }
    atomicAdd(&s_sel, my_sel);
    __syncthreads();
    if (tid == 0) atomicAdd(&L->t.ctrl[4], s_sel);
}

// Key of global-table slot `slot` as (lo, hi, special): the two side slots
// past the capacity hold the all-ones key and the null key.
__device__ __forceinline__ void slot_key(const QueryDesc *__restrict__ q, const GTable &t, uint64_t slot,
                                         uint64_t &lo, uint64_t &hi, int &special) {
    const uint64_t cap = t.mask + 1;
    lo = 0;
    hi = 0;
    special = 0;
    if (slot == cap) { special = 1; return; }
    if (slot == cap + 1) { special = 2; return; }
    lo = t.klo[slot];
    if (q->key_mode == KM_LANE128) hi = t.khi[slot];
}

// One thread per output group; waves of 64 consecutive groups pack validity
// (and Bool keys) through ballots.  Resets every slot it reads.
__global__ __launch_bounds__(256) void fgb_finalize(const Launch *__restrict__ L, int wide) {
    const QueryDesc *__restrict__ q = &L->q;
    const GTable &t = L->t;
    const Layout &lay = q->layout;
    GReader rd{t.arr};
    uint64_t ng = *(volatile uint32_t *)&t.ctrl[0];
    if (ng > t.list_cap) ng = t.list_cap;
    const uint64_t nr = (ng + 63) & ~63ull;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    const int lane = threadIdx.x & 63;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nr; i += stride) {
        const bool live = i < ng;
        const uint64_t slot = live ? t.list[i] : 0;
        uint64_t klo = 0, khi = 0;
        int special = 0;
        if (live) slot_key(q, t, slot, klo, khi, special);
        for (int o = 0; o < L->n_out; ++o) {
            const OutCol &oc = L->out[o];
            uint64_t bits = 0;
            bool valid = false;
            if (live) {
                if (oc.src == OS_KEY) key_field(q, oc.idx, klo, khi, special, bits, valid);
                else agg_value(q, oc, rd, slot, bits, valid);
                if (oc.tag != PMH_BOOL || oc.src != OS_KEY) store_tag(oc.data, i, oc.tag, bits);
            }
            if (oc.valid) {
                unsigned long long m = __ballot(live && valid);
                if (lane == 0) oc.valid[i >> 6] = m;
            }
            if (oc.src == OS_KEY && oc.tag == PMH_BOOL) {
                unsigned long long m = __ballot(live && (bits & 1));
                if (lane == 0) ((uint64_t *)oc.data)[i >> 6] = m;
            }
        }
        if (live) {
            // leave the slot clean for the next call
            t.klo[slot] = kEmpty;
            if (wide) { t.khi[slot] = kEmpty; t.kstate[slot] = 0; }
            for (int a = 0; a < lay.n_arrs; ++a) {
                int op = lay.init[a];
                if (mo_size(op) == 4) ((uint32_t *)t.arr[a])[slot] = (uint32_t)mo_identity(op);
                else ((uint64_t *)t.arr[a])[slot] = mo_identity(op);
            }
        }
    }
}

// Zeroes the control words once every kernel that reads them has finished.
__global__ void fgb_clear_ctrl(const Launch *__restrict__ L) {
    if (threadIdx.x < 8) L->t.ctrl[threadIdx.x] = 0;
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
namespace {

struct TableState {
    uint64_t cap = 0;
    std::string sig;  // layout signature the clean table was initialised for
    bool clean = false;
};
TableState g_table;

// g_table describes the "gb:table" slab; an allocation-free call's table is
// in the scratch array, starts re-initialised and leaves g_table unchanged.
struct TableScratchGuard {
    TableState saved;
    bool on;
    explicit TableScratchGuard(bool scratch) : saved(g_table), on(scratch) {
        if (on) g_table.clean = false;
    }
    ~TableScratchGuard() {
        if (on) g_table = saved;
    }
};

int agg_vkind(int tag) {
    if (tag == PMH_F64 || tag == PMH_F32) return VK_F;
    if (tag == PMH_I64) return VK_I64;
    if (tag == PMH_I32) return VK_I32;
    return -1;
}

const char *op_name(int op) {
    static const char *n[] = {"Sum", "Mean", "Count", "Min", "Max", "Len"};
    return n[op];
}

uint64_t next_pow2(uint64_t x) {
    uint64_t p = 1;
    while (p < x) p <<= 1;
    return p;
}

struct Lowered {
    std::vector<std::string> needed;  // column names in QueryDesc.cols order
    std::function<int(const std::string &)> col_index;
};

int lower_pred(const PlanPredNode &n, QueryDesc &q, const std::vector<Column> &cols,
               const std::function<int(const std::string &)> &col_index, bool want_bool) {
    auto push = [&](const PredNode &pn) {
        if (q.n_pred >= kMaxPredNodes) fail(PMH_E_NOTIMPL, "polars_metal: predicate has too many nodes");
        q.pred[q.n_pred] = pn;
        return q.n_pred++;
    };
    PredNode pn{};
    if (n.kind == "Column") {
        int ci = col_index(n.name);
        if (cols[ci].tag != PMH_BOOL)
            fail(PMH_E_VALUE, "polars_metal: predicate column " + n.name + " used as a boolean is not Bool");
        pn.kind = PK_COLBOOL;
        pn.lhs = ci;
        return push(pn);
    }
    if (n.kind == "LiteralBool") {
        pn.kind = PK_CONST;
        pn.cval = n.bval;
        return push(pn);
    }
    if (n.kind == "And" || n.kind == "Or") {
        int l = lower_pred(*n.lhs, q, cols, col_index, true);
        int r = lower_pred(*n.rhs, q, cols, col_index, true);
        pn.kind = n.kind == "And" ? PK_AND : PK_OR;
        pn.lhs = l;
        pn.rhs = r;
        return push(pn);
    }
    if (n.kind == "Compare") {
        static const char *ops[] = {"Eq", "Ne", "Lt", "Le", "Gt", "Ge"};
        pn.kind = PK_CMP;
        for (int i = 0; i < 6; ++i)
            if (n.op == ops[i]) pn.op = i;
        // operand dtype: explicit, else F64 if any side is float
        auto side_is_f = [&](const PlanPredNode &s) -> int {
            if (s.kind == "LiteralF64") return 1;
            if (s.kind == "LiteralI64") return 0;
            if (s.kind == "Column") {
                int ci = col_index(s.name);
                return (cols[ci].tag == PMH_F64 || cols[ci].tag == PMH_F32) ? 1 : 0;
            }
            fail(PMH_E_NOTIMPL, "polars_metal: Compare operand kind " + s.kind + " is not supported");
        };
        int lf = side_is_f(*n.lhs), rf = side_is_f(*n.rhs);
        bool is_f = lf || rf;
        if (!n.dtype.empty()) {
            if (n.dtype == "F64" || n.dtype == "F32") is_f = true;
            else if (n.dtype == "I64" || tag_from_wire(n.dtype) >= 0) {
                if (lf || rf) fail(PMH_E_VALUE, "polars_metal: I64 compare with a float operand");
                is_f = false;
            }
        }
        pn.is_f64 = is_f;
        auto side = [&](const PlanPredNode &s, int32_t &kind, int32_t &col, int64_t &iv, double &fv) {
            if (s.kind == "Column") {
                int ci = col_index(s.name);
                int tag = cols[ci].tag;
                if (tag == PMH_BOOL || tag == PMH_UTF8)
                    fail(PMH_E_VALUE, "polars_metal: compare on a " + std::string(tag_name(tag)) + " column");
                kind = OK_COL;
                col = ci;
            } else {
                kind = OK_LIT;
                iv = s.ival;
                fv = s.lit_is_int ? (double)s.ival : s.fval;
            }
        };
        side(*n.lhs, pn.lkind, pn.lhs, pn.li, pn.lf);
        side(*n.rhs, pn.rkind, pn.rhs, pn.ri, pn.rf);
        if (pn.lkind == OK_LIT && pn.rkind == OK_LIT)
            fail(PMH_E_VALUE, "polars_metal: Compare needs at least one Column operand");
        return push(pn);
    }
    (void)want_bool;
    fail(PMH_E_VALUE, "polars_metal: predicate node " + n.kind + " does not yield a boolean");
}

}  // namespace

void collect_pred_cols(const PlanPredNode &n, std::vector<std::string> &out) {
    if (n.kind == "Column") out.push_back(n.name);
    if (n.lhs) collect_pred_cols(*n.lhs, out);
    if (n.rhs) collect_pred_cols(*n.rhs, out);
}

void prepare_query(Prep &P, const GroupByPlan &plan, uint64_t n_rows, const pmh_column *cols, uint32_t n_cols) {
    if (n_rows >= (1ull << 31)) fail(PMH_E_VALUE, "polars_metal: row count must be below 2^31");
    if (plan.aggs.empty() && plan.keys.empty()) fail(PMH_E_VALUE, "polars_metal: plan parse error: empty plan");
    // ---- columns referenced by the plan
    std::vector<std::string> needed;
    for (auto &k : plan.keys) needed.push_back(k.first);
    for (auto &a : plan.aggs) {
        if (a.op == AGG_LEN) continue;
        if (a.expr) collect_expr_cols(*a.expr, needed);
        else needed.push_back(a.input_col);
    }
    if (plan.predicate) collect_pred_cols(*plan.predicate, needed);
    for (auto &c : plan.extra_cols) needed.push_back(c);
    // dtype checks before any device work (the same failures the reference
    // reports from its parser / agg_kind_for, legacy.rs:22-48)
    auto find_host = [&](const std::string &nm) -> const pmh_column * {
        for (uint32_t i = 0; i < n_cols; ++i)
            if (cols[i].name && nm == cols[i].name) return &cols[i];
        fail(PMH_E_KEY, "polars_metal: column not found: " + nm);
    };
    for (auto &k : plan.keys) {
        const pmh_column *c = find_host(k.first);
        int t = c->tag == PMH_UTF8 ? PMH_U32 : c->tag;
        int pt = k.second == PMH_UTF8 ? PMH_U32 : k.second;
        if (t != pt) fail(PMH_E_VALUE, "polars_metal: key " + k.first + " dtype mismatch between plan and column");
        if (tag_size(c->tag) < 0) fail(PMH_E_NOTIMPL, "polars_metal: unsupported key dtype");
    }
    // an expression aggregate folds a column of the expression's dtype
    std::function<int(const std::string &)> host_tag = [&](const std::string &nm) { return find_host(nm)->tag; };
    for (auto &a : plan.aggs) {
        if (a.op == AGG_LEN) continue;
        const int tag = a.expr ? expr_result_tag(*a.expr, host_tag) : find_host(a.input_col)->tag;
        if (agg_vkind(tag) < 0)
            fail(PMH_E_VALUE, std::string("polars_metal: unsupported (agg_op, dtype) combination: ") + op_name(a.op) +
                                  " / " + tag_name(tag));
    }
    for (auto &nm : needed) check_column(*find_host(nm), n_rows, nm);
    rt();
    std::vector<Column> dcols = stage_columns(cols, n_cols, n_rows, needed);
    // expression aggregates: each distinct tree evaluated once into a
    // device column
    std::vector<std::string> agg_input(plan.aggs.size());
    for (size_t ai = 0; ai < plan.aggs.size(); ++ai) {
        const PlanAgg &a = plan.aggs[ai];
        agg_input[ai] = a.input_col;
        if (a.op == AGG_LEN || !a.expr) continue;
        for (size_t aj = 0; aj < ai && agg_input[ai] == a.input_col; ++aj)
            if (plan.aggs[aj].expr && plan.aggs[aj].op != AGG_LEN && expr_equal(*plan.aggs[aj].expr, *a.expr))
                agg_input[ai] = agg_input[aj];
        if (agg_input[ai] != a.input_col) continue;
        const std::string nm = "expr:" + std::to_string(ai);
        dcols.push_back(materialize_expr(*a.expr, dcols, n_rows, nm));
        agg_input[ai] = nm;
    }
    auto col_index = [&](const std::string &nm) -> int {
        for (size_t i = 0; i < dcols.size(); ++i)
            if (dcols[i].name == nm) return (int)i;
        fail(PMH_E_KEY, "polars_metal: column not found: " + nm);
    };
    if ((int)dcols.size() > kMaxCols) fail(PMH_E_NOTIMPL, "polars_metal: too many columns in one query");
    P.dcols = dcols;

    QueryDesc &q = P.q;
    memset(&q, 0, sizeof(q));
    P.n_rows = n_rows;
    q.n_rows = n_rows;
    q.n_cols = (int)dcols.size();
    for (size_t i = 0; i < dcols.size(); ++i) {
        q.cols[i].data = dcols[i].d_data;
        q.cols[i].valid = dcols[i].d_valid;
        q.cols[i].tag = dcols[i].tag;
    }
    if (plan.predicate) {
        lower_pred(*plan.predicate, q, dcols, col_index, true);
        const PredNode &root = q.pred[q.n_pred - 1];
        if (root.kind != PK_CMP && root.kind != PK_AND && root.kind != PK_OR && root.kind != PK_COLBOOL &&
            root.kind != PK_CONST)
            fail(PMH_E_VALUE, "polars_metal: predicate does not yield a boolean");
    }
    // keys (groupby.rs:155-302 layout)
    q.n_keys = (int)plan.keys.size();
    if (q.n_keys > kMaxKeys) fail(PMH_E_VALUE, "polars_metal: too many key columns");
    int off = 0;
    for (int i = 0; i < q.n_keys; ++i) {
        int ci = col_index(plan.keys[i].first);
        KeyField &f = q.keys[i];
        f.col = ci;
        f.tag = dcols[ci].tag;
        f.bits = tag_data_bits(f.tag);
        f.null_off = off;
        f.data_off = off + 1;
        off += 1 + f.bits;
    }
    q.total_bits = off;
    if (off > 128)
        fail(PMH_E_VALUE, "polars_metal: dispatch_groupby failed: composite key width " + std::to_string(off) +
                              " bits exceeds 128-bit budget");
    if (q.n_keys == 0) q.key_mode = KM_NONE;
    else if (q.n_keys == 1 && q.keys[0].bits == 64) q.key_mode = KM_RAW64;
    else if (off <= 64) { q.key_mode = KM_LANE64; q.allones_possible = off == 64; }
    else { q.key_mode = KM_LANE128; q.allones_possible = off == 128; }
    const bool wide = q.key_mode == KM_LANE128;
    P.wide = wide;

    // values and state layout
    Layout &lay = q.layout;
    auto add_arr = [&](int op) {
        if (lay.n_arrs >= kMaxArrs) fail(PMH_E_NOTIMPL, "polars_metal: too many aggregations");
        lay.init[lay.n_arrs] = op;
        lay.esz[lay.n_arrs] = mo_size(op);
        return lay.n_arrs++;
    };
    lay.a_len = add_arr(MO_ADD32);
    std::vector<int> agg_val(plan.aggs.size(), -1);
    for (size_t ai = 0; ai < plan.aggs.size(); ++ai) {
        const PlanAgg &a = plan.aggs[ai];
        if (a.op == AGG_LEN) continue;
        int ci = col_index(agg_input[ai]);
        int vj = -1;
        for (int j = 0; j < q.n_vals; ++j)
            if (q.vals[j].col == ci) vj = j;
        if (vj < 0) {
            if (q.n_vals >= kMaxVals) fail(PMH_E_NOTIMPL, "polars_metal: too many aggregated columns");
            vj = q.n_vals++;
            q.vals[vj].col = ci;
            q.vals[vj].tag = dcols[ci].tag;
            q.vals[vj].kind = agg_vkind(dcols[ci].tag);
        }
        if (a.op == AGG_SUM || a.op == AGG_MEAN) q.vals[vj].need |= NEED_SUM;
        if (a.op == AGG_MIN || a.op == AGG_MAX) q.vals[vj].need |= NEED_MINMAX;
        agg_val[ai] = vj;
    }
    for (int j = 0; j < q.n_vals; ++j) {
        ValSpec &vs = q.vals[j];
        vs.a_cnt = add_arr(MO_ADD32);
        if (vs.need & NEED_SUM) {
            if (vs.kind == VK_F) {
                // five consecutive limb arrays, allocated back to back
                vs.a_limb = add_arr(MO_ADD64);
                for (int k = 1; k < kLimbs; ++k) add_arr(MO_ADD64);
                vs.a_res = add_arr(MO_ADDF64);
            } else {
                vs.a_sum = add_arr(MO_ADD64);
            }
        }
        if (vs.need & NEED_MINMAX) {
            if (vs.kind == VK_I32) { vs.a_min = add_arr(MO_MIN32); vs.a_max = add_arr(MO_MAX32); }
            else { vs.a_min = add_arr(MO_MIN64); vs.a_max = add_arr(MO_MAX64); }
            if (vs.kind == VK_F) vs.a_nan = add_arr(MO_OR32);
        }
    }
    // signature of the layout (a clean table is reusable only for the same one)
    std::string sig = std::to_string(q.key_mode) + ":";
    for (int a = 0; a < lay.n_arrs; ++a) sig += std::to_string(lay.init[a]) + ",";
    // limb k > 0 of an exact accumulator: addressed from limb 0
    std::vector<int> &limb_k = P.limb_k, &limb_base = P.limb_base;
    limb_k.assign(lay.n_arrs, 0);
    limb_base.assign(lay.n_arrs, -1);
    for (int j = 0; j < q.n_vals; ++j)
        if (q.vals[j].kind == VK_F && (q.vals[j].need & NEED_SUM))
            for (int k = 0; k < kLimbs; ++k) { limb_k[q.vals[j].a_limb + k] = k; limb_base[q.vals[j].a_limb + k] = q.vals[j].a_limb; }

    // ---- outputs: keys then aggs
    int n_out = q.n_keys + (int)plan.aggs.size();
    if (n_out > kOutMax) fail(PMH_E_NOTIMPL, "polars_metal: too many output columns");
    P.n_out = n_out;
    std::vector<std::string> &out_names = P.names;
    out_names.clear();
    memset(P.outs, 0, sizeof(P.outs));
    for (int i = 0; i < q.n_keys; ++i) {
        OutCol &oc = P.outs[i];
        oc.src = OS_KEY;
        oc.idx = i;
        oc.tag = dcols[q.keys[i].col].tag;
        out_names.push_back(plan.keys[i].first);
    }
    for (size_t ai = 0; ai < plan.aggs.size(); ++ai) {
        const PlanAgg &a = plan.aggs[ai];
        OutCol &oc = P.outs[q.n_keys + ai];
        oc.src = OS_AGG;
        oc.op = a.op;
        oc.idx = agg_val[ai];
        if (a.op == AGG_LEN || a.op == AGG_COUNT) oc.tag = PMH_U32;
        else {
            int vt = q.vals[agg_val[ai]].tag;
            if (a.op == AGG_MEAN) oc.tag = vt == PMH_F32 ? PMH_F32 : PMH_F64;  // Polars: mean of ints is Float64
            else oc.tag = vt;
        }
        out_names.push_back(a.alias);
    }
    P.sig = sig;
}

pmh_result *run_hash_global(Prep &P, const GroupByPlan &plan, uint32_t flags) {
    Runtime &R = rt();
    const TableScratchGuard table_guard(R.scratch_mode);
    const uint64_t n_rows = P.n_rows;
    const bool wide = P.wide;
    const int n_out = P.n_out;
    const std::vector<std::string> &out_names = P.names;
    const std::vector<int> &limb_k = P.limb_k, &limb_base = P.limb_base;
    const std::string &sig = P.sig;
    auto Lh = std::make_unique<Launch>();
    memset(Lh.get(), 0, sizeof(Launch));
    Launch &L = *Lh;
    L.q = P.q;
    QueryDesc &q = L.q;
    Layout &lay = q.layout;
    L.n_out = n_out;
    for (int i = 0; i < n_out; ++i) L.out[i] = P.outs[i];
    // ---- sizing
    const int cus = R.prop.multiProcessorCount > 0 ? R.prop.multiProcessorCount : 256;
    uint64_t hint = plan.opts.capacity_hint;
    uint64_t cap;
    if (R.scratch_mode) hint = std::max<uint64_t>(n_rows, 1);  // every row its own group fits: no resize
    if (hint > 0) cap = next_pow2(std::max<uint64_t>((uint64_t)(hint / 0.5), 64));
    else cap = next_pow2(std::min<uint64_t>(std::max<uint64_t>(n_rows / 64, 4096), 1ull << 21));
    const uint64_t initial_cap = cap;

    // LDS table size: largest power of two that fits the per-block budget
    size_t per_slot = 8;
    for (int a = 0; a < lay.n_arrs; ++a) per_slot += lay.esz[a];
    size_t lds_budget = std::min<size_t>(R.prop.sharedMemPerBlock ? R.prop.sharedMemPerBlock : 65536, 160 * 1024);
    lds_budget -= 64;
    uint32_t CL = 0;
    if (!wide && plan.opts.lds_slots != 0) {
        uint64_t want = plan.opts.lds_slots > 0 ? (uint64_t)plan.opts.lds_slots : (1u << 16);
        CL = 1;
        while ((uint64_t)CL * 2 <= want && (CL * 2 + 2) * per_slot <= lds_budget) CL *= 2;
        if (CL < 16) CL = 0;
    }
    size_t lds_bytes = 0;
    if (CL) {
        size_t o = 0;
        for (int a = 0; a < lay.n_arrs; ++a) {
            L.lds_off[a] = (uint32_t)o;  // arrays back to back: limbs are nslots apart
            o += (size_t)(CL + 2) * lay.esz[a];
        }
        lds_bytes = (size_t)(CL + 2) * 8 + o;
    }
    L.lds_slots = CL;
    int block = plan.opts.block > 0 ? (int)plan.opts.block : (wide ? 256 : 1024);
    if (block % 64 || block > 1024) fail(PMH_E_VALUE, "polars_metal: options.block must be a multiple of 64, <= 1024");
    uint64_t tile = plan.opts.tile > 0 ? (uint64_t)plan.opts.tile : (uint64_t)block * 8;
    L.tile = tile;
    int grid = plan.opts.grid > 0 ? (int)plan.opts.grid : 0;
    if (!grid) {
        int per_cu = 1;
        if (!wide && lds_bytes) per_cu = std::max<int>(1, (int)((160 * 1024) / (lds_bytes + 1024)));
        per_cu = std::min(per_cu, std::max(1, 2048 / block));
        if (wide) per_cu = 8;
        grid = cus * per_cu;
        uint64_t tiles = ceil_div(std::max<uint64_t>(n_rows, 1), tile);
        if ((uint64_t)grid > tiles) grid = (int)tiles;
    }
    const bool cached = q.n_cols <= kCacheCols;
    const void *narrow_fn = cached ? (const void *)fgb_main_narrow<4, true> : (const void *)fgb_main_narrow<1, false>;
    if (lds_bytes > 65536)
        PMH_HIP(hipFuncSetAttribute(narrow_fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes));

    // ---- passes (plain method: resize x4 on overflow)
    Launch *dL = (Launch *)slab("gb:launch", sizeof(Launch));
    uint32_t ctrl_h[8];
    int passes = 0;
    PooledEvent pe0, pe1;
    const hipEvent_t ev0 = pe0.ev, ev1 = pe1.ev;
    PMH_HIP(hipEventRecord(ev0, R.stream));
    for (;;) {
        ++passes;
        GTable &t = L.t;
        const uint64_t nslots = cap + 2;
        t.mask = cap - 1;
        t.limit = (uint64_t)(cap * 0.7);
        t.list_cap = nslots;
        // table memory: keys, [khi, kstate], list, ctrl, state arrays
        size_t bytes = nslots * 8 + (wide ? nslots * 12 : 0) + nslots * 4 + 64;
        for (int a = 0; a < lay.n_arrs; ++a) bytes += ((nslots * lay.esz[a]) + 255) & ~size_t(255);
        uint8_t *base = (uint8_t *)slab("gb:table", bytes + 4096);
        if (g_table.cap != cap || g_table.sig != sig || !g_table.clean) g_table.clean = false;
        size_t o = 0;
        auto carve = [&](size_t b) { uint8_t *p = base + o; o += (b + 255) & ~size_t(255); return p; };
        t.klo = (uint64_t *)carve(nslots * 8);
        if (wide) { t.khi = (uint64_t *)carve(nslots * 8); t.kstate = (uint32_t *)carve(nslots * 4); }
        t.list = (uint32_t *)carve(nslots * 4);
        t.ctrl = (uint32_t *)carve(64);
        for (int a = 0; a < lay.n_arrs; ++a) {
            if (limb_base[a] >= 0 && limb_k[a] > 0) t.arr[a] = (uint8_t *)t.arr[limb_base[a]] + (size_t)limb_k[a] * nslots * 8;
            else if (limb_base[a] >= 0) t.arr[a] = carve(nslots * 8 * kLimbs);
            else t.arr[a] = carve(nslots * lay.esz[a]);
        }
        // outputs (sized for the largest group count this pass can accept)
        const uint64_t maxg = nslots;
        const uint64_t vwords = ceil_div(maxg, 64) + 1;
        for (int oi = 0; oi < n_out; ++oi) {
            OutCol &oc = L.out[oi];
            size_t esz = tag_size(oc.tag);
            size_t db = oc.tag == PMH_BOOL ? vwords * 8 : maxg * esz;
            oc.data = slab("gb:out" + std::to_string(oi), db + 64);
            bool need_valid = !(oc.src == OS_AGG && (oc.op == AGG_LEN || oc.op == AGG_COUNT));
            oc.valid = need_valid ? (uint64_t *)slab("gb:outv" + std::to_string(oi), vwords * 8 + 64) : nullptr;
        }
        PMH_HIP(hipMemcpyAsync(dL, &L, sizeof(Launch), hipMemcpyHostToDevice, R.stream));
        if (!g_table.clean) {
            KernelTimer kt("fgb_init_table");
            int ig = (int)std::min<uint64_t>(ceil_div(nslots, 256), (uint64_t)cus * 16);
            hipLaunchKernelGGL(fgb_init_table, dim3(ig), dim3(256), 0, R.stream, dL, nslots, wide ? 1 : 0);
            PMH_HIP(hipGetLastError());
            g_table.cap = cap;
            g_table.sig = sig;
            g_table.clean = true;
        }
        {
            KernelTimer kt(wide ? "fgb_main_wide" : "fgb_main_narrow");
            if (wide)
                hipLaunchKernelGGL(fgb_main_wide, dim3(grid), dim3(block), 0, R.stream, dL);
            else if (cached)
                hipLaunchKernelGGL((fgb_main_narrow<4, true>), dim3(grid), dim3(block), lds_bytes, R.stream, dL);
            else
                hipLaunchKernelGGL((fgb_main_narrow<1, false>), dim3(grid), dim3(block), lds_bytes, R.stream, dL);
            PMH_HIP(hipGetLastError());
        }
        PMH_HIP(hipMemcpyAsync(ctrl_h, t.ctrl, sizeof(ctrl_h), hipMemcpyDeviceToHost, R.stream));
        PMH_HIP(hipStreamSynchronize(R.stream));
        if (ctrl_h[1] == 0) {
            uint64_t fg = std::max<uint64_t>(1, std::min<uint64_t>(ceil_div(ctrl_h[0], 256), (uint64_t)cus * 8));
            {
                KernelTimer kt("fgb_finalize");
                hipLaunchKernelGGL(fgb_finalize, dim3(fg), dim3(256), 0, R.stream, dL, wide ? 1 : 0);
                PMH_HIP(hipGetLastError());
            }
            hipLaunchKernelGGL(fgb_clear_ctrl, dim3(1), dim3(64), 0, R.stream, dL);
            PMH_HIP(hipGetLastError());
            break;
        }
        // overflow: the table is dirty; grow and re-run
        g_table.clean = false;
        if (cap >= (1ull << 31)) fail(PMH_E_COMPUTE, "polars_metal: hash table capacity exhausted");
        cap *= 4;
    }
    PMH_HIP(hipEventRecord(ev1, R.stream));
    PMH_HIP(hipEventSynchronize(ev1));
    float dev_ms = 0.f;
    (void)hipEventElapsedTime(&dev_ms, ev0, ev1);

    uint64_t ng = std::min<uint64_t>(ctrl_h[0], cap + 2);
    pmh_result *res = new_result(n_out);
    res->n_rows = ng;
    res->stats.method = PMH_METHOD_HASH;
    res->stats.n_passes = passes;
    res->stats.initial_capacity = initial_cap;
    res->stats.final_capacity = cap;
    res->stats.n_groups = ng;
    res->stats.n_selected = ctrl_h[4];
    res->stats.device_ms = dev_ms;
    const uint64_t vbytes = bitmap_bytes_padded(ng);
    for (int oi = 0; oi < n_out; ++oi) {
        const OutCol &oc = L.out[oi];
        pmh_out_column &c = res->columns[oi];
        set_out_name(c, out_names[oi]);
        c.tag = oc.tag;
        size_t esz = tag_size(oc.tag);
        c.data_len = oc.tag == PMH_BOOL ? bitmap_bytes_padded(ng) : ng * esz;
        c.valid_len = vbytes;
        if (flags & PMH_OUT_DEVICE) {
            c.location = PMH_DEVICE;
            c.data = oc.data;
            c.valid = (uint8_t *)oc.valid;  // nullptr for count/len: all valid
        } else {
            c.location = PMH_HOST;
            c.data = malloc(c.data_len ? c.data_len : 1);
            c.valid = (uint8_t *)malloc(vbytes);
            if (c.data_len) download_host(c.data, oc.data, c.data_len);
            if (oc.valid) {
                download_host(c.valid, oc.valid, vbytes);
            } else {
                memset(c.valid, 0xFF, vbytes);  // counts / lens are never null (groupby/mod.rs:695-706)
            }
        }
    }
    // host outputs wait for their copies; device outputs were complete at ev1
    if (!(flags & PMH_OUT_DEVICE)) PMH_HIP(hipStreamSynchronize(R.stream));
    // Bool key data and bitmaps come back as whole 64-bit words; clear the
    // bits past the last group so padding is zero like pack_valid_bitmap's.
    if (!(flags & PMH_OUT_DEVICE)) {
        for (int oi = 0; oi < n_out; ++oi) {
            pmh_out_column &c = res->columns[oi];
            const OutCol &oc = L.out[oi];
            auto clear_tail = [&](uint8_t *b, uint64_t len) {
                for (uint64_t bit = ng; bit < len * 8; ++bit) b[bit >> 3] &= (uint8_t) ~(1u << (bit & 7));
            };
            if (oc.valid) clear_tail(c.valid, c.valid_len);
            if (oc.tag == PMH_BOOL && oc.src == OS_KEY) clear_tail((uint8_t *)c.data, c.data_len);
        }
    }
    collect_kernel_times();
    return res;
}

pmh_result *run_filter_groupby(const GroupByPlan &plan, uint64_t n_rows, const pmh_column *cols,
                               uint32_t n_cols, uint32_t flags) {
    Prep P;
    prepare_query(P, plan, n_rows, cols, n_cols);
    const std::string &m = plan.opts.method;
    if (m == "hash_global") return run_hash_global(P, plan, flags);
    if (m == "auto" || m == "hash") {
        // nullptr: the shape does not fit the partitioned engine's LDS plans,
        // or its fallback table overflowed; the global-table engine (which
        // resizes) then runs the query
        pmh_result *r = run_hash_partitioned(P, plan, flags);
        if (!r && P.narrow_miss) {
            // a key outside the sampled narrowing ranges: narrowed from
            // every row this time
            P.narrow_miss = false;
            P.narrow_exact = true;
            r = run_hash_partitioned(P, plan, flags);
        }
        if (r) return r;
        if (P.route == 1) {  // the learned sizer chose the sort engine
            r = run_sort(P, plan, flags);
            r->stats.sizer_choice = 2;
            return r;
        }
        if (P.overflowed) {
            // sized for too few keys: once more with four times the
            // estimate (more partitions and a larger fallback table share)
            // before the global-table engine
            P.d_retry = std::min(4.0 * std::max(P.d_used, 1024.0), (double)n_rows + 1.0);
            P.overflowed = false;
            r = run_hash_partitioned(P, plan, flags);
            if (r) {
                r->stats.n_passes = 2;
                return r;
            }
        }
    }
    if (m == "sort") return run_sort(P, plan, flags);
    return run_hash_global(P, plan, flags);
}

// The partitioned engine's sample sketch alone: the learned sizer's
// features and the sketch fields of the stats, or (info) the whole sketch
// with the installed model's outputs.
void run_sketch(GroupByPlan &plan, uint64_t n_rows, const pmh_column *cols, uint32_t n_cols, float *feat,
                pmh_exec_stats *stats, pmh_sketch_info *info) {
    plan.opts.method = "hash";
    plan.opts.sketch_only = true;
    Prep P;
    prepare_query(P, plan, n_rows, cols, n_cols);
    P.sketch_out = info;
    pmh_result *r = run_hash_partitioned(P, plan, 0);
    if (!r) fail(PMH_E_NOTIMPL, "polars_metal: no sample sketch for this query shape");
    if (stats) *stats = r->stats;
    pmh_result_free(r);
    if (feat) memcpy(feat, P.feat, sizeof(P.feat));
}

void lower_predicate_into(QueryDesc &q, const GroupByPlan &plan, const Prep &P) {
    q.n_pred = 0;
    if (!plan.predicate) return;
    auto col_index = [&](const std::string &nm) -> int {
        for (size_t i = 0; i < P.dcols.size(); ++i)
            if (P.dcols[i].name == nm) return (int)i;
        fail(PMH_E_KEY, "polars_metal: column not found: " + nm);
    };
    lower_pred(*plan.predicate, q, P.dcols, col_index, true);
}

}  // namespace pmh
