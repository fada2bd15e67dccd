// pmh_internal.h -- types shared by the host orchestration and the gfx950
// kernels of the filtered group-by backend.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <functional>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "pmh.h"

namespace pmh {

// ---------------------------------------------------------------------------
// errors: C++ exceptions inside the library, status codes at the C ABI
// ---------------------------------------------------------------------------
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};
[[noreturn]] inline void fail(int code, const std::string &msg) { throw Error(code, msg); }
void set_last_error(const std::string &msg);

#define PMH_HIP(expr)                                                                   \
    do {                                                                                \
        hipError_t _e = (expr);                                                         \
        if (_e != hipSuccess)                                                           \
            ::pmh::fail(PMH_E_COMPUTE, std::string("polars_metal: HIP error in ") + #expr + \
                                           ": " + hipGetErrorString(_e));               \
    } while (0)

// ---------------------------------------------------------------------------
// limits
// ---------------------------------------------------------------------------
constexpr int kMaxPredNodes = 32;
constexpr int kMaxKeys = 16;
constexpr int kMaxVals = 8;   // distinct aggregated input columns per query
constexpr int kMaxCols = 32;  // distinct columns referenced by one query
constexpr int kMaxArrs = 96;  // per-group state arrays
constexpr int kLimbs = 5;     // exact f64 sum: 5 x 32-bit chunks at scale 2^-104
constexpr int kFixScale = 104;

// ---------------------------------------------------------------------------
// dtypes
// ---------------------------------------------------------------------------
__host__ __device__ inline int tag_size(int tag) {
    switch (tag) {
    case PMH_I64: case PMH_F64: case PMH_U64: return 8;
    case PMH_I32: case PMH_F32: case PMH_U32: case PMH_UTF8: return 4;
    case PMH_I16: case PMH_U16: return 2;
    case PMH_I8: case PMH_U8: return 1;
    case PMH_BOOL: return 0;  // bit-packed
    default: return -1;
    }
}
__host__ __device__ inline int tag_data_bits(int tag) {
    int s = tag_size(tag);
    return s == 0 ? 1 : s * 8;
}
const char *tag_name(int tag);
int tag_from_wire(const std::string &s);  // -1 when unknown

// ---------------------------------------------------------------------------
// device-side query description (uploaded once per call, read with scalar
// loads by every kernel)
// ---------------------------------------------------------------------------
struct DevCol {
    const void *data;
    const uint8_t *valid;  // nullptr: all valid
    int32_t tag;
    int32_t pad;
};

enum PredKind : int32_t { PK_COLBOOL = 0, PK_CMP = 1, PK_AND = 2, PK_OR = 3, PK_CONST = 4 };
enum OpndKind : int32_t { OK_COL = 0, OK_LIT = 1 };

// Predicate in topological order; node k may reference only nodes < k.  The
// result of node k lives in bit k of two per-row 32-bit masks, so evaluation
// needs no stack and no dynamically indexed registers.
struct PredNode {
    int32_t kind;
    int32_t op;       // pmh_cmp_op for PK_CMP
    int32_t is_f64;   // PK_CMP: compare as f64 (TotalOrd) instead of i64
    int32_t lkind, rkind;
    int32_t lhs, rhs; // column index (OK_COL / PK_COLBOOL) or child node index
    int32_t cval;     // PK_CONST: literal bool
    int64_t li, ri;
    double lf, rf;
};

enum KeyMode : int32_t {
    KM_NONE = 0,    // no key columns: one group
    KM_LANE64 = 1,  // encoded lane fits in 64 bits
    KM_RAW64 = 2,   // one 64-bit key column: raw bits, null and all-ones via side slots
    KM_LANE128 = 3  // encoded lane up to 128 bits
};

// A key column's field of the lane: its null bit at null_off (-1: the
// column holds no null, so no bit), its `bits` data bits at data_off.  With
// xform 0 the data bits are the value's raw bits; a narrowed lane (xform 1:
// raw bits as unsigned, 2: signed, sign-extended with the sign bit flipped)
// stores the value's image minus bias, the image of the column's smallest
// value, in as many bits as the column's range needs.
struct KeyField {
    int32_t col, tag, bits, null_off, data_off, xform;
    uint64_t bias;
};
enum KeyXform : int32_t { KX_RAW = 0, KX_UNSIGNED = 1, KX_SIGNED = 2 };
__host__ __device__ inline bool tag_signed(int tag) {
    return tag == PMH_I64 || tag == PMH_I32 || tag == PMH_I16 || tag == PMH_I8;
}

enum ValKind : int32_t { VK_F = 0, VK_I64 = 1, VK_I32 = 2 };
enum ValNeed : int32_t { NEED_SUM = 1, NEED_MINMAX = 2 };

struct ValSpec {
    int32_t col, tag, kind, need;
    int32_t a_cnt, a_limb, a_res, a_sum, a_min, a_max, a_nan, pad;
};

enum ArrInit : int32_t { AI_ZERO = 0, AI_I32MAX = 1, AI_I32MIN = 2, AI_I64MAX = 3, AI_I64MIN = 4 };

struct Layout {
    int32_t n_arrs;
    int32_t a_len;
    int32_t esz[kMaxArrs];
    int32_t init[kMaxArrs];
    int32_t limbs_contig;  // limb arrays of one value column are consecutive
};

// Expression program (pmh_expr.hip): nodes in postfix order; each node's
// value has dtype `tag`; a binary node converts both operands (of dtypes
// ta, tb) to `ctag` first and computes in it.
enum ExprKind : int32_t { EX_COL = 0, EX_LIT = 1, EX_ADD = 2, EX_SUB = 3, EX_MUL = 4, EX_DIV = 5 };
constexpr int kMaxExprNodes = 31;  // a binary tree four levels deep
constexpr int kMaxExprCols = 16;
struct ExprNode {
    int32_t kind, tag, ctag, ta, tb, col;
    uint64_t lit;  // EX_LIT: the literal's bits in dtype `tag`
};
struct ExprProg {
    int32_t n_nodes, n_cols;
    DevCol cols[kMaxExprCols];
    ExprNode nodes[kMaxExprNodes];
};

struct QueryDesc {
    uint64_t n_rows;
    int32_t n_cols;
    int32_t n_pred;  // 0: keep all rows
    DevCol cols[kMaxCols];
    PredNode pred[kMaxPredNodes];
    int32_t key_mode, n_keys, total_bits, allones_possible;
    KeyField keys[kMaxKeys];
    int32_t n_vals;
    int32_t narrow_check;  // the narrowed lane's ranges come from a sample: the scan flags keys outside them
    ValSpec vals[kMaxVals];
    Layout layout;
};

// ---------------------------------------------------------------------------
// host-side plan
// ---------------------------------------------------------------------------
enum AggOpHost { AGG_SUM, AGG_MEAN, AGG_COUNT, AGG_MIN, AGG_MAX, AGG_LEN };

// Parsed AggExpr (crates/polars-metal-core/src/plan/mod.rs:133-142): a
// column, an f64 or i64 literal, or Add/Sub/Mul/Div of two subtrees, at most
// four binary levels deep (plan/mod.rs:169-178).
struct PlanExpr {
    enum Kind { COLUMN, LIT_F64, LIT_I64, BINARY } kind = COLUMN;
    std::string name;  // COLUMN
    int op = 0;        // BINARY: ExprKind EX_ADD .. EX_DIV
    int64_t ival = 0;  // LIT_I64
    double fval = 0;   // LIT_F64
    std::shared_ptr<PlanExpr> lhs, rhs;
};
constexpr int kMaxExprDepth = 4;

struct PlanAgg {
    std::string input_col;
    int op;
    std::string alias;
    // kind "Expression": the per-row expression the reducer `op` folds
    std::shared_ptr<PlanExpr> expr;
};

// Every column an expression reads, in tree order.
void collect_expr_cols(const PlanExpr &e, std::vector<std::string> &out);
bool expr_equal(const PlanExpr &a, const PlanExpr &b);  // same tree, literals by bits

struct PlanPredNode {  // parsed PredicateAst
    std::string kind;
    std::string name, dtype, op;
    int64_t ival = 0;
    double fval = 0.0;
    bool bval = false;
    bool lit_is_int = false;
    std::unique_ptr<PlanPredNode> lhs, rhs;
};

struct Options {
    std::string method = "auto";  // auto | hash | hash_global | sort
    int64_t capacity_hint = 0;    // 0: plain-method default
    int64_t lds_slots = -1;       // -1: auto, 0: no block-local table
    int64_t block = 0;            // threads per block, 0: auto
    int64_t grid = 0;             // blocks, 0: auto
    int64_t tile = 0;             // rows per block per tile, 0: auto
    int64_t partitions = 0;       // partitioned engine: tail partitions (power of two), 0: auto
    int64_t tail_slots = 0;       // partitioned engine: per-partition LDS table slots, 0: auto
    int64_t sample = -1;          // partitioned engine: sampled rows for the hot set, -1: auto
    int64_t distinct_hint = 0;    // partitioned engine: expected distinct keys (a sizer's prediction), 0: sketch
    int64_t fast_scan = 1;        // partitioned engine: 1 use the specialised scan where the query fits it, 0 never
    int64_t scan_rows = 4;        // specialised scan: rows per lane per step (4 or 8)
    int64_t scan_blocks = 1;      // partitioned scan: resident blocks per CU (the hot table shares LDS)
    int64_t narrow_keys = 1;      // partitioned engine: narrow a lane wider than 64 bits to its key ranges, 1 from
                                  // a sample the specialised scan checks (else every row), 2 from every row, 0 never
    int64_t sort_buckets = 1;     // sort engine: 1 radix passes over the key's top bits, then LDS-sorted buckets
                                  // where that saves passes; 0 radix passes over the whole key
    int64_t sort_carry = 1;       // sort engine: 1 records carry the aggregated values (no gather through
                                  // the row in the fold) where the fast scan encodes them; 0 rows
    int64_t sort_sample_bits = 1;  // sort engine: 1 a large query without a predicate takes its varying key bits
                                   // from a sample (the encode checks every record), 0 a pass over every row
    int64_t sort_fold = 1;        // sort engine: 1 fold each LDS-sorted bucket's groups in place (no sorted
                                  // records written back), 0 write the bucket back and fold in a later pass
    int64_t sort_msd = 1;         // sort engine: 1 the top-bit passes (at most two) run most significant digit
                                  // first, unstable and merged per XCD; 0 stable LSD passes
    bool sketch_only = false;     // partitioned engine: stop after the sample sketch (feature extraction)
};

struct GroupByPlan {
    std::vector<std::pair<std::string, int>> keys;
    std::vector<PlanAgg> aggs;
    std::unique_ptr<PlanPredNode> predicate;
    Options opts;
    // columns staged beside the keys, aggregate inputs and predicate (a
    // batch template lists the columns its queries' own predicates read)
    std::vector<std::string> extra_cols;
};

// Appends every column a PredicateAst reads, in tree order.
void collect_pred_cols(const PlanPredNode &n, std::vector<std::string> &out);

GroupByPlan parse_groupby_plan(const std::string &json, bool allow_predicate);
std::string compute_lifting_plan_json(const std::string &json);

// ---------------------------------------------------------------------------
// device runtime
// ---------------------------------------------------------------------------
struct Runtime {
    int device = -1;
    hipStream_t stream = nullptr;
    hipDeviceProp_t prop{};
    std::map<std::string, std::pair<void *, size_t>> slabs;
    void *pinned = nullptr;
    size_t pinned_size = 0;
    bool profiling = false;
    std::map<std::string, std::pair<int64_t, double>> ktimes;
    std::vector<std::pair<std::string, std::pair<hipEvent_t, hipEvent_t>>> pending;
    std::vector<hipEvent_t> event_pool;
    // Allocation-free calls carve every slab from the library's static
    // device array `scratch` (its device address, resolved once at set-up)
    // and use the mailbox and events made at set-up.
    uint8_t *scratch = nullptr;
    size_t scratch_cap = 0;
    bool scratch_mode = false;
    size_t scratch_used = 0;
    std::map<std::string, std::pair<size_t, size_t>> scratch_slabs;  // name -> (offset, bytes)
};

Runtime &rt();  // initialises the device on first use (throws PMH_E_COMPUTE)
void *slab(const std::string &name, size_t bytes);  // grow-only named device buffer
void *pinned_host(size_t bytes);

// Bytes of the static scratch array, and the most rows an allocation-free
// call takes (its slabs at that size fit the array for every engine).
constexpr size_t kScratchBytes = PMH_NOALLOC_SCRATCH_BYTES;
constexpr uint64_t kNoAllocMaxRows = PMH_NOALLOC_MAX_ROWS;

// While alive, slabs come from the scratch array, bump-allocated by name
// (a name asked again with no more bytes gets its earlier range), and a
// request beyond the array fails with PMH_E_VALUE.  Restores slab() on exit.
struct ScratchScope {
    ScratchScope();
    ~ScratchScope();
};

// Copies an engine's device result into caller buffers (data, then the
// validity bitmap padded to 4 bytes, all ones where the result has none)
// and waits for the copies.  Sizes are checked by the caller.
void copy_outputs_to_caller(const pmh_result &r, pmh_out_column *outs, uint32_t n_outs);

// An event from the runtime's pool (made at set-up), returned on scope exit;
// with the pool empty a normal call makes one, an allocation-free call fails.
struct PooledEvent {
    PooledEvent();
    ~PooledEvent();
    PooledEvent(const PooledEvent &) = delete;
    PooledEvent &operator=(const PooledEvent &) = delete;
    hipEvent_t ev = nullptr;
};
// Launch bracketing for the device-time instrumentation.
struct KernelTimer {
    KernelTimer(const char *name);
    ~KernelTimer();
    std::string name;
    hipEvent_t a = nullptr, b = nullptr;
};
void collect_kernel_times();

// Column resolved for one call: device pointers plus the host descriptor.
struct Column {
    std::string name;
    int tag;
    const void *d_data;
    const uint8_t *d_valid;
    uint64_t n_rows;
};

// Upload (or pass through) the columns the plan references.
// Refuses a column whose value or validity buffer is shorter than n_rows
// need, or whose tag is unknown (PMH_E_VALUE), before any device work.
void check_column(const pmh_column &c, uint64_t n_rows, const std::string &nm);
std::vector<Column> stage_columns(const pmh_column *cols, uint32_t n_cols, uint64_t n_rows,
                                  const std::vector<std::string> &needed);

// Host bytes onto the device through the pinned staging ring, on the
// library's stream (pmh_ingest.hip).  Returns once src is no longer read;
// the DMA completes in stream order.
void upload_host(void *dst, const void *src, size_t bytes);
uint64_t h2d_bytes();  // bytes upload_host has moved, since the process began
// Device bytes into host memory (after the work queued on the stream): a
// copy of 1 MiB or more goes through the same ring and is complete on
// return; a smaller one completes by the next synchronisation of the stream.
void download_host(void *dst, const void *src, size_t bytes);
// The device column cache (pmh_ingest.hip).
void column_resident(int32_t tag, const pmh_chunk *chunks, uint32_t n_chunks, uint64_t stamp,
                     pmh_resident_column *out);
void column_release(uint64_t entry);
void column_evict(uint64_t entry);
void colcache_clear();
void colcache_set_capacity(uint64_t bytes);
void colcache_stats(pmh_colcache_info *out);

// Utf8 dictionaries on the device (pmh_utf8.hip).
pmh_utf8_dict *utf8_encode(const pmh_utf8_column &c);
std::vector<uint8_t> utf8_decode(const pmh_utf8_dict &d, const uint32_t *codes, const uint8_t *valid, uint64_t n,
                                 int location);
void utf8_dict_free(pmh_utf8_dict *d);
void utf8_dict_codes(const pmh_utf8_dict &d, const uint32_t **codes, const uint8_t **valid, uint64_t *n_rows,
                     uint64_t *n_distinct);

// result helpers
pmh_result *new_result(size_t n_cols);
void set_out_name(pmh_out_column &c, const std::string &name);

// Expression aggregates (pmh_expr.hip).  expr_result_tag: the dtype of the
// expression's value under Polars' arithmetic rules, given the dtype of each
// column (PMH_E_NOTIMPL for a shape whose rules are not restated here).
// materialize_expr: evaluates it for every row into a device column (slab
// `slab_name`); a row is null where any column it reads is null.
int expr_result_tag(const PlanExpr &e, const std::function<int(const std::string &)> &col_tag);
Column materialize_expr(const PlanExpr &e, const std::vector<Column> &cols, uint64_t n_rows,
                        const std::string &slab_name);

// group-by entry (pmh_groupby.hip)
pmh_result *run_filter_groupby(const GroupByPlan &plan, uint64_t n_rows, const pmh_column *cols,
                               uint32_t n_cols, uint32_t flags);

// sample sketch only: the learned sizer's features (pmh_groupby.hip)
void run_sketch(GroupByPlan &plan, uint64_t n_rows, const pmh_column *cols, uint32_t n_cols, float *feat,
                pmh_exec_stats *stats, pmh_sketch_info *info = nullptr);

// filter entries (pmh_filter.hip)
pmh_result *run_filter_compact(const uint8_t *pd, const uint8_t *pv, uint64_t n_rows,
                               const pmh_column *cols, uint32_t n_cols);
void run_cmp(int is_f64, int col_col, const void *lhs, const uint8_t *lv, const void *rhs,
             const uint8_t *rv, int64_t rhs_i, double rhs_f, int op, uint64_t n_rows,
             uint8_t *out_d, uint8_t *out_v);
void run_bool(int is_or, const uint8_t *ld, const uint8_t *lv, const uint8_t *rd,
              const uint8_t *rv, uint64_t n_rows, uint8_t *out_d, uint8_t *out_v);

inline uint64_t ceil_div(uint64_t a, uint64_t b) { return (a + b - 1) / b; }
inline uint64_t bitmap_bytes_padded(uint64_t n) {
    uint64_t b = ((n + 7) / 8 + 3) & ~uint64_t(3);
    return b < 4 ? 4 : b;
}

}  // namespace pmh
