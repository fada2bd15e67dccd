/*
 * pmh.h -- C ABI of the MI355X (gfx950) execution backend for polars-metal's
 * relational hot path: predicate filter, stream compaction, hash / u128
 * radix-sort group-by, fused sum/count/len/min/max/mean.
 *
 * Every entry point below replaces one function of the reference's PyO3
 * module `polars_metal._native` (crates/polars-metal-core/src/lib.rs:57-83);
 * the comment on each names the reference function it stands in for.  Plans
 * travel as JSON text with exactly the keys of the reference's Python dicts.
 *
 * Conventions
 *   - Status: every function returns PMH_OK (0) or a PMH_E_* code; the
 *     message is in pmh_last_error() (thread-local).  The codes map onto the
 *     reference's Python exception classes (see pmh_status).
 *   - Buffers: Arrow little-endian value buffers and LSB-first validity
 *     bitmaps.  A NULL validity pointer means "all valid".  Inputs may live
 *     in host memory or in device memory (pmh_column.location).
 *   - Output validity bitmaps are padded to a multiple of 4 bytes, minimum 4
 *     (crates/polars-metal-core/src/udf/common.rs:74-84).
 *   - Group order in results is unspecified; callers sort by key
 *     (crates/polars-metal-core/src/udf/groupby/mod.rs:720-746).
 *   - Row counts must be below 2^31; every entry returns PMH_E_VALUE for
 *     2^31 rows or more, before any device work.  The reference accepts up
 *     to u32::MAX (crates/polars-metal-kernels/src/filter.rs:73-77,
 *     groupby.rs:349-350); this backend narrows that by one bit because its
 *     kernels address rows with signed 32-bit offsets.
 *   - Calls are serialised per process; no entry point may be called from two
 *     host threads at once on the same device.
 */
#ifndef PMH_H
#define PMH_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes.  Mapping to the reference's exceptions:
 *   PMH_E_VALUE    -> ValueError        (plan parse, bad dtype/op, groupby/mod.rs:755-756)
 *   PMH_E_KEY      -> KeyError          (missing column, groupby/mod.rs:809-814)
 *   PMH_E_NOTIMPL  -> NotImplementedError (unsupported key dtype, groupby/mod.rs:952-954)
 *   PMH_E_RUNTIME  -> RuntimeError      (compaction failure, udf/compact.rs:97-101)
 *   PMH_E_COMPUTE  -> polars.exceptions.ComputeError (device failure, error.rs:19-33) */
typedef enum {
    PMH_OK = 0,
    PMH_E_VALUE = 1,
    PMH_E_KEY = 2,
    PMH_E_NOTIMPL = 3,
    PMH_E_RUNTIME = 4,
    PMH_E_COMPUTE = 5
} pmh_status;

/* Wire dtype tags (crates/polars-metal-core/src/plan/mod.rs:78-101).
 * Utf8 key columns arrive as u32 dictionary codes. */
typedef enum {
    PMH_I64 = 0, PMH_F64 = 1, PMH_BOOL = 2, PMH_I32 = 3, PMH_F32 = 4, PMH_I8 = 5,
    PMH_I16 = 6, PMH_U8 = 7, PMH_U16 = 8, PMH_U32 = 9, PMH_U64 = 10, PMH_UTF8 = 11
} pmh_dtype;

/* Comparison operators (crates/polars-metal-kernels/src/cmp.rs CompareOp). */
typedef enum { PMH_EQ = 0, PMH_NE = 1, PMH_LT = 2, PMH_LE = 3, PMH_GT = 4, PMH_GE = 5 } pmh_cmp_op;

typedef enum { PMH_HOST = 0, PMH_DEVICE = 1 } pmh_location;

/* One input column: the C form of the reference's
 * `(name, dtype_tag, data: bytes, valid: bytes)` tuple
 * (crates/polars-metal-core/src/udf/groupby/mod.rs:720-746). */
typedef struct pmh_column {
    const char *name;
    int32_t tag;        /* pmh_dtype */
    int32_t location;   /* pmh_location of data and valid */
    const void *data;
    uint64_t data_len;  /* bytes */
    const uint8_t *valid;
    uint64_t valid_len; /* bytes; ignored when valid == NULL */
} pmh_column;

/* One output column: the C form of the reference's output tuple
 * `(name, tag, data, valid)` (groupby/mod.rs:1109-1143). */
typedef struct pmh_out_column {
    const char *name;
    int32_t tag;
    int32_t location;
    void *data;
    uint64_t data_len;
    uint8_t *valid;
    uint64_t valid_len;
} pmh_out_column;

/* Group-by path selection recorded for each call (no reference
 * counterpart: the reference only logs debug lines,
 * python/polars_metal/_callback.py:22-125). */
typedef enum { PMH_METHOD_HASH = 0, PMH_METHOD_SORT = 1 } pmh_method;

typedef struct pmh_exec_stats {
    int32_t method;           /* pmh_method that produced the result */
    int32_t n_passes;         /* hash passes run (1 + resizes) */
    uint64_t initial_capacity;/* first hash-table capacity tried */
    uint64_t final_capacity;  /* capacity of the pass that succeeded */
    uint64_t n_groups;
    uint64_t n_selected;      /* rows that passed the predicate */
    double device_ms;         /* device time of the call (HIP events) */
    /* row-sample sketch of the partitioned hash engine (zeros otherwise) */
    uint32_t sample_rows;     /* sampled rows that passed the predicate */
    uint32_t sample_distinct; /* distinct keys among them */
    uint32_t sample_f1;       /* keys seen once */
    uint32_t sample_f2;       /* keys seen twice */
    uint32_t hot_keys;        /* keys folded in LDS by every scan block */
    uint32_t partitions;      /* tail partitions */
    uint64_t tail_rows;       /* selected rows not covered by the hot keys */
    double distinct_estimate; /* distinct keys the engine sized its partitions for */
    int32_t scan_kernel;      /* partitioned engine's row scan: 0 generic, 1 specialised */
    int32_t sizer_choice;     /* learned sizer under method "auto": 0 not used, 1 chose hash, 2 chose sort */
} pmh_exec_stats;

typedef struct pmh_result {
    uint64_t n_rows;          /* rows in every output column */
    uint32_t n_columns;
    pmh_out_column *columns;
    pmh_exec_stats stats;
} pmh_result;

/* Output placement flags for the fused entry. */
#define PMH_OUT_HOST 0u
/* Keep outputs in device memory owned by the library; they stay valid until
 * the next call on the device or pmh_result_free. */
#define PMH_OUT_DEVICE 1u

/* Last error message of the calling thread ("" if none). */
const char *pmh_last_error(void);

/* Replaces _native.version_string (crates/polars-metal-core/src/lib.rs:31-35). */
const char *pmh_version_string(void);

/* Replaces _native.device_name (crates/polars-metal-core/src/lib.rs:36-40).
 * Writes the device name into buf (NUL-terminated). */
int pmh_device_name(char *buf, size_t buf_len);

/* Replaces _native.execute_groupby
 * (crates/polars-metal-core/src/udf/groupby/mod.rs:747-1144).
 * plan_json: {"keys": [[name, tag], ...], "aggs": [{"kind": "Simple"|"Length",
 * "input_col": ..., "op": "Sum|Mean|Count|Min|Max|Len", "output_alias": ...}],
 * "options"?: {...}}.  Outputs: keys in `keys` order, then aggs in `aggs`
 * order (mod.rs:1109-1143).
 * An agg of kind "Expression" carries an AggExpr tree under "expr"
 * (Column / LiteralF64 / LiteralI64 / Binary Add|Sub|Mul|Div, at most four
 * binary levels; mod.rs:278-360, plan/mod.rs:133-178); the tree is evaluated
 * per row on the device with Polars' arithmetic dtypes and folded by "op",
 * where the reference evaluates it through Polars on the host first
 * (python/polars_metal/_udf.py:1164-1209).  Operand dtypes Polars types
 * outside the rules restated in pmh_expr.hip return PMH_E_NOTIMPL. */
int pmh_execute_groupby(const char *plan_json, uint64_t n_rows, const pmh_column *columns,
                        uint32_t n_columns, pmh_result **out);

/* Fused filter + group-by.  No single reference function: it replaces the
 * pair _udf._filter_via_polars (python/polars_metal/_udf.py:550-597, Polars
 * CPU compaction) + _native.execute_groupby (groupby/mod.rs:747) on the
 * GroupBy(Filter(Scan)) shape.  plan_json is the execute_groupby plan plus an
 * optional "predicate" in the reference's PredicateAst dict form
 * (crates/polars-metal-core/src/udf/predicate.rs:92-200).
 * flags: PMH_OUT_HOST or PMH_OUT_DEVICE. */
int pmh_execute_filter_groupby(const char *plan_json, uint64_t n_rows, const pmh_column *columns,
                               uint32_t n_columns, uint32_t flags, pmh_result **out);

/* Allocation-free form of pmh_execute_filter_groupby.  No counterpart in
 * the reference; it is for callers that cannot allocate device memory
 * (serving loops, harnesses that forbid hipMalloc, code next to HIP graph
 * capture).  Every intermediate buffer lives in the library's static
 * device array of PMH_NOALLOC_SCRATCH_BYTES, one per device, held until
 * the process ends; the mailbox and events it uses are made at set-up.
 *   - columns: device-resident (PMH_DEVICE) only; n_rows at most
 *     PMH_NOALLOC_MAX_ROWS.  Larger inputs, host columns or a plan whose
 *     buffers exceed the array are refused with PMH_E_VALUE.
 *   - outputs: n_outputs caller-owned device columns, keys then aggregates
 *     in plan order.  On entry data/data_len and valid/valid_len (valid may
 *     be NULL: no bitmap wanted) are the caller's buffers and their sizes in
 *     bytes; on return tag, location, data_len and valid_len describe what
 *     was written (Bool data and bitmaps padded to 4 bytes, min 4, as for
 *     pmh_execute_filter_groupby).  name is not touched.
 *   - *n_groups receives the group count.  If an output is too small for
 *     it the call returns PMH_E_VALUE with *n_groups set and no output
 *     written, so the caller can size its buffers and call again.
 * Calls on one device share its array and each clears what it needs
 * first; calls that overlap on one device race.  Synchronous: outputs are
 * complete on return. */
#define PMH_NOALLOC_MAX_ROWS 262144u
#define PMH_NOALLOC_SCRATCH_BYTES (256u << 20)
int pmh_execute_filter_groupby_noalloc(const char *plan_json, uint64_t n_rows, const pmh_column *columns,
                                       uint32_t n_columns, pmh_out_column *outputs, uint32_t n_outputs,
                                       uint64_t *n_groups);

/* Replaces _native.execute_filter_compact
 * (crates/polars-metal-core/src/udf/compact.rs:46-147): keep = pred_data AND
 * pred_valid (bit-packed, host memory); every column (I64/F64/Bool/I32/F32/
 * I16/I8/U8/U16/U32/U64) is compacted stably.  Result column i holds column i
 * compacted; result->n_rows is n_out. */
int pmh_execute_filter_compact(const uint8_t *pred_data, const uint8_t *pred_valid, uint64_t n_rows,
                               const pmh_column *columns, uint32_t n_columns, pmh_result **out);

/* Replace _native.cmp_{i64,f64}_col_scalar / _col_col
 * (crates/polars-metal-core/src/udf/compare.rs:23-145).  Host buffers.
 * Outputs are bit-packed, ceil(n/8) bytes rounded up to 4 (min 4); out_data
 * is valid AND (lhs op rhs); float compares use Polars' TotalOrd (NaN equal
 * to NaN and above every number). */
int pmh_cmp_i64_col_scalar(const int64_t *lhs, const uint8_t *lhs_valid, int64_t rhs, int op,
                           uint64_t n_rows, uint8_t *out_data, uint8_t *out_valid);
int pmh_cmp_f64_col_scalar(const double *lhs, const uint8_t *lhs_valid, double rhs, int op,
                           uint64_t n_rows, uint8_t *out_data, uint8_t *out_valid);
int pmh_cmp_i64_col_col(const int64_t *lhs, const uint8_t *lhs_valid, const int64_t *rhs,
                        const uint8_t *rhs_valid, int op, uint64_t n_rows, uint8_t *out_data,
                        uint8_t *out_valid);
int pmh_cmp_f64_col_col(const double *lhs, const uint8_t *lhs_valid, const double *rhs,
                        const uint8_t *rhs_valid, int op, uint64_t n_rows, uint8_t *out_data,
                        uint8_t *out_valid);

/* Replace _native.bool_and_dispatch / bool_or_dispatch
 * (crates/polars-metal-core/src/udf/logical.rs:26,44): Kleene AND / OR. */
int pmh_bool_and(const uint8_t *ld, const uint8_t *lv, const uint8_t *rd, const uint8_t *rv,
                 uint64_t n_rows, uint8_t *out_data, uint8_t *out_valid);
int pmh_bool_or(const uint8_t *ld, const uint8_t *lv, const uint8_t *rd, const uint8_t *rv,
                uint64_t n_rows, uint8_t *out_data, uint8_t *out_valid);

/* Replaces _native.compute_lifting_plan
 * (crates/polars-metal-core/src/router_udf.rs:22-40).  Host-only; no device
 * needed.  *out_json receives {"<Kind>#<seq>": "gpu_lift"|"cpu_leave"|
 * "fallback:<reason>"}; free it with pmh_string_free. */
int pmh_compute_lifting_plan(const char *plan_json, char **out_json);

/* Replaces _native.warmup_common_fused_signatures
 * (crates/polars-metal-core/src/udf/groupby/mod.rs:187-201): initialises the
 * device, stream and scratch arena; returns the number of kernels warmed. */
int pmh_warmup_common_fused_signatures(void);

/* hipMalloc-backed Arrow buffers on the library's device: the rework of
 * polars-metal-buffer's MetalDevice (crates/polars-metal-buffer/src/device.rs:
 * new_buffer_uninit :144, new_buffer_zeroed :55, new_buffer_from_bytes :104).
 * Pointers returned here may be passed as PMH_DEVICE columns. */
int pmh_buffer_alloc(uint64_t bytes, void **out_device_ptr);
int pmh_buffer_zeroed(uint64_t bytes, void **out_device_ptr);
int pmh_buffer_from_host(const void *src, uint64_t bytes, void **out_device_ptr);
int pmh_buffer_to_host(const void *device_ptr, void *dst, uint64_t bytes);
int pmh_buffer_free(void *device_ptr);
/* Device-to-device copy of `bytes` (synchronous).  Replaces the device
 * buffer copies of the reference's buffer crate
 * (crates/polars-metal-buffer/src/device.rs:55-153); also the known-size
 * copy that calibrates HBM traffic counters. */
int pmh_buffer_copy(void *dst_device, const void *src_device, uint64_t bytes);

/* ------------------------------------------------------------------------
 * Utf8 keys: the dictionary built and read on the device.  Replaces the
 * reference's host-side build_dict_nullable (crates/polars-metal-buffer/src/
 * dict.rs:40-66, called from crates/polars-metal-core/src/udf/groupby/
 * mod.rs:818-936) and decode_dict (dict.rs:69-74) with its packed output
 * form (mod.rs:646-670).  A key column's codes go to the fused entry as a
 * PMH_DEVICE column of tag PMH_UTF8; the key column of its result decodes
 * back to strings.
 * ---------------------------------------------------------------------- */
typedef struct pmh_utf8_column {
    uint64_t n_rows;
    int32_t location;           /* PMH_HOST or PMH_DEVICE: offsets, bytes and valid alike */
    int32_t large;              /* 0: int32 offsets (Arrow string); 1: int64 (large_string) */
    const void *offsets;        /* n_rows + 1 entries; offsets[0] may be nonzero (a slice) */
    const uint8_t *bytes;       /* the string buffer */
    uint64_t bytes_len;         /* bytes readable at `bytes`; offsets[n_rows] must not exceed it */
    const uint8_t *valid;       /* NULL: every row valid */
    uint32_t valid_bit_offset;  /* bit of valid[0] holding row 0 (0..7) */
    uint32_t pad;
} pmh_utf8_column;

typedef struct pmh_utf8_dict pmh_utf8_dict;

/* Dictionary-encodes a Utf8 column on the device.  Code k is the k-th
 * distinct string in row order (first seen wins, as build_dict_nullable);
 * null rows get code 0 and stay null; "" is a value of its own.  A valid
 * row with decreasing or out-of-buffer offsets, or bytes that are not
 * well-formed UTF-8, fails with PMH_E_VALUE naming the first such row (as
 * mod.rs:884-911).  Rows: below 2^31. */
int pmh_utf8_encode(const pmh_utf8_column *col, pmh_utf8_dict **out);
/* The device codes (u32 per row) and their validity bitmap (row i at bit i;
 * NULL when the column had none), both valid until pmh_utf8_dict_free: the
 * PMH_DEVICE key column for the fused entry.  Also the rows and the
 * distinct strings. */
int pmh_utf8_dict_codes(const pmh_utf8_dict *d, const uint32_t **codes_device, const uint8_t **valid_device,
                        uint64_t *n_rows, uint64_t *n_distinct);
/* The strings of n codes (with their validity bitmap, NULL: all valid; both
 * at `location`) in the packed wire form: u32 n, int32 offsets[n + 1], the
 * bytes; a null row has length 0.  *out_packed is the library's (release it
 * with pmh_string_free). */
int pmh_utf8_decode(const pmh_utf8_dict *d, const uint32_t *codes, const uint8_t *valid, uint64_t n,
                    int32_t location, uint8_t **out_packed, uint64_t *out_len);
void pmh_utf8_dict_free(pmh_utf8_dict *d);

/* ------------------------------------------------------------------------
 * Host columns onto the device.  Every host-to-device copy of the library
 * (host columns of the fused entry, pmh_buffer_from_host, cache misses)
 * goes through one pinned staging ring: slot-sized pieces copied by a pool
 * of host threads, each DMA overlapping the next piece's copy.  Replaces
 * the reference's StagingPool (crates/polars-metal-buffer/src/staging.rs:
 * 24-58).
 *
 * The device column cache replaces its zero-copy Arrow bridge
 * (crates/polars-metal-buffer/src/bridge.rs:56-64, which on unified memory
 * reads host bytes in place): a column is uploaded once and found again by
 * its tag, its chunks' host addresses, row counts and bit offsets, and the
 * caller's stamp, so a second query on the same buffers copies nothing.  The
 * library cannot see host writes: a caller whose bytes at an address change,
 * or whose buffer is freed and the address reused, passes a new stamp.
 * A column of several chunks (the reference concatenates Arrow chunks on the
 * host, python/polars_metal/_udf.py:1441-1456) is assembled on the device.
 * ---------------------------------------------------------------------- */
typedef struct pmh_chunk {
    uint64_t n_rows;
    const void *data;          /* host: the chunk's first value; Bool: the byte holding its first bit */
    const uint8_t *valid;      /* host validity bitmap, or NULL: every row of the chunk valid */
    uint32_t data_bit_offset;  /* Bool only: bit of data[0] holding the first row (0..7) */
    uint32_t valid_bit_offset; /* bit of valid[0] holding the first row (0..7) */
} pmh_chunk;

typedef struct pmh_resident_column {
    uint64_t entry;           /* handle for pmh_column_release / pmh_column_evict */
    uint64_t n_rows;
    const void *data;         /* device: the chunks' values, concatenated */
    uint64_t data_len;
    const uint8_t *valid;     /* device bitmap (row i at bit i), or NULL: no chunk had one */
    uint64_t valid_len;
    int32_t hit;              /* 1: found resident, no bytes copied */
    int32_t pad;
} pmh_resident_column;

typedef struct pmh_colcache_info {
    uint64_t entries;
    uint64_t bytes_resident;
    uint64_t capacity;        /* bytes; unheld entries beyond it are evicted least recently used */
    uint64_t hits, misses;
    uint64_t bytes_uploaded;  /* host bytes the cache's misses copied */
    uint64_t h2d_bytes;       /* host bytes every copy of the library moved to the device */
} pmh_colcache_info;

/* The device copy of a host column given as chunks, uploaded on a miss.  The
 * entry is held (never evicted) until pmh_column_release; its device
 * pointers stay valid while it is held.  Pass them to the fused entry as
 * PMH_DEVICE columns. */
int pmh_column_resident(int32_t tag, const pmh_chunk *chunks, uint32_t n_chunks, uint64_t stamp,
                        pmh_resident_column *out);
int pmh_column_release(uint64_t entry);
/* Frees an unheld entry now (an entry already evicted is not an error). */
int pmh_column_evict(uint64_t entry);
/* Frees every unheld entry. */
int pmh_colcache_clear(void);
/* Default: a quarter of device memory. */
int pmh_colcache_set_capacity(uint64_t bytes);
int pmh_colcache_stats(pmh_colcache_info *out);

void pmh_result_free(pmh_result *r);
void pmh_string_free(char *s);

/* Device-time instrumentation (no reference counterpart; the reference's
 * UDF returns empty timings, python/polars_metal/_udf.py:85-87).  When
 * enabled, every kernel launch is bracketed by HIP events on the launch
 * stream; pmh_kernel_times returns, per kernel name, the number of launches
 * and the summed milliseconds as JSON {"name": [launches, ms], ...}. */
int pmh_set_profiling(int enabled);
int pmh_kernel_times(char **out_json);

/* ------------------------------------------------------------------------
 * Learned sizer.  No counterpart in the reference, which sizes nothing from
 * data (its nearest precedents are fixed thresholds, cost.rs:38 and
 * groupby.rs:493); these entries exist for the cardinality / path model
 * that method "auto" may consult.  The model is a 3-layer fp32 MLP
 * (PMH_MLP_IN -> 64 -> 64 -> 2) trained and run on the device.  Outputs per
 * sample: [log2(groups + 1) / 32, logit(sort engine faster than hash)].
 * ---------------------------------------------------------------------- */
#define PMH_MLP_IN 16
#define PMH_MLP_HIDDEN 64
#define PMH_MLP_PARAMS (PMH_MLP_IN * 64 + 64 + 64 * 64 + 64 + 64 * 2 + 2)

typedef struct pmh_mlp pmh_mlp;

/* Runs only the partitioned engine's 64K-row sample sketch of a group-by
 * plan (predicate included) and returns its PMH_MLP_IN features and the
 * sketch fields of *stats.  No counterpart in the reference. */
int pmh_sketch_features(const char *plan_json, uint64_t n_rows, const pmh_column *cols, uint32_t n_cols,
                        float *features, pmh_exec_stats *stats);
/* The same sketch whole: its counts, the PMH_MLP_IN features and plain
 * distinct estimate the sketch block derives from them, and the installed
 * model's outputs as the sketch block computed them on the device.  Counts
 * of a hash-subsampled recount (a sample whose fingerprint table dropped
 * more than an eighth of it) are scaled back by 8.  No counterpart in the
 * reference. */
typedef struct pmh_sketch_info {
    uint32_t sample_size;     /* rows sampled */
    uint32_t sample_rows;     /* sampled rows that passed the predicate */
    uint32_t distinct;        /* distinct keys counted among them (side keys excluded) */
    uint32_t f1, f2;          /* keys counted once, twice */
    uint32_t dropped;         /* sampled rows whose key found no fingerprint slot */
    uint32_t null_keys;       /* sampled rows with a side key (a null or all-ones lane) */
    uint32_t hot_keys;        /* keys placed in the hot table */
    uint32_t lane_bits;       /* bits of the packed group key */
    uint32_t scale;           /* 1, or 8 after a hash-subsampled recount */
    uint32_t hist[1024];      /* keys counted c times (c clipped at 1023; hist[0] unused) */
    float features[PMH_MLP_IN];
    float prediction[2];      /* NaN, NaN with no model installed */
    double distinct_estimate; /* the plain estimate the partitions are sized from without a model */
} pmh_sketch_info;
int pmh_sketch_inspect(const char *plan_json, uint64_t n_rows, const pmh_column *cols, uint32_t n_cols,
                       pmh_sketch_info *info);

/* A device-resident model initialised from PMH_MLP_PARAMS host floats. */
int pmh_mlp_new(const float *init_params, pmh_mlp **out);
void pmh_mlp_free(pmh_mlp *m);
int pmh_mlp_params(pmh_mlp *m, float *out_params);
/* y[2n] = model(x[n][PMH_MLP_IN]) (host arrays). */
int pmh_mlp_forward(pmh_mlp *m, const float *x, uint32_t n, float *y);
/* Gradient of the mean loss over (x, y) at the current parameters into
 * grad[PMH_MLP_PARAMS]; *loss gets the mean loss.  Targets y[2n]:
 * [log2(groups + 1) / 32, 1 sort faster | 0 hash faster | -1 unknown]. */
int pmh_mlp_grad(pmh_mlp *m, const float *x, const float *y, uint32_t n, float lambda, float *grad, float *loss);
/* Uploads a training set, then runs minibatch Adam over it on the device. */
int pmh_mlp_set_data(pmh_mlp *m, const float *x, const float *y, uint32_t n);
int pmh_mlp_train(pmh_mlp *m, uint32_t epochs, uint32_t batch, float lr, float lambda, uint32_t seed,
                  float *last_epoch_loss);
/* Installs PMH_MLP_PARAMS floats as the model method "auto" consults after
 * its sample sketch (NULL uninstalls).  The model only picks the engine and
 * sizes the partitioned tables; the overflow fallback keeps every result
 * exact whatever it predicts. */
int pmh_sizer_install(const float *params);

/* ------------------------------------------------------------------------
 * Batched group-by.  No counterpart in the reference, which runs one query
 * per call with its own dispatches and blocking waits (command.rs:137-254);
 * this entry is for callers with many queries of one shape in flight.
 * ---------------------------------------------------------------------- */
typedef struct pmh_batch_query {
    uint64_t n_rows;
    const pmh_column *columns;  /* this query's columns (device memory), matched by name */
    uint32_t n_columns;
    const char *predicate_json; /* PredicateAst dict as JSON, or NULL: keep every row */
} pmh_batch_query;

typedef struct pmh_batch_stats {
    double device_ms;        /* sketch to last output, HIP events */
    uint32_t n_queries, n_tiles;
    uint64_t n_partitions;
    uint32_t n_rerun;        /* queries re-run alone (output capacity or fallback overflow) */
    uint32_t tile_rows, part_slots;
    int32_t graph_reused;    /* 1: the captured graph of an earlier batch was replayed */
    int32_t learned;         /* 1: the installed sizer sized the partitions */
    uint32_t n_fallback_keys; /* keys that missed their partition's LDS table (global fallback table) */
} pmh_batch_stats;

/* Runs n_queries filtered group-bys that share template_plan_json's keys
 * and aggregates (its predicate is ignored; each query brings its own):
 * one sketch launch, then one captured graph for every query.  results
 * receives n_queries pmh_result pointers (free each with pmh_result_free).
 * With PMH_OUT_DEVICE, a query the batch could not finish (its groups
 * outgrew the rows sized from its sketch) is re-run alone and comes back
 * with host columns: each column's location says where it is.
 * Keys: one lane of at most 64 bits; at most two aggregated columns; at
 * most 8 columns referenced by the keys, the aggregates and the queries'
 * predicates together; predicates: conjunctions of column-vs-literal
 * comparisons (PMH_E_NOTIMPL otherwise).  Each query holds fewer than 2^31
 * rows, and each of its columns is checked against its own row count
 * (buffer lengths, NULL data) before any launch (PMH_E_VALUE otherwise).
 * flags as pmh_execute_filter_groupby. */
int pmh_execute_batch(const char *template_plan_json, const pmh_batch_query *queries, uint32_t n_queries,
                      uint32_t flags, pmh_result **results, pmh_batch_stats *stats);

#ifdef __cplusplus
}
#endif
#endif /* PMH_H */
