/*
 * pmh_oracle.c -- CPU oracle for the filtered group-by hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, loads or
 * calls this file.  It exists so that tests/, the smoke check and the CPU
 * baseline leg of bench.py can check (and time) the HIP path against a plain,
 * single-threaded restatement of the reference's algorithm.
 *
 * What it restates (reference paths are relative to the reference checkout):
 *   - predicate evaluation: Compare on I64/F64 operands with TotalOrd float
 *     semantics (docs/open-questions.md:55-66), Kleene AND/OR
 *     (shaders/logical_bool.metal:1-31), "null predicate drops the row"
 *     (tests/python_integration/test_filter_edges.py:46-92);
 *   - stable compaction (crates/polars-metal-kernels/src/pipeline.rs:87-270);
 *   - composite key encoding into u128 lanes: per key one null bit, then the
 *     data bits LSB first (crates/polars-metal-kernels/src/groupby.rs:155-302);
 *   - the CPU HashMap build with first-occurrence group ids
 *     (groupby.rs:497-520) and decode-by-first-row (groupby.rs:1833-1985);
 *   - the aggregators' null rules (groupby.rs:1268-1700, run_one_agg
 *     groupby.rs:2189-2274): sum skips nulls and an all-null group sums to 0;
 *     count counts non-nulls; len counts rows; min/max skip nulls and an
 *     all-null group is null; f64 min/max are poisoned by NaN
 *     (groupby.rs:1467-1469); i64/i32 sums wrap.
 *
 * f64 sums: the reference folds with Rayon in a non-deterministic order
 * (groupby.rs:1286-1326).  The oracle instead returns the correctly rounded
 * exact sum, computed with Shewchuk's non-overlapping partials (J. R.
 * Shewchuk, "Adaptive Precision Floating-Point Arithmetic and Fast Robust
 * Geometric Predicates", 1997) and the half-even final correction of
 * R. Hettinger's msum recipe -- the method behind Python's math.fsum.  A
 * second mode (sum_mode = 1) sums sequentially in row order, which is what
 * a one-thread run of the reference's fold computes.  Either way a zero
 * sum is +0.0, as the fold's +0.0 identity makes it.  Float min/max order
 * -0.0 below +0.0, so the result does not depend on row order (the
 * reference's Rayon fold keeps whichever zero it met first), and a group
 * holding a NaN yields the canonical quiet NaN.  F32 values widen exactly
 * to f64; F32 sums and means are the f64 result rounded once more to f32.
 *
 * Parity pins (tests/test_oracle.py): predicate truth tables vs the
 * reference's exhaustive Kleene table and TotalOrd examples; compaction vs a
 * brute-force numpy restatement; group sums vs math.fsum; group identity vs
 * numpy.unique; closed-form arithmetic-series groups; golden fixtures under
 * tests/golden/.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "pmh_oracle.h"

#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------ */
/* column access                                                       */
/* ------------------------------------------------------------------ */

static int valid_at(const uint8_t *valid, uint64_t i) {
    if (!valid) return 1;
    return (valid[i >> 3] >> (i & 7)) & 1;
}

/* Physical value of row i widened to int64 (integer tags and Bool). */
static int64_t load_i64(int tag, const void *data, uint64_t i) {
    switch (tag) {
    case ORC_I64: return ((const int64_t *)data)[i];
    case ORC_U64: return (int64_t)((const uint64_t *)data)[i];
    case ORC_I32: return ((const int32_t *)data)[i];
    case ORC_U32: return ((const uint32_t *)data)[i];
    case ORC_I16: return ((const int16_t *)data)[i];
    case ORC_U16: return ((const uint16_t *)data)[i];
    case ORC_I8: return ((const int8_t *)data)[i];
    case ORC_U8: return ((const uint8_t *)data)[i];
    case ORC_BOOL: return (((const uint8_t *)data)[i >> 3] >> (i & 7)) & 1;
    default: return 0;
    }
}

static double load_f64(int tag, const void *data, uint64_t i) {
    if (tag == ORC_F64) return ((const double *)data)[i];
    if (tag == ORC_F32) return (double)((const float *)data)[i];
    return (double)load_i64(tag, data, i);
}

/* Raw data bits of row i, zero-extended (the key encoder's view). */
static uint64_t load_bits(int tag, const void *data, uint64_t i) {
    switch (tag) {
    case ORC_I64: case ORC_U64: case ORC_F64: return ((const uint64_t *)data)[i];
    case ORC_I32: case ORC_U32: case ORC_F32: return ((const uint32_t *)data)[i];
    case ORC_I16: case ORC_U16: return ((const uint16_t *)data)[i];
    case ORC_I8: case ORC_U8: return ((const uint8_t *)data)[i];
    case ORC_BOOL: return (((const uint8_t *)data)[i >> 3] >> (i & 7)) & 1;
    default: return 0;
    }
}

int orc_data_bits(int tag) {
    switch (tag) {
    case ORC_I64: case ORC_U64: case ORC_F64: return 64;
    case ORC_I32: case ORC_U32: case ORC_F32: return 32;
    case ORC_I16: case ORC_U16: return 16;
    case ORC_I8: case ORC_U8: return 8;
    case ORC_BOOL: return 1;
    default: return -1;
    }
}

/* ------------------------------------------------------------------ */
/* predicate                                                           */
/* ------------------------------------------------------------------ */

typedef struct { int is_f; int64_t i; double f; int b; int valid; } slot_t;

/* TotalOrd three-way compare on doubles: NaN == NaN, NaN above all
 * non-NaN values, -0.0 == +0.0 (docs/open-questions.md:57-64). */
static int tot_cmp_f64(double a, double b) {
    int an = isnan(a), bn = isnan(b);
    if (an && bn) return 0;
    if (an) return 1;
    if (bn) return -1;
    if (a < b) return -1;
    if (a > b) return 1;
    return 0;
}

static int apply_op(int op, int c) {
    switch (op) {
    case ORC_EQ: return c == 0;
    case ORC_NE: return c != 0;
    case ORC_LT: return c < 0;
    case ORC_LE: return c <= 0;
    case ORC_GT: return c > 0;
    case ORC_GE: return c >= 0;
    default: return 0;
    }
}

/* Row r of a postfix predicate program -> (*d, *v); -1 on a malformed
 * program (which fails on every row alike). */
static int eval_row(const orc_node *nodes, int n_nodes, const orc_col *cols, uint64_t r,
                    uint8_t *d, uint8_t *v) {
    slot_t st[32];
    {
        int sp = 0;
        for (int k = 0; k < n_nodes; k++) {
            const orc_node *nd = &nodes[k];
            slot_t s = {0, 0, 0.0, 0, 0};
            switch (nd->kind) {
            case ORC_N_COLUMN: {
                const orc_col *c = &cols[nd->col];
                s.valid = valid_at(c->valid, r);
                if (nd->dtype == ORC_F64) { s.is_f = 1; s.f = load_f64(c->tag, c->data, r); }
                else if (nd->dtype == ORC_BOOL) { s.b = (int)load_i64(ORC_BOOL, c->data, r); }
                else { s.i = load_i64(c->tag, c->data, r); }
                st[sp++] = s; break;
            }
            case ORC_N_LIT_I64: s.valid = 1; s.i = nd->ival; s.f = (double)nd->ival; st[sp++] = s; break;
            case ORC_N_LIT_F64: s.valid = 1; s.is_f = 1; s.f = nd->fval; st[sp++] = s; break;
            case ORC_N_LIT_BOOL: s.valid = 1; s.b = nd->ival != 0; st[sp++] = s; break;
            case ORC_N_COMPARE: {
                slot_t rhs = st[--sp], lhs = st[--sp];
                s.valid = lhs.valid && rhs.valid;
                if (s.valid) {
                    int c;
                    if (nd->dtype == ORC_F64) {
                        double a = lhs.is_f ? lhs.f : (double)lhs.i;
                        double b = rhs.is_f ? rhs.f : (double)rhs.i;
                        c = tot_cmp_f64(a, b);
                    } else {
                        c = (lhs.i < rhs.i) ? -1 : (lhs.i > rhs.i ? 1 : 0);
                    }
                    s.b = apply_op(nd->op, c);
                }
                st[sp++] = s; break;
            }
            case ORC_N_AND: case ORC_N_OR: {
                /* Kleene 3VL (shaders/logical_bool.metal:1-31): for AND a
                 * known false dominates, for OR a known true dominates,
                 * otherwise any null makes the result null. */
                slot_t rhs = st[--sp], lhs = st[--sp];
                int dom = nd->kind == ORC_N_AND ? 0 : 1;
                int lk = lhs.valid && lhs.b == dom, rk = rhs.valid && rhs.b == dom;
                if (lk || rk) { s.valid = 1; s.b = dom; }
                else if (lhs.valid && rhs.valid) { s.valid = 1; s.b = !dom; }
                else { s.valid = 0; s.b = 0; }
                st[sp++] = s; break;
            }
            default: return -1;
            }
            if (sp >= 32) return -1;
        }
        if (sp != 1) return -1;
        *v = (uint8_t)st[0].valid;
        *d = (uint8_t)(st[0].valid && st[0].b);
    }
    return 0;
}

/* Rows are independent: with OpenMP the row loop runs on every host core
 * (the CPU baseline); each row's result is the same either way. */
int orc_eval_predicate(const orc_node *nodes, int n_nodes, const orc_col *cols,
                       uint64_t n_rows, uint8_t *out_data, uint8_t *out_valid) {
    if (n_rows == 0) return 0;
    uint8_t d0, v0;
    if (eval_row(nodes, n_nodes, cols, 0, &d0, &v0) != 0) return -1;
    out_data[0] = d0;
    out_valid[0] = v0;
#pragma omp parallel for schedule(static)
    for (uint64_t r = 1; r < n_rows; r++) (void)eval_row(nodes, n_nodes, cols, r, &out_data[r], &out_valid[r]);
    return 0;
}

/* Stable compaction of one fixed-width column by a keep mask (one byte per
 * row).  Returns the number of kept rows; out_valid gets one byte per kept
 * row. */
uint64_t orc_compact(const uint8_t *keep, uint64_t n_rows, int elem_size,
                     const void *data, const uint8_t *valid,
                     void *out_data, uint8_t *out_valid) {
    uint64_t o = 0;
    for (uint64_t r = 0; r < n_rows; r++) {
        if (!keep[r]) continue;
        if (elem_size > 0)
            memcpy((char *)out_data + o * elem_size, (const char *)data + r * elem_size, elem_size);
        out_valid[o] = (uint8_t)valid_at(valid, r);
        o++;
    }
    return o;
}

/* ------------------------------------------------------------------ */
/* exact summation                                                     */
/* ------------------------------------------------------------------ */

typedef struct {
    double *p;      /* non-overlapping partials, increasing magnitude */
    int n, cap;
    double inline_p[4];
    double special; /* sum of the non-finite inputs (IEEE) */
    int has_special;
} xsum_t;

static void xsum_init(xsum_t *s) {
    s->p = s->inline_p; s->n = 0; s->cap = 4; s->special = 0.0; s->has_special = 0;
}

static void xsum_free(xsum_t *s) {
    if (s->p != s->inline_p) free(s->p);
}

static void xsum_add(xsum_t *s, double x) {
    if (!isfinite(x)) { s->special += x; s->has_special = 1; return; }
    int i = 0;
    for (int j = 0; j < s->n; j++) {
        double y = s->p[j];
        if (fabs(x) < fabs(y)) { double t = x; x = y; y = t; }
        double hi = x + y;
        double lo = y - (hi - x);
        if (lo != 0.0) s->p[i++] = lo;
        x = hi;
    }
    if (i >= s->cap) {
        int nc = s->cap * 2;
        double *np = (double *)malloc(sizeof(double) * nc);
        memcpy(np, s->p, sizeof(double) * i);
        if (s->p != s->inline_p) free(s->p);
        s->p = np; s->cap = nc;
    }
    s->p[i++] = x;
    s->n = i;
}

static double xsum_result(const xsum_t *s) {
    if (s->has_special) return s->special; /* inf, -inf or nan */
    int n = s->n;
    double hi = 0.0, lo = 0.0;
    if (n > 0) {
        hi = s->p[--n];
        while (n > 0) {
            double x = hi, y = s->p[--n];
            hi = x + y;
            double yr = hi - x;
            lo = y - yr;
            if (lo != 0.0) break;
        }
        if (n > 0 && ((lo < 0.0 && s->p[n - 1] < 0.0) || (lo > 0.0 && s->p[n - 1] > 0.0))) {
            double y = lo * 2.0, x = hi + y, yr = x - hi;
            if (y == yr) hi = x;
        }
    }
    return hi;
}

double orc_exact_sum(const double *x, uint64_t n) {
    xsum_t s; xsum_init(&s);
    for (uint64_t i = 0; i < n; i++) xsum_add(&s, x[i]);
    double r = xsum_result(&s);
    xsum_free(&s);
    return r;
}

/* ------------------------------------------------------------------ */
/* key encoding + hash build                                           */
/* ------------------------------------------------------------------ */

int orc_encode_keys(int n_keys, const orc_col *keys, uint64_t n_rows,
                    uint64_t *lo, uint64_t *hi) {
    int off = 0;
    int null_off[16], data_off[16];
    if (n_keys <= 0 || n_keys > 16) return -1;
    for (int k = 0; k < n_keys; k++) {
        int b = orc_data_bits(keys[k].tag);
        if (b < 0) return -1;
        null_off[k] = off; data_off[k] = off + 1;
        off += 1 + b;
        if (off > 128) return -2; /* KeyEncodeError::TooWide */
    }
#pragma omp parallel for schedule(static)
    for (uint64_t r = 0; r < n_rows; r++) {
        uint64_t l = 0, h = 0;
        for (int k = 0; k < n_keys; k++) {
            int nb, sh; uint64_t v;
            if (!valid_at(keys[k].valid, r)) { nb = null_off[k]; v = 1; sh = nb; }
            else { v = load_bits(keys[k].tag, keys[k].data, r); sh = data_off[k]; }
            /* OR v << sh into the 128-bit lane */
            if (sh < 64) {
                l |= v << sh;
                if (sh > 0) h |= v >> (64 - sh);
            } else {
                h |= v << (sh - 64);
            }
        }
        lo[r] = l; hi[r] = h;
    }
    return off;
}

typedef struct {
    uint64_t *klo, *khi;
    uint32_t *gid;
    uint8_t *used;
    uint64_t mask;
} omap_t;

static uint64_t mix64(uint64_t x) {
    x ^= x >> 33; x *= 0xff51afd7ed558ccdULL;
    x ^= x >> 33; x *= 0xc4ceb9fe1a85ec53ULL;
    x ^= x >> 33; return x;
}

struct orc_gb {
    uint64_t n_groups;
    uint64_t *lane_lo, *lane_hi;
    uint32_t *first_row;
    uint32_t *len;
    int n_vals;
    int val_tags[ORC_MAX_VALS];
    uint32_t *count[ORC_MAX_VALS];
    int64_t *sum_i[ORC_MAX_VALS];
    double *sum_f[ORC_MAX_VALS];
    int64_t *min_i[ORC_MAX_VALS], *max_i[ORC_MAX_VALS];
    double *min_f[ORC_MAX_VALS], *max_f[ORC_MAX_VALS];
};

static int is_float_tag(int t) { return t == ORC_F64 || t == ORC_F32; }

/* 0x7FF8000000000000: the quiet NaN that Rust's f64::NAN denotes. */
static double canon_nan(void) {
    const uint64_t b = 0x7FF8000000000000ULL;
    double d;
    memcpy(&d, &b, sizeof d);
    return d;
}

/* a < b on non-NaN doubles with -0.0 ordered below +0.0.  The reference
 * folds min/max with a strict IEEE compare in Rayon's unspecified order
 * (groupby.rs:1471-1520), so a group holding both zeros gets whichever came
 * first; this rule picks -0.0 for min and +0.0 for max in any order. */
static int zero_signed_lt(double a, double b) {
    if (a < b) return 1;
    return a == 0.0 && b == 0.0 && signbit(a) && !signbit(b);
}

orc_gb *orc_groupby(uint64_t n_rows, const uint8_t *keep, int n_keys, const orc_col *keys,
                    int n_vals, const orc_col *vals, int sum_mode) {
    if (n_vals > ORC_MAX_VALS || n_vals < 0) return NULL;
    uint64_t *lo = (uint64_t *)malloc(sizeof(uint64_t) * (n_rows ? n_rows : 1));
    uint64_t *hi = (uint64_t *)malloc(sizeof(uint64_t) * (n_rows ? n_rows : 1));
    if (n_keys > 0) {
        if (orc_encode_keys(n_keys, keys, n_rows, lo, hi) < 0) { free(lo); free(hi); return NULL; }
    } else {
        memset(lo, 0, sizeof(uint64_t) * n_rows); memset(hi, 0, sizeof(uint64_t) * n_rows);
    }
    /* first-occurrence group ids (groupby.rs:497-520) */
    uint64_t cap = 16;
    while (cap < 2 * (n_rows + 1)) cap <<= 1;
    omap_t m;
    m.klo = (uint64_t *)malloc(sizeof(uint64_t) * cap);
    m.khi = (uint64_t *)malloc(sizeof(uint64_t) * cap);
    m.gid = (uint32_t *)malloc(sizeof(uint32_t) * cap);
    m.used = (uint8_t *)calloc(cap, 1);
    m.mask = cap - 1;
    uint32_t *r2g = (uint32_t *)malloc(sizeof(uint32_t) * (n_rows ? n_rows : 1));
    uint32_t *first = (uint32_t *)malloc(sizeof(uint32_t) * (n_rows ? n_rows : 1));
    uint64_t ng = 0;
    for (uint64_t r = 0; r < n_rows; r++) {
        if (keep && !keep[r]) { r2g[r] = UINT32_MAX; continue; }
        uint64_t h = mix64(lo[r] ^ mix64(hi[r] + 0x9e3779b97f4a7c15ULL)) & m.mask;
        for (;;) {
            if (!m.used[h]) {
                m.used[h] = 1; m.klo[h] = lo[r]; m.khi[h] = hi[r]; m.gid[h] = (uint32_t)ng;
                first[ng] = (uint32_t)r; r2g[r] = (uint32_t)ng; ng++; break;
            }
            if (m.klo[h] == lo[r] && m.khi[h] == hi[r]) { r2g[r] = m.gid[h]; break; }
            h = (h + 1) & m.mask;
        }
    }
    free(m.klo); free(m.khi); free(m.gid); free(m.used);

    orc_gb *g = (orc_gb *)calloc(1, sizeof(orc_gb));
    g->n_groups = ng;
    uint64_t na = ng ? ng : 1;
    g->lane_lo = (uint64_t *)malloc(sizeof(uint64_t) * na);
    g->lane_hi = (uint64_t *)malloc(sizeof(uint64_t) * na);
    g->first_row = (uint32_t *)malloc(sizeof(uint32_t) * na);
    g->len = (uint32_t *)calloc(na, sizeof(uint32_t));
    for (uint64_t i = 0; i < ng; i++) {
        g->first_row[i] = first[i]; g->lane_lo[i] = lo[first[i]]; g->lane_hi[i] = hi[first[i]];
    }
    for (uint64_t r = 0; r < n_rows; r++) if (r2g[r] != UINT32_MAX) g->len[r2g[r]]++;
    g->n_vals = n_vals;
    for (int j = 0; j < n_vals; j++) {
        const orc_col *c = &vals[j];
        int fl = is_float_tag(c->tag);
        g->val_tags[j] = c->tag;
        g->count[j] = (uint32_t *)calloc(na, sizeof(uint32_t));
        if (fl) {
            g->sum_f[j] = (double *)calloc(na, sizeof(double));
            g->min_f[j] = (double *)malloc(sizeof(double) * na);
            g->max_f[j] = (double *)malloc(sizeof(double) * na);
            for (uint64_t i = 0; i < ng; i++) { g->min_f[j][i] = INFINITY; g->max_f[j][i] = -INFINITY; }
        } else {
            g->sum_i[j] = (int64_t *)calloc(na, sizeof(int64_t));
            g->min_i[j] = (int64_t *)malloc(sizeof(int64_t) * na);
            g->max_i[j] = (int64_t *)malloc(sizeof(int64_t) * na);
            for (uint64_t i = 0; i < ng; i++) { g->min_i[j][i] = INT64_MAX; g->max_i[j][i] = INT64_MIN; }
        }
        xsum_t *xs = NULL;
        if (fl && sum_mode == 0) {
            xs = (xsum_t *)malloc(sizeof(xsum_t) * na);
            for (uint64_t i = 0; i < na; i++) xsum_init(&xs[i]);
        }
        for (uint64_t r = 0; r < n_rows; r++) {
            uint32_t gi = r2g[r];
            if (gi == UINT32_MAX || !valid_at(c->valid, r)) continue;
            g->count[j][gi]++;
            if (fl) {
                double v = load_f64(c->tag, c->data, r);
                if (xs) xsum_add(&xs[gi], v); else g->sum_f[j][gi] += v;
                /* NaN poisons min and max and the result is the canonical
                 * quiet NaN, Rust's f64::NAN (groupby.rs:1467-1469, :1494) */
                if (isnan(v) || isnan(g->min_f[j][gi])) g->min_f[j][gi] = canon_nan();
                else if (zero_signed_lt(v, g->min_f[j][gi])) g->min_f[j][gi] = v;
                if (isnan(v) || isnan(g->max_f[j][gi])) g->max_f[j][gi] = canon_nan();
                else if (zero_signed_lt(g->max_f[j][gi], v)) g->max_f[j][gi] = v;
            } else {
                int64_t v = load_i64(c->tag, c->data, r);
                g->sum_i[j][gi] = (int64_t)((uint64_t)g->sum_i[j][gi] + (uint64_t)v); /* wrapping */
                if (v < g->min_i[j][gi]) g->min_i[j][gi] = v;
                if (v > g->max_i[j][gi]) g->max_i[j][gi] = v;
            }
        }
        if (xs) {
            /* The reference's fold starts every group at +0.0 (groupby.rs:1304,
             * 1318), so a group whose values are all -0.0 sums to +0.0, not to
             * the -0.0 an fsum-style sum keeps. */
            for (uint64_t i = 0; i < ng; i++) {
                const double r = xsum_result(&xs[i]);
                g->sum_f[j][i] = r == 0.0 ? 0.0 : r;
                xsum_free(&xs[i]);
            }
            free(xs);
        }
    }
    free(lo); free(hi); free(r2g); free(first);
    return g;
}

uint64_t orc_gb_n_groups(const orc_gb *g) { return g->n_groups; }

void orc_gb_lanes(const orc_gb *g, uint64_t *lo, uint64_t *hi, uint32_t *first_row, uint32_t *len) {
    memcpy(lo, g->lane_lo, sizeof(uint64_t) * g->n_groups);
    memcpy(hi, g->lane_hi, sizeof(uint64_t) * g->n_groups);
    memcpy(first_row, g->first_row, sizeof(uint32_t) * g->n_groups);
    memcpy(len, g->len, sizeof(uint32_t) * g->n_groups);
}

/* Per value column j: count, and either (sum_i, min_i, max_i) or
 * (sum_f, min_f, max_f) depending on the column's tag. */
void orc_gb_val(const orc_gb *g, int j, uint32_t *count, int64_t *sum_i, int64_t *min_i,
                int64_t *max_i, double *sum_f, double *min_f, double *max_f) {
    uint64_t n = g->n_groups;
    memcpy(count, g->count[j], sizeof(uint32_t) * n);
    if (is_float_tag(g->val_tags[j])) {
        memcpy(sum_f, g->sum_f[j], sizeof(double) * n);
        memcpy(min_f, g->min_f[j], sizeof(double) * n);
        memcpy(max_f, g->max_f[j], sizeof(double) * n);
    } else {
        memcpy(sum_i, g->sum_i[j], sizeof(int64_t) * n);
        memcpy(min_i, g->min_i[j], sizeof(int64_t) * n);
        memcpy(max_i, g->max_i[j], sizeof(int64_t) * n);
    }
}

void orc_gb_free(orc_gb *g) {
    if (!g) return;
    free(g->lane_lo); free(g->lane_hi); free(g->first_row); free(g->len);
    for (int j = 0; j < g->n_vals; j++) {
        free(g->count[j]); free(g->sum_i[j]); free(g->sum_f[j]);
        free(g->min_i[j]); free(g->max_i[j]); free(g->min_f[j]); free(g->max_f[j]);
    }
    free(g);
}

/* ------------------------------------------------------------------ */
/* all-core group-by (the CPU baseline)                                */
/* ------------------------------------------------------------------ */

/* The same group-by on every host core, bit-identical to orc_groupby in
 * the same sum_mode: selected rows are hash-partitioned by their lane with
 * a stable scatter (thread chunks in row order), so each partition holds
 * its rows in row order, and each partition then folds its groups exactly
 * as orc_groupby does -- every group's f64 sum adds the same values in the
 * same order.  Groups come out partition by partition, each partition's in
 * first-occurrence order.  n_threads <= 0 takes OpenMP's default. */
orc_gb *orc_groupby_par(uint64_t n_rows, const uint8_t *keep, int n_keys, const orc_col *keys,
                        int n_vals, const orc_col *vals, int sum_mode, int n_threads) {
#ifndef _OPENMP
    return orc_groupby(n_rows, keep, n_keys, keys, n_vals, vals, sum_mode);
#else
    if (n_vals > ORC_MAX_VALS || n_vals < 0) return NULL;
    if (n_threads <= 0) n_threads = omp_get_max_threads();
    const int T = n_threads;
    enum { PBITS = 12, NP = 1 << PBITS };
    const uint64_t na = n_rows ? n_rows : 1;
    uint64_t *lo = (uint64_t *)malloc(sizeof(uint64_t) * na);
    uint64_t *hi = (uint64_t *)malloc(sizeof(uint64_t) * na);
    if (n_keys > 0) {
        if (orc_encode_keys(n_keys, keys, n_rows, lo, hi) < 0) { free(lo); free(hi); return NULL; }
    } else {
        memset(lo, 0, sizeof(uint64_t) * na); memset(hi, 0, sizeof(uint64_t) * na);
    }
    uint64_t *cnt = (uint64_t *)calloc((size_t)T * NP, sizeof(uint64_t));
    const uint64_t per = (n_rows + T - 1) / T;
#define ORC_PART(r) (mix64(lo[r] ^ mix64(hi[r] + 0x9e3779b97f4a7c15ULL)) >> (64 - PBITS))
#pragma omp parallel num_threads(T)
    {
        const int t = omp_get_thread_num();
        const uint64_t r0 = (uint64_t)t * per, r1 = r0 + per < n_rows ? r0 + per : n_rows;
        uint64_t *c = cnt + (size_t)t * NP;
        for (uint64_t r = r0; r < r1; r++)
            if (!keep || keep[r]) c[ORC_PART(r)]++;
    }
    /* partition-major offsets: partition p's rows, thread chunk by chunk */
    uint64_t *pstart = (uint64_t *)malloc(sizeof(uint64_t) * (NP + 1));
    uint64_t run = 0;
    for (int p = 0; p < NP; p++) {
        pstart[p] = run;
        for (int t = 0; t < T; t++) {
            const uint64_t c = cnt[(size_t)t * NP + p];
            cnt[(size_t)t * NP + p] = run;
            run += c;
        }
    }
    pstart[NP] = run;
    const uint64_t n_sel = run;
    uint32_t *idx = (uint32_t *)malloc(sizeof(uint32_t) * (n_sel ? n_sel : 1));
#pragma omp parallel num_threads(T)
    {
        const int t = omp_get_thread_num();
        const uint64_t r0 = (uint64_t)t * per, r1 = r0 + per < n_rows ? r0 + per : n_rows;
        uint64_t *c = cnt + (size_t)t * NP;
        for (uint64_t r = r0; r < r1; r++)
            if (!keep || keep[r]) idx[c[ORC_PART(r)]++] = (uint32_t)r;
    }
#undef ORC_PART
    free(cnt);
    /* pass 1: group ids local to each partition, first rows in place */
    uint32_t *lgid = (uint32_t *)malloc(sizeof(uint32_t) * (n_sel ? n_sel : 1));
    uint32_t *firsts = (uint32_t *)malloc(sizeof(uint32_t) * (n_sel ? n_sel : 1));
    uint64_t *png = (uint64_t *)calloc(NP + 1, sizeof(uint64_t));
#pragma omp parallel num_threads(T)
    {
        uint64_t cap = 0;
        uint64_t *tk_lo = NULL, *tk_hi = NULL;
        uint32_t *tg = NULL;
        uint8_t *used = NULL;
#pragma omp for schedule(dynamic, 16)
        for (int p = 0; p < NP; p++) {
            const uint64_t s0 = pstart[p], s1 = pstart[p + 1];
            uint64_t want = 16;
            while (want < 2 * (s1 - s0)) want <<= 1;
            /* grow by distinct keys: a hot partition of few keys stays small */
            uint64_t tcap = want < 1024 ? want : 1024;
            if (tcap > cap) {
                free(tk_lo); free(tk_hi); free(tg); free(used);
                cap = tcap;
                tk_lo = (uint64_t *)malloc(8 * cap); tk_hi = (uint64_t *)malloc(8 * cap);
                tg = (uint32_t *)malloc(4 * cap); used = (uint8_t *)malloc(cap);
            }
            uint64_t mask = tcap - 1;
            memset(used, 0, tcap);
            uint32_t ng = 0;
            for (uint64_t i = s0; i < s1; i++) {
                const uint32_t r = idx[i];
                uint64_t h = mix64(lo[r] ^ mix64(hi[r] + 0x51afd7ed558ccdULL)) & mask;
                for (;;) {
                    if (!used[h]) {
                        used[h] = 1; tk_lo[h] = lo[r]; tk_hi[h] = hi[r]; tg[h] = ng;
                        firsts[s0 + ng] = r; lgid[i] = ng++; break;
                    }
                    if (tk_lo[h] == lo[r] && tk_hi[h] == hi[r]) { lgid[i] = tg[h]; break; }
                    h = (h + 1) & mask;
                }
                if (2 * (uint64_t)ng > tcap && tcap < want) {
                    /* rehash into twice the slots */
                    const uint64_t ncap = tcap * 2;
                    uint64_t *nlo = (uint64_t *)malloc(8 * ncap), *nhi = (uint64_t *)malloc(8 * ncap);
                    uint32_t *ng_ = (uint32_t *)malloc(4 * ncap);
                    uint8_t *nu = (uint8_t *)calloc(ncap, 1);
                    for (uint64_t j = 0; j < tcap; j++) {
                        if (!used[j]) continue;
                        uint64_t q = mix64(tk_lo[j] ^ mix64(tk_hi[j] + 0x51afd7ed558ccdULL)) & (ncap - 1);
                        while (nu[q]) q = (q + 1) & (ncap - 1);
                        nu[q] = 1; nlo[q] = tk_lo[j]; nhi[q] = tk_hi[j]; ng_[q] = tg[j];
                    }
                    free(tk_lo); free(tk_hi); free(tg); free(used);
                    tk_lo = nlo; tk_hi = nhi; tg = ng_; used = nu;
                    tcap = cap = ncap;
                    mask = tcap - 1;
                }
            }
            png[p + 1] = ng;
        }
        free(tk_lo); free(tk_hi); free(tg); free(used);
    }
    for (int p = 0; p < NP; p++) png[p + 1] += png[p];
    const uint64_t ng = png[NP];
    orc_gb *g = (orc_gb *)calloc(1, sizeof(orc_gb));
    g->n_groups = ng;
    const uint64_t ga = ng ? ng : 1;
    g->lane_lo = (uint64_t *)malloc(sizeof(uint64_t) * ga);
    g->lane_hi = (uint64_t *)malloc(sizeof(uint64_t) * ga);
    g->first_row = (uint32_t *)malloc(sizeof(uint32_t) * ga);
    g->len = (uint32_t *)calloc(ga, sizeof(uint32_t));
    g->n_vals = n_vals;
    for (int j = 0; j < n_vals; j++) {
        const int fl = is_float_tag(vals[j].tag);
        g->val_tags[j] = vals[j].tag;
        g->count[j] = (uint32_t *)calloc(ga, sizeof(uint32_t));
        if (fl) {
            g->sum_f[j] = (double *)calloc(ga, sizeof(double));
            g->min_f[j] = (double *)malloc(sizeof(double) * ga);
            g->max_f[j] = (double *)malloc(sizeof(double) * ga);
        } else {
            g->sum_i[j] = (int64_t *)calloc(ga, sizeof(int64_t));
            g->min_i[j] = (int64_t *)malloc(sizeof(int64_t) * ga);
            g->max_i[j] = (int64_t *)malloc(sizeof(int64_t) * ga);
        }
    }
    /* pass 2: fold each partition's rows, in row order, into its groups */
#pragma omp parallel for num_threads(T) schedule(dynamic, 16)
    for (int p = 0; p < NP; p++) {
        const uint64_t s0 = pstart[p], s1 = pstart[p + 1], g0 = png[p], gn = png[p + 1] - png[p];
        for (uint64_t k = 0; k < gn; k++) {
            const uint32_t r = firsts[s0 + k];
            g->first_row[g0 + k] = r; g->lane_lo[g0 + k] = lo[r]; g->lane_hi[g0 + k] = hi[r];
        }
        for (uint64_t i = s0; i < s1; i++) g->len[g0 + lgid[i]]++;
        for (int j = 0; j < n_vals; j++) {
            const orc_col *c = &vals[j];
            const int fl = is_float_tag(c->tag);
            if (fl) for (uint64_t k = 0; k < gn; k++) { g->min_f[j][g0 + k] = INFINITY; g->max_f[j][g0 + k] = -INFINITY; }
            else for (uint64_t k = 0; k < gn; k++) { g->min_i[j][g0 + k] = INT64_MAX; g->max_i[j][g0 + k] = INT64_MIN; }
            xsum_t *xs = NULL;
            if (fl && sum_mode == 0) {
                xs = (xsum_t *)malloc(sizeof(xsum_t) * (gn ? gn : 1));
                for (uint64_t k = 0; k < gn; k++) xsum_init(&xs[k]);
            }
            for (uint64_t i = s0; i < s1; i++) {
                const uint32_t r = idx[i];
                if (!valid_at(c->valid, r)) continue;
                const uint64_t gi = g0 + lgid[i];
                g->count[j][gi]++;
                if (fl) {
                    const double v = load_f64(c->tag, c->data, r);
                    if (xs) xsum_add(&xs[lgid[i]], v); else g->sum_f[j][gi] += v;
                    if (isnan(v) || isnan(g->min_f[j][gi])) g->min_f[j][gi] = canon_nan();
                    else if (zero_signed_lt(v, g->min_f[j][gi])) g->min_f[j][gi] = v;
                    if (isnan(v) || isnan(g->max_f[j][gi])) g->max_f[j][gi] = canon_nan();
                    else if (zero_signed_lt(g->max_f[j][gi], v)) g->max_f[j][gi] = v;
                } else {
                    const int64_t v = load_i64(c->tag, c->data, r);
                    g->sum_i[j][gi] = (int64_t)((uint64_t)g->sum_i[j][gi] + (uint64_t)v);
                    if (v < g->min_i[j][gi]) g->min_i[j][gi] = v;
                    if (v > g->max_i[j][gi]) g->max_i[j][gi] = v;
                }
            }
            if (xs) {
                for (uint64_t k = 0; k < gn; k++) {
                    const double rr = xsum_result(&xs[k]);
                    g->sum_f[j][g0 + k] = rr == 0.0 ? 0.0 : rr;
                    xsum_free(&xs[k]);
                }
                free(xs);
            }
        }
    }
    free(lo); free(hi); free(idx); free(lgid); free(firsts); free(png); free(pstart);
    return g;
#endif
}
