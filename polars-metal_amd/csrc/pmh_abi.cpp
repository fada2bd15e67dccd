// pmh_abi.cpp -- the extern "C" entry points declared in include/pmh.h.
// Each converts library exceptions into a status code plus a thread-local
// message, the C form of the reference's PyErr mapping
// (crates/polars-metal-core/src/error.rs:19-33).
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <set>
#include <sstream>

#include "pmh_internal.h"
#include "pmh_mlp.h"
#include "pmh_table.h"

namespace pmh {
const char *last_error();
}

using namespace pmh;

namespace {
std::mutex g_call_mutex;

template <class F>
int guarded(F &&f) {
    std::lock_guard<std::mutex> lk(g_call_mutex);
    try {
        set_last_error("");
        f();
        return PMH_OK;
    } catch (const Error &e) {
        set_last_error(e.what());
        return e.code;
    } catch (const std::bad_alloc &) {
        set_last_error("polars_metal: host allocation failed");
        return PMH_E_COMPUTE;
    } catch (const std::exception &e) {
        set_last_error(std::string("polars_metal: ") + e.what());
        return PMH_E_RUNTIME;
    }
}

char *dup_string(const std::string &s) {
    char *p = (char *)malloc(s.size() + 1);
    memcpy(p, s.c_str(), s.size() + 1);
    return p;
}
}  // namespace

extern "C" {

const char *pmh_last_error(void) { return last_error(); }

const char *pmh_version_string(void) { return "0.1.0-mi355x"; }

int pmh_device_name(char *buf, size_t buf_len) {
    return guarded([&] {
        if (!buf || buf_len == 0) fail(PMH_E_VALUE, "polars_metal: device_name needs a buffer");
        Runtime &r = rt();
        std::string nm = r.prop.name;
        if (r.prop.gcnArchName[0]) nm += std::string(" (") + r.prop.gcnArchName + ")";
        strncpy(buf, nm.c_str(), buf_len - 1);
        buf[buf_len - 1] = 0;
    });
}

int pmh_execute_groupby(const char *plan_json, uint64_t n_rows, const pmh_column *columns, uint32_t n_columns,
                        pmh_result **out) {
    return guarded([&] {
        if (!out) fail(PMH_E_VALUE, "polars_metal: out is NULL");
        *out = nullptr;
        if (!plan_json) fail(PMH_E_VALUE, "polars_metal: plan parse error: plan is NULL");
        GroupByPlan plan = parse_groupby_plan(plan_json, false);
        *out = run_filter_groupby(plan, n_rows, columns, n_columns, PMH_OUT_HOST);
    });
}

int pmh_execute_filter_groupby(const char *plan_json, uint64_t n_rows, const pmh_column *columns,
                               uint32_t n_columns, uint32_t flags, pmh_result **out) {
    return guarded([&] {
        if (!out) fail(PMH_E_VALUE, "polars_metal: out is NULL");
        *out = nullptr;
        if (!plan_json) fail(PMH_E_VALUE, "polars_metal: plan parse error: plan is NULL");
        if (flags & ~PMH_OUT_DEVICE) fail(PMH_E_VALUE, "polars_metal: unknown flags");
        GroupByPlan plan = parse_groupby_plan(plan_json, true);
        *out = run_filter_groupby(plan, n_rows, columns, n_columns, flags);
    });
}

int pmh_execute_filter_groupby_noalloc(const char *plan_json, uint64_t n_rows, const pmh_column *columns,
                                       uint32_t n_columns, pmh_out_column *outputs, uint32_t n_outputs,
                                       uint64_t *n_groups) {
    return guarded([&] {
        if (!n_groups || (n_outputs && !outputs)) fail(PMH_E_VALUE, "polars_metal: outputs and n_groups are required");
        *n_groups = 0;
        if (!plan_json) fail(PMH_E_VALUE, "polars_metal: plan parse error: plan is NULL");
        if (n_rows > PMH_NOALLOC_MAX_ROWS)
            fail(PMH_E_VALUE, "polars_metal: allocation-free group-by takes at most " +
                                  std::to_string(PMH_NOALLOC_MAX_ROWS) + " rows, got " + std::to_string(n_rows));
        for (uint32_t i = 0; i < n_columns; ++i)
            if (columns[i].location != PMH_DEVICE)
                fail(PMH_E_VALUE, "polars_metal: allocation-free group-by takes device-resident columns only");
        GroupByPlan plan = parse_groupby_plan(plan_json, true);
        const uint32_t n_out = (uint32_t)(plan.keys.size() + plan.aggs.size());
        if (n_outputs != n_out)
            fail(PMH_E_VALUE, "polars_metal: allocation-free group-by needs " + std::to_string(n_out) +
                                  " output columns (keys, then aggregates), got " + std::to_string(n_outputs));
        for (uint32_t i = 0; i < n_outputs; ++i)
            if (!outputs[i].data && outputs[i].data_len)
                fail(PMH_E_VALUE, "polars_metal: output " + std::to_string(i) + " has no data buffer");
        pmh_result *r = nullptr;
        {
            const ScratchScope scope;
            r = run_filter_groupby(plan, n_rows, columns, n_columns, PMH_OUT_DEVICE);
            std::unique_ptr<pmh_result, void (*)(pmh_result *)> own(r, pmh_result_free);
            *n_groups = r->n_rows;
            const uint64_t ng = r->n_rows, vbytes = bitmap_bytes_padded(ng);
            for (uint32_t i = 0; i < n_outputs; ++i) {
                const pmh_out_column &c = r->columns[i];
                const uint64_t need = c.tag == PMH_BOOL ? vbytes : ng * (uint64_t)tag_size(c.tag);
                if (outputs[i].data_len < need || (outputs[i].valid && outputs[i].valid_len < vbytes))
                    fail(PMH_E_VALUE, "polars_metal: output " + std::to_string(i) + " (" + c.name + ") holds fewer than " +
                                          std::to_string(ng) + " groups (needs " + std::to_string(need) +
                                          " data bytes and " + std::to_string(vbytes) + " validity bytes)");
            }
            copy_outputs_to_caller(*r, outputs, n_outputs);
        }
    });
}

int pmh_execute_filter_compact(const uint8_t *pred_data, const uint8_t *pred_valid, uint64_t n_rows,
                               const pmh_column *columns, uint32_t n_columns, pmh_result **out) {
    return guarded([&] {
        if (!out) fail(PMH_E_VALUE, "polars_metal: out is NULL");
        *out = nullptr;
        *out = run_filter_compact(pred_data, pred_valid, n_rows, columns, n_columns);
    });
}

int pmh_cmp_i64_col_scalar(const int64_t *lhs, const uint8_t *lhs_valid, int64_t rhs, int op, uint64_t n_rows,
                           uint8_t *out_data, uint8_t *out_valid) {
    return guarded([&] { run_cmp(0, 0, lhs, lhs_valid, nullptr, nullptr, rhs, 0.0, op, n_rows, out_data, out_valid); });
}

int pmh_cmp_f64_col_scalar(const double *lhs, const uint8_t *lhs_valid, double rhs, int op, uint64_t n_rows,
                           uint8_t *out_data, uint8_t *out_valid) {
    return guarded([&] { run_cmp(1, 0, lhs, lhs_valid, nullptr, nullptr, 0, rhs, op, n_rows, out_data, out_valid); });
}

int pmh_cmp_i64_col_col(const int64_t *lhs, const uint8_t *lhs_valid, const int64_t *rhs, const uint8_t *rhs_valid,
                        int op, uint64_t n_rows, uint8_t *out_data, uint8_t *out_valid) {
    return guarded([&] { run_cmp(0, 1, lhs, lhs_valid, rhs, rhs_valid, 0, 0.0, op, n_rows, out_data, out_valid); });
}

int pmh_cmp_f64_col_col(const double *lhs, const uint8_t *lhs_valid, const double *rhs, const uint8_t *rhs_valid,
                        int op, uint64_t n_rows, uint8_t *out_data, uint8_t *out_valid) {
    return guarded([&] { run_cmp(1, 1, lhs, lhs_valid, rhs, rhs_valid, 0, 0.0, op, n_rows, out_data, out_valid); });
}

int pmh_bool_and(const uint8_t *ld, const uint8_t *lv, const uint8_t *rd, const uint8_t *rv, uint64_t n_rows,
                 uint8_t *out_data, uint8_t *out_valid) {
    return guarded([&] { run_bool(0, ld, lv, rd, rv, n_rows, out_data, out_valid); });
}

int pmh_bool_or(const uint8_t *ld, const uint8_t *lv, const uint8_t *rd, const uint8_t *rv, uint64_t n_rows,
                uint8_t *out_data, uint8_t *out_valid) {
    return guarded([&] { run_bool(1, ld, lv, rd, rv, n_rows, out_data, out_valid); });
}

int pmh_compute_lifting_plan(const char *plan_json, char **out_json) {
    return guarded([&] {
        if (!out_json) fail(PMH_E_VALUE, "polars_metal: out_json is NULL");
        *out_json = nullptr;
        if (!plan_json) fail(PMH_E_VALUE, "polars_metal: plan is NULL");
        *out_json = dup_string(compute_lifting_plan_json(plan_json));
    });
}

int pmh_warmup_common_fused_signatures(void) {
    int n = 0;
    int st = guarded([&] {
        Runtime &r = rt();
        // touch the arena so the first query does not pay for it
        slab("gb:launch", 16384);
        PMH_HIP(hipStreamSynchronize(r.stream));
        n = 1;
    });
    return st == PMH_OK ? n : -st;
}

int pmh_buffer_alloc(uint64_t bytes, void **out) {
    return guarded([&] {
        if (!out) fail(PMH_E_VALUE, "polars_metal: out is NULL");
        rt();
        *out = nullptr;
        PMH_HIP(hipMalloc(out, bytes ? bytes : 1));
    });
}

int pmh_buffer_zeroed(uint64_t bytes, void **out) {
    return guarded([&] {
        if (!out) fail(PMH_E_VALUE, "polars_metal: out is NULL");
        Runtime &r = rt();
        *out = nullptr;
        PMH_HIP(hipMalloc(out, bytes ? bytes : 1));
        if (bytes) PMH_HIP(hipMemsetAsync(*out, 0, bytes, r.stream));
        PMH_HIP(hipStreamSynchronize(r.stream));
    });
}

int pmh_buffer_from_host(const void *src, uint64_t bytes, void **out) {
    return guarded([&] {
        if (!out) fail(PMH_E_VALUE, "polars_metal: out is NULL");
        if (bytes && !src) fail(PMH_E_VALUE, "polars_metal: source is NULL");
        Runtime &r = rt();
        *out = nullptr;
        PMH_HIP(hipMalloc(out, bytes ? bytes : 1));
        upload_host(*out, src, bytes);
        PMH_HIP(hipStreamSynchronize(r.stream));
    });
}

int pmh_buffer_to_host(const void *dev, void *dst, uint64_t bytes) {
    return guarded([&] {
        if (bytes && (!dev || !dst)) fail(PMH_E_VALUE, "polars_metal: NULL buffer");
        Runtime &r = rt();
        if (bytes) PMH_HIP(hipMemcpyAsync(dst, dev, bytes, hipMemcpyDeviceToHost, r.stream));
        PMH_HIP(hipStreamSynchronize(r.stream));
    });
}

int pmh_buffer_copy(void *dst, const void *src, uint64_t bytes) {
    return guarded([&] {
        if (bytes && (!dst || !src)) fail(PMH_E_VALUE, "polars_metal: NULL buffer");
        Runtime &r = rt();
        if (bytes) PMH_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, r.stream));
        PMH_HIP(hipStreamSynchronize(r.stream));
    });
}

int pmh_buffer_free(void *dev) {
    return guarded([&] {
        Runtime &r = rt();
        if (dev) {
            PMH_HIP(hipStreamSynchronize(r.stream));
            PMH_HIP(hipFree(dev));
        }
    });
}

int pmh_column_resident(int32_t tag, const pmh_chunk *chunks, uint32_t n_chunks, uint64_t stamp,
                        pmh_resident_column *out) {
    return guarded([&] {
        if (!out) fail(PMH_E_VALUE, "polars_metal: out is NULL");
        column_resident(tag, chunks, n_chunks, stamp, out);
    });
}

int pmh_column_release(uint64_t entry) { return guarded([&] { column_release(entry); }); }
int pmh_column_evict(uint64_t entry) { return guarded([&] { column_evict(entry); }); }
int pmh_colcache_clear(void) { return guarded([&] { colcache_clear(); }); }
int pmh_colcache_set_capacity(uint64_t bytes) { return guarded([&] { colcache_set_capacity(bytes); }); }

int pmh_colcache_stats(pmh_colcache_info *out) {
    return guarded([&] {
        if (!out) fail(PMH_E_VALUE, "polars_metal: out is NULL");
        colcache_stats(out);
    });
}

int pmh_utf8_encode(const pmh_utf8_column *col, pmh_utf8_dict **out) {
    return guarded([&] {
        if (!col || !out) fail(PMH_E_VALUE, "polars_metal: utf8_encode needs a column and an out pointer");
        *out = nullptr;
        *out = utf8_encode(*col);
    });
}

int pmh_utf8_dict_codes(const pmh_utf8_dict *d, const uint32_t **codes, const uint8_t **valid, uint64_t *n_rows,
                        uint64_t *n_distinct) {
    return guarded([&] {
        if (!d) fail(PMH_E_VALUE, "polars_metal: dictionary is NULL");
        utf8_dict_codes(*d, codes, valid, n_rows, n_distinct);
    });
}

int pmh_utf8_decode(const pmh_utf8_dict *d, const uint32_t *codes, const uint8_t *valid, uint64_t n, int32_t location,
                    uint8_t **out_packed, uint64_t *out_len) {
    return guarded([&] {
        if (!d || !out_packed || !out_len) fail(PMH_E_VALUE, "polars_metal: utf8_decode needs a dictionary and outputs");
        if (n && !codes) fail(PMH_E_VALUE, "polars_metal: codes is NULL");
        *out_packed = nullptr;
        *out_len = 0;
        std::vector<uint8_t> v = utf8_decode(*d, codes, valid, n, location);
        auto *p = (uint8_t *)malloc(v.size() ? v.size() : 1);
        if (!p) fail(PMH_E_RUNTIME, "polars_metal: host allocation failed");
        memcpy(p, v.data(), v.size());
        *out_packed = p;
        *out_len = v.size();
    });
}

void pmh_utf8_dict_free(pmh_utf8_dict *d) {
    try {
        utf8_dict_free(d);
    } catch (...) {
    }
}

void pmh_result_free(pmh_result *r) {
    if (!r) return;
    for (uint32_t i = 0; i < r->n_columns; ++i) {
        pmh_out_column &c = r->columns[i];
        free((void *)c.name);
        if (c.location == PMH_HOST) {
            free(c.data);
            free(c.valid);
        }
    }
    free(r->columns);
    free(r);
}

void pmh_string_free(char *s) { free(s); }

int pmh_set_profiling(int enabled) {
    return guarded([&] {
        Runtime &r = rt();
        r.profiling = enabled != 0;
        if (!enabled) r.ktimes.clear();
    });
}

int pmh_kernel_times(char **out_json) {
    return guarded([&] {
        if (!out_json) fail(PMH_E_VALUE, "polars_metal: out_json is NULL");
        Runtime &r = rt();
        collect_kernel_times();
        std::ostringstream os;
        os << "{";
        bool first = true;
        for (auto &kv : r.ktimes) {
            if (!first) os << ", ";
            first = false;
            os << "\"" << kv.first << "\": [" << kv.second.first << ", " << kv.second.second << "]";
        }
        os << "}";
        *out_json = dup_string(os.str());
    });
}

int pmh_sketch_features(const char *plan_json, uint64_t n_rows, const pmh_column *cols, uint32_t n_cols,
                        float *features, pmh_exec_stats *stats) {
    return guarded([&] {
        if (!features) fail(PMH_E_VALUE, "polars_metal: features is NULL");
        if (!plan_json) fail(PMH_E_VALUE, "polars_metal: plan parse error: plan is NULL");
        GroupByPlan plan = parse_groupby_plan(plan_json, true);
        run_sketch(plan, n_rows, cols, n_cols, features, stats);
    });
}

int pmh_sketch_inspect(const char *plan_json, uint64_t n_rows, const pmh_column *cols, uint32_t n_cols,
                       pmh_sketch_info *info) {
    return guarded([&] {
        if (!info) fail(PMH_E_VALUE, "polars_metal: info is NULL");
        if (!plan_json) fail(PMH_E_VALUE, "polars_metal: plan parse error: plan is NULL");
        *info = pmh_sketch_info{};
        GroupByPlan plan = parse_groupby_plan(plan_json, true);
        run_sketch(plan, n_rows, cols, n_cols, nullptr, nullptr, info);
    });
}

struct pmh_mlp {
    pmh::Mlp *m;
};

int pmh_mlp_new(const float *init_params, pmh_mlp **out) {
    return guarded([&] {
        if (!out || !init_params) fail(PMH_E_VALUE, "polars_metal: mlp_new needs parameters and an out pointer");
        *out = nullptr;
        *out = new pmh_mlp{mlp_new(init_params)};
    });
}

void pmh_mlp_free(pmh_mlp *m) {
    if (!m) return;
    guarded([&] { mlp_free(m->m); });
    delete m;
}

int pmh_mlp_params(pmh_mlp *m, float *out_params) {
    return guarded([&] {
        if (!m || !out_params) fail(PMH_E_VALUE, "polars_metal: mlp_params needs a model and a buffer");
        mlp_get_params(m->m, out_params);
    });
}

int pmh_mlp_forward(pmh_mlp *m, const float *x, uint32_t n, float *y) {
    return guarded([&] {
        if (!m || (n && (!x || !y))) fail(PMH_E_VALUE, "polars_metal: mlp_forward needs a model and buffers");
        mlp_forward_host(m->m, x, n, y);
    });
}

int pmh_mlp_grad(pmh_mlp *m, const float *x, const float *y, uint32_t n, float lambda, float *grad, float *loss) {
    return guarded([&] {
        if (!m || !x || !y || !grad) fail(PMH_E_VALUE, "polars_metal: mlp_grad needs a model and buffers");
        const float l = mlp_grad_host(m->m, x, y, n, lambda, grad);
        if (loss) *loss = l;
    });
}

int pmh_mlp_set_data(pmh_mlp *m, const float *x, const float *y, uint32_t n) {
    return guarded([&] {
        if (!m || (n && (!x || !y))) fail(PMH_E_VALUE, "polars_metal: mlp_set_data needs a model and buffers");
        mlp_set_data(m->m, x, y, n);
    });
}

int pmh_mlp_train(pmh_mlp *m, uint32_t epochs, uint32_t batch, float lr, float lambda, uint32_t seed,
                  float *last_epoch_loss) {
    return guarded([&] {
        if (!m) fail(PMH_E_VALUE, "polars_metal: mlp_train needs a model");
        const float l = mlp_train(m->m, epochs, batch, lr, lambda, seed);
        if (last_epoch_loss) *last_epoch_loss = l;
    });
}

int pmh_sizer_install(const float *params) {
    return guarded([&] { sizer_install(params); });
}

int pmh_execute_batch(const char *template_plan_json, const pmh_batch_query *queries, uint32_t n_queries,
                      uint32_t flags, pmh_result **results, pmh_batch_stats *stats) {
    return guarded([&] {
        if (!results || (n_queries && !queries)) fail(PMH_E_VALUE, "polars_metal: batch needs queries and results");
        if (!template_plan_json) fail(PMH_E_VALUE, "polars_metal: plan parse error: plan is NULL");
        if (flags & ~PMH_OUT_DEVICE) fail(PMH_E_VALUE, "polars_metal: unknown flags");
        for (uint32_t i = 0; i < n_queries; ++i) results[i] = nullptr;
        GroupByPlan tmpl = parse_groupby_plan(template_plan_json, true);
        tmpl.predicate.reset();
        for (auto &a : tmpl.aggs)
            if (a.expr)
                fail(PMH_E_NOTIMPL, "polars_metal: a batch folds columns only; run expression aggregates "
                                    "through pmh_execute_filter_groupby");
        std::vector<BatchQueryIn> qs(n_queries);
        for (uint32_t i = 0; i < n_queries; ++i)
            qs[i] = BatchQueryIn{queries[i].n_rows, queries[i].columns, queries[i].n_columns, queries[i].predicate_json};
        // columns the queries' predicates read are staged with the template
        // even when no key or aggregate names them
        std::set<std::string> texts, have;
        for (auto &k : tmpl.keys) have.insert(k.first);
        for (auto &a : tmpl.aggs) if (a.op != AGG_LEN) have.insert(a.input_col);
        for (uint32_t i = 0; i < n_queries; ++i) {
            if (!queries[i].predicate_json || !texts.insert(queries[i].predicate_json).second) continue;
            const std::string js = std::string("{\"keys\": [], \"aggs\": [], \"predicate\": ") +
                                   queries[i].predicate_json + "}";
            GroupByPlan pq = parse_groupby_plan(js, true);
            if (!pq.predicate) continue;
            std::vector<std::string> names;
            collect_pred_cols(*pq.predicate, names);
            for (auto &nm : names)
                if (have.insert(nm).second) tmpl.extra_cols.push_back(nm);
        }
        // every query's row count and buffers, on the host before the device
        // is touched
        std::vector<std::string> needed;
        for (auto &k : tmpl.keys) needed.push_back(k.first);
        for (auto &a : tmpl.aggs) if (a.op != AGG_LEN) needed.push_back(a.input_col);
        for (auto &c : tmpl.extra_cols) needed.push_back(c);
        for (uint32_t i = 0; i < n_queries; ++i) {
            const std::string qn = "polars_metal: batch query " + std::to_string(i);
            if (qs[i].n_rows >= (1ull << 31)) fail(PMH_E_VALUE, qn + " holds 2^31 rows or more");
            for (auto &nm : needed) {
                const pmh_column *col = nullptr;
                for (uint32_t k = 0; k < qs[i].n_cols && qs[i].cols; ++k)
                    if (qs[i].cols[k].name && nm == qs[i].cols[k].name) { col = &qs[i].cols[k]; break; }
                if (!col) fail(PMH_E_KEY, qn + " has no column " + nm);
                check_column(*col, qs[i].n_rows, nm);
            }
        }
        std::vector<pmh_result *> out;
        try {
            run_batch(tmpl, qs, flags, out, stats);
        } catch (...) {
            for (auto *r : out) if (r) pmh_result_free(r);
            throw;
        }
        for (uint32_t i = 0; i < n_queries; ++i) results[i] = out[i];
    });
}

}  // extern "C"
