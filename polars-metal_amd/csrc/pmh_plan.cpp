// pmh_plan.cpp -- plan wire format: a small JSON reader, the group-by plan
// parser and the lifting router.
//
// The JSON keys are exactly those of the reference's Python plan dicts:
//   group-by plan   crates/polars-metal-core/src/udf/groupby/mod.rs:362-501
//   predicate AST   crates/polars-metal-core/src/udf/predicate.rs:92-200
//   router input    crates/polars-metal-core/src/router_udf.rs:22-175
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <sstream>

#include "pmh_internal.h"

namespace pmh {

// ---------------------------------------------------------------------------
// JSON value + recursive-descent reader
// ---------------------------------------------------------------------------
namespace {

struct JVal {
    enum T { NUL, BOOL, INT, DBL, STR, ARR, OBJ } t = NUL;
    bool b = false;
    int64_t i = 0;
    double d = 0.0;
    std::string s;
    std::vector<JVal> a;
    std::vector<std::pair<std::string, JVal>> o;
    const JVal *get(const char *k) const {
        if (t != OBJ) return nullptr;
        for (auto &kv : o)
            if (kv.first == k) return &kv.second;
        return nullptr;
    }
    bool is_num() const { return t == INT || t == DBL; }
    double num() const { return t == INT ? (double)i : d; }
};

struct Reader {
    const char *p, *e;
    [[noreturn]] void bad(const char *what) {
        fail(PMH_E_VALUE, std::string("polars_metal: plan parse error: malformed JSON (") + what + ")");
    }
    void ws() {
        while (p < e && (*p == ' ' || *p == '\n' || *p == '\t' || *p == '\r')) ++p;
    }
    bool lit(const char *s) {
        size_t n = strlen(s);
        if ((size_t)(e - p) >= n && memcmp(p, s, n) == 0) { p += n; return true; }
        return false;
    }
    JVal value(int depth) {
        if (depth > 64) bad("nesting too deep");
        ws();
        if (p >= e) bad("unexpected end");
        JVal v;
        char c = *p;
        if (c == '{') {
            ++p; v.t = JVal::OBJ; ws();
            if (p < e && *p == '}') { ++p; return v; }
            for (;;) {
                ws();
                if (p >= e || *p != '"') bad("object key");
                std::string k = str();
                ws();
                if (p >= e || *p != ':') bad("colon");
                ++p;
                v.o.emplace_back(k, value(depth + 1));
                ws();
                if (p < e && *p == ',') { ++p; continue; }
                if (p < e && *p == '}') { ++p; break; }
                bad("object separator");
            }
        } else if (c == '[') {
            ++p; v.t = JVal::ARR; ws();
            if (p < e && *p == ']') { ++p; return v; }
            for (;;) {
                v.a.push_back(value(depth + 1));
                ws();
                if (p < e && *p == ',') { ++p; continue; }
                if (p < e && *p == ']') { ++p; break; }
                bad("array separator");
            }
        } else if (c == '"') {
            v.t = JVal::STR; v.s = str();
        } else if (lit("true")) { v.t = JVal::BOOL; v.b = true; }
        else if (lit("false")) { v.t = JVal::BOOL; v.b = false; }
        else if (lit("null")) { v.t = JVal::NUL; }
        else if (lit("NaN")) { v.t = JVal::DBL; v.d = NAN; }
        else if (lit("Infinity")) { v.t = JVal::DBL; v.d = INFINITY; }
        else if (lit("-Infinity")) { v.t = JVal::DBL; v.d = -INFINITY; }
        else {
            const char *s = p;
            bool flt = false;
            if (p < e && (*p == '-' || *p == '+')) ++p;
            while (p < e && ((*p >= '0' && *p <= '9') || *p == '.' || *p == 'e' || *p == 'E' ||
                             *p == '-' || *p == '+')) {
                if (*p == '.' || *p == 'e' || *p == 'E') flt = true;
                ++p;
            }
            if (p == s) bad("unexpected character");
            std::string num(s, p);
            if (!flt) {
                errno = 0;
                char *end = nullptr;
                long long x = strtoll(num.c_str(), &end, 10);
                if (errno == 0 && end && *end == 0) { v.t = JVal::INT; v.i = x; return v; }
            }
            char *end = nullptr;
            v.t = JVal::DBL; v.d = strtod(num.c_str(), &end);
            if (!end || *end != 0) bad("number");
        }
        return v;
    }
    std::string str() {
        std::string out;
        ++p;  // opening quote
        while (p < e && *p != '"') {
            if (*p == '\\') {
                ++p;
                if (p >= e) bad("escape");
                char c = *p++;
                switch (c) {
                case 'n': out += '\n'; break;
                case 't': out += '\t'; break;
                case 'r': out += '\r'; break;
                case 'b': out += '\b'; break;
                case 'f': out += '\f'; break;
                case 'u': {
                    if (e - p < 4) bad("unicode escape");
                    unsigned cp = (unsigned)strtoul(std::string(p, p + 4).c_str(), nullptr, 16);
                    p += 4;
                    if (cp >= 0xD800 && cp <= 0xDBFF && e - p >= 6 && p[0] == '\\' && p[1] == 'u') {
                        unsigned lo = (unsigned)strtoul(std::string(p + 2, p + 6).c_str(), nullptr, 16);
                        if (lo >= 0xDC00 && lo <= 0xDFFF) {
                            cp = 0x10000 + ((cp - 0xD800) << 10) + (lo - 0xDC00);
                            p += 6;
                        }
                    }
                    if (cp < 0x80) out += (char)cp;
                    else if (cp < 0x800) { out += (char)(0xC0 | (cp >> 6)); out += (char)(0x80 | (cp & 0x3F)); }
                    else if (cp < 0x10000) {
                        out += (char)(0xE0 | (cp >> 12)); out += (char)(0x80 | ((cp >> 6) & 0x3F));
                        out += (char)(0x80 | (cp & 0x3F));
                    } else {
                        out += (char)(0xF0 | (cp >> 18)); out += (char)(0x80 | ((cp >> 12) & 0x3F));
                        out += (char)(0x80 | ((cp >> 6) & 0x3F)); out += (char)(0x80 | (cp & 0x3F));
                    }
                    break;
                }
                default: out += c;
                }
            } else {
                out += *p++;
            }
        }
        if (p >= e) bad("unterminated string");
        ++p;
        return out;
    }
};

JVal parse_json(const std::string &s) {
    Reader r{s.data(), s.data() + s.size()};
    JVal v = r.value(0);
    r.ws();
    if (r.p != r.e) r.bad("trailing characters");
    return v;
}

[[noreturn]] void parse_err(const std::string &what) {
    fail(PMH_E_VALUE, "polars_metal: plan parse error: " + what);
}

std::string need_str(const JVal *v, const char *field) {
    if (!v || v->t != JVal::STR) parse_err(std::string("wrong type for field ") + field);
    return v->s;
}

int parse_op(const std::string &s) {
    if (s == "Sum") return AGG_SUM;
    if (s == "Mean") return AGG_MEAN;
    if (s == "Count") return AGG_COUNT;
    if (s == "Min") return AGG_MIN;
    if (s == "Max") return AGG_MAX;
    if (s == "Len") return AGG_LEN;
    return -1;
}

std::unique_ptr<PlanPredNode> parse_pred(const JVal &v, int depth) {
    if (depth > 16) parse_err("predicate nested too deeply");
    if (v.t != JVal::OBJ) parse_err("predicate must be an object");
    auto n = std::make_unique<PlanPredNode>();
    const JVal *k = v.get("kind");
    if (!k || k->t != JVal::STR) fail(PMH_E_KEY, "predicate dict missing 'kind'");
    n->kind = k->s;
    if (n->kind == "Column") {
        const JVal *nm = v.get("name");
        const JVal *dt = v.get("dtype");
        if (!nm) fail(PMH_E_KEY, "Column missing 'name'");
        if (!dt) fail(PMH_E_KEY, "Column missing 'dtype'");
        n->name = need_str(nm, "name");
        n->dtype = need_str(dt, "dtype");
        if (tag_from_wire(n->dtype) < 0)
            fail(PMH_E_VALUE, "polars_metal: unknown MetalDtype tag \"" + n->dtype + "\"");
    } else if (n->kind == "LiteralI64") {
        const JVal *x = v.get("value");
        if (!x) fail(PMH_E_KEY, "LiteralI64 missing 'value'");
        if (x->t != JVal::INT) parse_err("LiteralI64 value must be an integer");
        n->ival = x->i; n->fval = (double)x->i; n->lit_is_int = true;
    } else if (n->kind == "LiteralF64") {
        const JVal *x = v.get("value");
        if (!x) fail(PMH_E_KEY, "LiteralF64 missing 'value'");
        if (!x->is_num()) parse_err("LiteralF64 value must be a number");
        n->fval = x->num();
    } else if (n->kind == "LiteralBool") {
        const JVal *x = v.get("value");
        if (!x) fail(PMH_E_KEY, "LiteralBool missing 'value'");
        if (x->t != JVal::BOOL) parse_err("LiteralBool value must be a bool");
        n->bval = x->b;
    } else if (n->kind == "Compare" || n->kind == "And" || n->kind == "Or") {
        if (n->kind == "Compare") {
            const JVal *op = v.get("op");
            if (!op) fail(PMH_E_KEY, "Compare missing 'op'");
            n->op = need_str(op, "op");
            static const char *ops[] = {"Eq", "Ne", "Lt", "Le", "Gt", "Ge"};
            bool ok = false;
            for (auto o : ops) ok |= n->op == o;
            if (!ok) parse_err("unknown compare op \"" + n->op + "\"");
            if (const JVal *dt = v.get("dtype")) n->dtype = need_str(dt, "dtype");
        }
        const JVal *l = v.get("lhs");
        const JVal *r = v.get("rhs");
        if (!l) fail(PMH_E_KEY, n->kind + " missing 'lhs'");
        if (!r) fail(PMH_E_KEY, n->kind + " missing 'rhs'");
        n->lhs = parse_pred(*l, depth + 1);
        n->rhs = parse_pred(*r, depth + 1);
    } else {
        fail(PMH_E_NOTIMPL, "polars_metal: predicate kind \"" + n->kind + "\" lands in a later phase");
    }
    return n;
}

// parse_agg_expr_dict (crates/polars-metal-core/src/udf/groupby/mod.rs:278-360)
// with AggExpr::validate's depth cap of four binary levels (plan/mod.rs:169-178).
std::shared_ptr<PlanExpr> parse_expr(const JVal &v, int depth) {
    if (v.t != JVal::OBJ) parse_err("wrong type for field expr(dict)");
    auto e = std::make_shared<PlanExpr>();
    const JVal *k = v.get("kind");
    if (!k || k->t != JVal::STR) parse_err("wrong type for field expr.kind");
    if (k->s == "Column") {
        const JVal *nm = v.get("name");
        if (!nm || nm->t != JVal::STR) parse_err("wrong type for field expr.name");
        e->kind = PlanExpr::COLUMN;
        e->name = nm->s;
    } else if (k->s == "LiteralF64") {
        const JVal *x = v.get("value");
        if (!x || !x->is_num()) parse_err("wrong type for field expr.value(f64)");
        e->kind = PlanExpr::LIT_F64;
        e->fval = x->num();
    } else if (k->s == "LiteralI64") {
        const JVal *x = v.get("value");
        if (!x || x->t != JVal::INT) parse_err("wrong type for field expr.value(i64)");
        e->kind = PlanExpr::LIT_I64;
        e->ival = x->i;
    } else if (k->s == "Binary") {
        const JVal *op = v.get("op");
        if (!op || op->t != JVal::STR) parse_err("wrong type for field expr.op");
        e->kind = PlanExpr::BINARY;
        if (op->s == "Add") e->op = EX_ADD;
        else if (op->s == "Sub") e->op = EX_SUB;
        else if (op->s == "Mul") e->op = EX_MUL;
        else if (op->s == "Div") e->op = EX_DIV;
        else parse_err("unknown agg op binary op " + op->s);
        if (depth >= kMaxExprDepth) parse_err("wrong type for field expr(too deep)");
        const JVal *l = v.get("lhs");
        const JVal *r = v.get("rhs");
        if (!l) parse_err("wrong type for field expr.lhs");
        if (!r) parse_err("wrong type for field expr.rhs");
        e->lhs = parse_expr(*l, depth + 1);
        e->rhs = parse_expr(*r, depth + 1);
    } else {
        parse_err("wrong type for field expr.kind");
    }
    return e;
}

}  // namespace

bool expr_equal(const PlanExpr &a, const PlanExpr &b) {
    if (a.kind != b.kind) return false;
    switch (a.kind) {
    case PlanExpr::COLUMN: return a.name == b.name;
    case PlanExpr::LIT_I64: return a.ival == b.ival;
    case PlanExpr::LIT_F64: return memcmp(&a.fval, &b.fval, sizeof(double)) == 0;
    default: return a.op == b.op && expr_equal(*a.lhs, *b.lhs) && expr_equal(*a.rhs, *b.rhs);
    }
}

void collect_expr_cols(const PlanExpr &e, std::vector<std::string> &out) {
    if (e.kind == PlanExpr::COLUMN) out.push_back(e.name);
    if (e.lhs) collect_expr_cols(*e.lhs, out);
    if (e.rhs) collect_expr_cols(*e.rhs, out);
}

const char *tag_name(int tag) {
    static const char *names[] = {"I64", "F64", "Bool", "I32", "F32", "I8", "I16", "U8", "U16", "U32", "U64", "Utf8"};
    if (tag < 0 || tag > PMH_UTF8) return "?";
    return names[tag];
}

int tag_from_wire(const std::string &s) {
    // crates/polars-metal-core/src/plan/mod.rs:78-101 (Date aliases I32)
    if (s == "I64") return PMH_I64;
    if (s == "F64") return PMH_F64;
    if (s == "Bool") return PMH_BOOL;
    if (s == "I32" || s == "Date") return PMH_I32;
    if (s == "F32") return PMH_F32;
    if (s == "I8") return PMH_I8;
    if (s == "I16") return PMH_I16;
    if (s == "U8") return PMH_U8;
    if (s == "U16") return PMH_U16;
    if (s == "U32") return PMH_U32;
    if (s == "U64") return PMH_U64;
    if (s == "Utf8" || s == "String") return PMH_UTF8;
    return -1;
}

// Mirrors parse_groupby_plan (crates/polars-metal-core/src/udf/groupby/mod.rs:362-501):
// same required fields, the same legacy inference of "kind" from op, and the
// same error wording ("plan parse error: ...").
GroupByPlan parse_groupby_plan(const std::string &json, bool allow_predicate) {
    JVal root = parse_json(json);
    if (root.t != JVal::OBJ) parse_err("plan must be an object");
    GroupByPlan plan;
    const JVal *keys = root.get("keys");
    if (!keys) parse_err("missing field keys");
    if (keys->t != JVal::ARR) parse_err("wrong type for field keys");
    for (auto &e : keys->a) {
        if (e.t != JVal::ARR || e.a.size() < 2) parse_err("wrong type for field key entry");
        if (e.a[0].t != JVal::STR) parse_err("wrong type for field key name");
        if (e.a[1].t != JVal::STR) parse_err("wrong type for field key dtype");
        int tag = tag_from_wire(e.a[1].s);
        if (tag < 0) parse_err("unknown dtype " + e.a[1].s);
        plan.keys.emplace_back(e.a[0].s, tag);
    }
    const JVal *aggs = root.get("aggs");
    if (!aggs) parse_err("missing field aggs");
    if (aggs->t != JVal::ARR) parse_err("wrong type for field aggs");
    for (auto &e : aggs->a) {
        if (e.t != JVal::OBJ) parse_err("wrong type for field agg entry");
        std::string kind;
        if (const JVal *k = e.get("kind"); k && k->t == JVal::STR) kind = k->s;
        else {
            const JVal *op = e.get("op");
            kind = (op && op->t == JVal::STR && op->s == "Len") ? "Length" : "Simple";
        }
        const JVal *alias = e.get("output_alias");
        if (!alias || alias->t != JVal::STR) parse_err("wrong type for field output_alias");
        PlanAgg a;
        a.alias = alias->s;
        if (kind == "Length") {
            a.op = AGG_LEN;
        } else if (kind == "Simple") {
            if (const JVal *ic = e.get("input_col"); ic && ic->t == JVal::STR) a.input_col = ic->s;
            const JVal *op = e.get("op");
            if (!op || op->t != JVal::STR) parse_err("wrong type for field op");
            a.op = parse_op(op->s);
            if (a.op < 0) parse_err("unknown agg op " + op->s);
        } else if (kind == "Expression") {
            // The reference evaluates the expression through Polars on the
            // host before the call (python/polars_metal/_udf.py:1164-1209);
            // here it is evaluated on the device (pmh_expr.hip).
            const JVal *op = e.get("op");
            if (!op || op->t != JVal::STR) parse_err("wrong type for field op");
            a.op = parse_op(op->s);
            if (a.op < 0) parse_err("unknown agg op " + op->s);
            const JVal *ex = e.get("expr");
            if (!ex) parse_err("wrong type for field expr");
            a.expr = parse_expr(*ex, 0);
            if (a.op == AGG_LEN) a.expr.reset();
        } else {
            parse_err("unknown agg op kind=" + kind);
        }
        plan.aggs.push_back(a);
    }
    if (const JVal *pr = root.get("predicate"); pr && pr->t != JVal::NUL) {
        if (!allow_predicate) parse_err("execute_groupby takes no predicate; use the fused entry");
        plan.predicate = parse_pred(*pr, 0);
    }
    if (const JVal *op = root.get("options"); op && op->t == JVal::OBJ) {
        for (auto &kv : op->o) {
            const JVal &v = kv.second;
            if (kv.first == "method") {
                if (v.t != JVal::STR || (v.s != "auto" && v.s != "hash" && v.s != "hash_global" && v.s != "sort"))
                    parse_err("options.method must be auto|hash|hash_global|sort");
                plan.opts.method = v.s;
            } else {
                if (v.t != JVal::INT) parse_err("options." + kv.first + " must be an integer");
                if (kv.first == "capacity_hint") plan.opts.capacity_hint = v.i;
                else if (kv.first == "lds_slots") plan.opts.lds_slots = v.i;
                else if (kv.first == "block") plan.opts.block = v.i;
                else if (kv.first == "grid") plan.opts.grid = v.i;
                else if (kv.first == "tile") plan.opts.tile = v.i;
                else if (kv.first == "partitions") plan.opts.partitions = v.i;
                else if (kv.first == "tail_slots") plan.opts.tail_slots = v.i;
                else if (kv.first == "sample") plan.opts.sample = v.i;
                else if (kv.first == "distinct_hint") plan.opts.distinct_hint = v.i;
                else if (kv.first == "fast_scan") plan.opts.fast_scan = v.i;
                else if (kv.first == "scan_rows") plan.opts.scan_rows = v.i;
                else if (kv.first == "scan_blocks") plan.opts.scan_blocks = v.i;
                else if (kv.first == "narrow_keys") plan.opts.narrow_keys = v.i;
                else if (kv.first == "sort_buckets") plan.opts.sort_buckets = v.i;
                else if (kv.first == "sort_carry") plan.opts.sort_carry = v.i;
                else if (kv.first == "sort_fold") plan.opts.sort_fold = v.i;
                else if (kv.first == "sort_sample_bits") plan.opts.sort_sample_bits = v.i;
                else if (kv.first == "sort_msd") plan.opts.sort_msd = v.i;
                else parse_err("unknown option " + kv.first);
            }
        }
    }
    return plan;
}

// ---------------------------------------------------------------------------
// router (crates/polars-metal-core/src/router_udf.rs:22-175, cost.rs:38-111)
// ---------------------------------------------------------------------------
namespace {

// Smallest input at which the MI355X path is lifted.  The reference tuned
// 100_000 rows on M2 Ultra (cost.rs:38); a fused device launch here costs
// tens of microseconds, so the same threshold is kept.
constexpr int64_t kGpuMinRows = 100000;

struct Route {
    std::vector<std::pair<std::string, std::string>> out;
    int seq = 0;
};

int key_width_bits(int tag) { return 1 + tag_data_bits(tag); }

int64_t peek_rows(const JVal &d) {
    const JVal *k = d.get("kind");
    if (!k || k->t != JVal::STR) fail(PMH_E_KEY, "missing 'kind'");
    if (k->s == "Scan") {
        const JVal *n = d.get("n_rows");
        if (!n) fail(PMH_E_KEY, "Scan: missing n_rows");
        if (n->t != JVal::INT) parse_err("Scan.n_rows must be an integer");
        return n->i;
    }
    if (k->s == "Project" || k->s == "Filter" || k->s == "GroupBy" || k->s == "Sort" || k->s == "HStack") {
        const JVal *in = d.get("input");
        if (!in || in->t != JVal::OBJ) fail(PMH_E_KEY, "missing 'input' in row-count peek");
        return peek_rows(*in);
    }
    return 0;
}

std::string route(const JVal &d, Route &r, std::string *decision) {
    const JVal *k = d.get("kind");
    if (!k || k->t != JVal::STR) fail(PMH_E_KEY, "router: missing 'kind'");
    const std::string &kind = k->s;
    auto child = [&](std::string *cd) -> std::string {
        const JVal *in = d.get("input");
        if (!in || in->t != JVal::OBJ) fail(PMH_E_KEY, kind + ": missing input");
        return route(*in, r, cd);
    };
    std::string dec;
    if (kind == "Scan") {
        dec = "cpu_leave";
    } else if (kind == "Project" || kind == "Sort" || kind == "HStack") {
        std::string cd;
        child(&cd);
        dec = cd;  // inherit (cost.rs decide_project)
    } else if (kind == "Filter") {
        // The reference leaves every Filter on the CPU (cost.rs:40-43)
        // because its compaction ran through Polars.  Here compaction and the
        // fused filter+group-by are device kernels, so a Filter is lifted on
        // the same row threshold as a GroupBy.
        int64_t n = peek_rows(d);
        std::string cd;
        child(&cd);
        dec = n > kGpuMinRows ? "gpu_lift" : "cpu_leave";
    } else if (kind == "GroupBy") {
        const JVal *in = d.get("input");
        if (!in || in->t != JVal::OBJ) fail(PMH_E_KEY, "GroupBy: missing input");
        int64_t n = peek_rows(*in);
        std::string cd;
        route(*in, r, &cd);
        int bits = 0;
        if (const JVal *keys = d.get("keys"); keys && keys->t == JVal::ARR) {
            for (auto &e : keys->a) {
                if (e.t == JVal::ARR && e.a.size() >= 2 && e.a[1].t == JVal::STR) {
                    int tag = tag_from_wire(e.a[1].s);
                    if (tag >= 0) bits += key_width_bits(tag);
                }
            }
        }
        if (bits > 128) {
            dec = "fallback:composite key total " + std::to_string(bits) + " bits; supports <= 128";
        } else {
            dec = n > kGpuMinRows ? "gpu_lift" : "cpu_leave";
        }
    } else {
        if (const JVal *in = d.get("input"); in && in->t == JVal::OBJ) route(*in, r, nullptr);
        dec = "fallback:unsupported IR node: " + kind;
    }
    std::string id = kind + "#" + std::to_string(r.seq++);
    r.out.emplace_back(id, dec);
    if (decision) *decision = dec;
    return id;
}

std::string json_escape(const std::string &s) {
    std::string o;
    for (char c : s) {
        if (c == '"' || c == '\\') { o += '\\'; o += c; }
        else if ((unsigned char)c < 0x20) {
            char buf[8];
            snprintf(buf, sizeof buf, "\\u%04x", (unsigned char)c);
            o += buf;
        } else o += c;
    }
    return o;
}

}  // namespace

std::string compute_lifting_plan_json(const std::string &json) {
    JVal root = parse_json(json);
    if (root.t != JVal::OBJ) parse_err("plan must be an object");
    Route r;
    route(root, r, nullptr);
    std::ostringstream os;
    os << "{";
    for (size_t i = 0; i < r.out.size(); ++i) {
        if (i) os << ", ";
        os << "\"" << json_escape(r.out[i].first) << "\": \"" << json_escape(r.out[i].second) << "\"";
    }
    os << "}";
    return os.str();
}

}  // namespace pmh
