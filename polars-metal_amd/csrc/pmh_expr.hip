// pmh_expr.hip -- expression aggregates: the per-row value of an AggExpr
// tree (crates/polars-metal-core/src/plan/mod.rs:133-142), evaluated on the
// device into a column that the group-by then folds like any other.
//
// The reference evaluates these trees through Polars on the host before the
// group-by call (python/polars_metal/_udf.py:1164-1209) and hands the result
// over as a Simple agg; here one launch evaluates the tree for every row.
//
// Arithmetic follows Polars:
//   * Add / Sub / Mul compute in the operands' numeric supertype; integer
//     results wrap in that width; float results round to nearest in that
//     width, one rounding per operation (no fused multiply-add).
//   * Div is true division: Float64 when both operands are integers, the
//     float supertype otherwise.
//   * A literal takes its type from the other operand: an integer literal
//     becomes that integer type (refused if it does not fit) or that float
//     type; a float literal becomes Float32 beside a Float32 and Float64
//     otherwise.  Literal-only subtrees fold on the host.
//   * Numeric supertypes: the wider of two signed or two unsigned types; a
//     signed and an unsigned type give the signed type when it is wider,
//     else the signed type twice the unsigned one's width; an integer of at
//     most 16 bits with Float32 gives Float32, a wider one Float64; anything
//     with Float64 gives Float64.
//   * A row is null when any column the tree reads is null there.
// Shapes outside these rules (Bool or string operands, a signed type with
// UInt64, a literal-only tree) raise PMH_E_NOTIMPL.
#include <cstring>

#include "pmh_device.h"

namespace pmh {

namespace {

constexpr int kExprBlock = 256;
constexpr int kExprStack = kMaxExprDepth + 1;  // postfix stack of a tree four binary levels deep

__host__ __device__ inline bool ex_is_int(int t) {
    return t == PMH_I8 || t == PMH_I16 || t == PMH_I32 || t == PMH_I64 || t == PMH_U8 || t == PMH_U16 ||
           t == PMH_U32 || t == PMH_U64;
}

// A column value as 64 bits: integers sign- or zero-extended, F32 bits in
// the low half, F64 bits.
__device__ __forceinline__ uint64_t ex_load(const DevCol &c, int tag, uint64_t i) {
    switch (tag) {
    case PMH_I64: case PMH_U64: case PMH_F64: return ((const uint64_t *)c.data)[i];
    case PMH_I32: return (uint64_t)(int64_t)((const int32_t *)c.data)[i];
    case PMH_U32: case PMH_F32: return ((const uint32_t *)c.data)[i];
    case PMH_I16: return (uint64_t)(int64_t)((const int16_t *)c.data)[i];
    case PMH_U16: return ((const uint16_t *)c.data)[i];
    case PMH_I8: return (uint64_t)(int64_t)((const int8_t *)c.data)[i];
    default: return ((const uint8_t *)c.data)[i];  // PMH_U8
    }
}

// Truncates a wrapped 64-bit integer result to the width of `tag` and
// extends it back (sign for signed types).
__device__ __forceinline__ uint64_t ex_wrap(uint64_t v, int tag) {
    switch (tag) {
    case PMH_I32: return (uint64_t)(int64_t)(int32_t)(uint32_t)v;
    case PMH_U32: return (uint32_t)v;
    case PMH_I16: return (uint64_t)(int64_t)(int16_t)(uint16_t)v;
    case PMH_U16: return (uint16_t)v;
    case PMH_I8: return (uint64_t)(int64_t)(int8_t)(uint8_t)v;
    case PMH_U8: return (uint8_t)v;
    default: return v;
    }
}

// Value of dtype `from` as dtype `to` (a widening: supertypes never narrow).
__device__ __forceinline__ uint64_t ex_conv(uint64_t v, int from, int to) {
    if (from == to || ex_is_int(to)) return v;
    if (to == PMH_F32) {  // from an integer of at most 16 bits: exact
        float f = tag_signed(from) ? (float)(int64_t)v : (float)v;
        return __float_as_uint(f);
    }
    double d;  // to == PMH_F64
    if (from == PMH_F32) d = (double)__uint_as_float((uint32_t)v);
    else if (tag_signed(from)) d = (double)(int64_t)v;
    else d = (double)v;
    return (uint64_t)__double_as_longlong(d);
}

__device__ __forceinline__ uint64_t ex_binary(int kind, int ctag, uint64_t a, uint64_t b) {
#pragma clang fp contract(off)
    if (ctag == PMH_F64) {
        double x = __longlong_as_double((long long)a), y = __longlong_as_double((long long)b), r;
        if (kind == EX_ADD) r = x + y;
        else if (kind == EX_SUB) r = x - y;
        else if (kind == EX_MUL) r = x * y;
        else r = x / y;
        return (uint64_t)__double_as_longlong(r);
    }
    if (ctag == PMH_F32) {
        float x = __uint_as_float((uint32_t)a), y = __uint_as_float((uint32_t)b), r;
        if (kind == EX_ADD) r = x + y;
        else if (kind == EX_SUB) r = x - y;
        else if (kind == EX_MUL) r = x * y;
        else r = x / y;
        return __float_as_uint(r);
    }
    uint64_t r = kind == EX_ADD ? a + b : kind == EX_SUB ? a - b : a * b;
    return ex_wrap(r, ctag);
}

// One thread per row; the program sits in the kernel arguments, so every
// node is read with scalar loads and the stack pointer is uniform.
__global__ __launch_bounds__(kExprBlock) void ex_eval(const ExprProg P, uint64_t n, void *__restrict__ out,
                                                      uint64_t *__restrict__ out_valid) {
    const int root_tag = P.nodes[P.n_nodes - 1].tag;
    const int root_size = tag_size(root_tag);
    for (uint64_t base = (uint64_t)blockIdx.x * kExprBlock; base < n; base += (uint64_t)gridDim.x * kExprBlock) {
        const uint64_t i = base + threadIdx.x;
        const bool in = i < n;
        const uint64_t r = in ? i : base;  // base < n: an in-bounds row for idle lanes
        uint64_t s[kExprStack];
        int sp = 0;
        for (int k = 0; k < P.n_nodes; ++k) {
            const ExprNode &nd = P.nodes[k];
            uint64_t v;
            if (nd.kind == EX_COL) {
                v = ex_load(P.cols[nd.col], nd.tag, r);
            } else if (nd.kind == EX_LIT) {
                v = nd.lit;
            } else {
                uint64_t b = 0, a = 0;
#pragma unroll
                for (int j = 0; j < kExprStack; ++j) {
                    if (j == sp - 1) b = s[j];
                    if (j == sp - 2) a = s[j];
                }
                sp -= 2;
                v = ex_binary(nd.kind, nd.ctag, ex_conv(a, nd.ta, nd.ctag), ex_conv(b, nd.tb, nd.ctag));
            }
#pragma unroll
            for (int j = 0; j < kExprStack; ++j)
                if (j == sp) s[j] = v;
            ++sp;
        }
        bool ok = in;
        for (int c = 0; c < P.n_cols; ++c) {
            const uint8_t *vb = P.cols[c].valid;
            if (vb) ok = ok && ((vb[r >> 3] >> (r & 7)) & 1);
        }
        const uint64_t v = ok ? s[0] : 0;
        if (in) {
            if (root_size == 8) ((uint64_t *)out)[i] = v;
            else if (root_size == 4) ((uint32_t *)out)[i] = (uint32_t)v;
            else if (root_size == 2) ((uint16_t *)out)[i] = (uint16_t)v;
            else ((uint8_t *)out)[i] = (uint8_t)v;
        }
        if (out_valid) {
            const unsigned long long m = __ballot(ok);
            const uint64_t w = (base >> 6) + (threadIdx.x >> 6);
            if ((threadIdx.x & 63) == 0 && (w << 6) < n) out_valid[w] = m;
        }
    }
}

// ---------------------------------------------------------------------------
// host: types and the postfix program
// ---------------------------------------------------------------------------
int int_bits(int t) { return tag_size(t) * 8; }

int signed_of_bits(int bits) {
    return bits <= 8 ? PMH_I8 : bits <= 16 ? PMH_I16 : bits <= 32 ? PMH_I32 : PMH_I64;
}

[[noreturn]] void expr_unsupported(const std::string &what) {
    fail(PMH_E_NOTIMPL, "polars_metal: expression aggregate: " + what);
}

// Numeric supertype of two concrete dtypes (the rules in the file comment).
int supertype(int a, int b) {
    if (a == b) return a;
    const bool fa = a == PMH_F32 || a == PMH_F64, fb = b == PMH_F32 || b == PMH_F64;
    if (fa && fb) return PMH_F64;
    if (fa || fb) {
        const int f = fa ? a : b, i = fa ? b : a;
        if (f == PMH_F64) return PMH_F64;
        return int_bits(i) <= 16 ? PMH_F32 : PMH_F64;
    }
    const bool sa = tag_signed(a), sb = tag_signed(b);
    if (sa == sb) return int_bits(a) >= int_bits(b) ? a : b;
    const int s = sa ? a : b, u = sa ? b : a;
    if (int_bits(s) > int_bits(u)) return s;
    if (u == PMH_U64) expr_unsupported("a signed integer with UInt64 has no 64-bit supertype");
    return signed_of_bits(2 * int_bits(u));
}

// A subtree's type during lowering: a concrete dtype, or a literal whose
// type is still open (folded value kept).
struct ETy {
    int tag = -1;          // concrete dtype, or -1 for an open literal
    bool lit_float = false;
    int64_t ival = 0;
    double fval = 0;
};

bool int_fits(int64_t v, int tag) {
    switch (tag) {
    case PMH_I8: return v >= INT8_MIN && v <= INT8_MAX;
    case PMH_I16: return v >= INT16_MIN && v <= INT16_MAX;
    case PMH_I32: return v >= INT32_MIN && v <= INT32_MAX;
    case PMH_I64: return true;
    case PMH_U8: return v >= 0 && v <= UINT8_MAX;
    case PMH_U16: return v >= 0 && v <= UINT16_MAX;
    case PMH_U32: return v >= 0 && v <= (int64_t)UINT32_MAX;
    case PMH_U64: return v >= 0;
    default: return false;
    }
}

// The literal's bits in dtype `tag` (ints sign- or zero-extended as the
// kernel holds them).
uint64_t lit_bits(const ETy &l, int tag) {
    if (tag == PMH_F64) {
        double d = l.lit_float ? l.fval : (double)l.ival;
        uint64_t u;
        memcpy(&u, &d, 8);
        return u;
    }
    if (tag == PMH_F32) {
        float f = l.lit_float ? (float)l.fval : (float)l.ival;
        uint32_t u;
        memcpy(&u, &f, 4);
        return u;
    }
    if (l.lit_float) expr_unsupported("a float literal cannot take an integer type");
    if (!int_fits(l.ival, tag))
        expr_unsupported("integer literal " + std::to_string(l.ival) + " does not fit " + tag_name(tag));
    return (uint64_t)l.ival;
}

struct ExprBuilder {
    ExprProg prog{};
    std::vector<std::string> names;
    const std::function<int(const std::string &)> &col_tag;
    bool emit;
    int depth = 0, max_depth = 0;

    ExprBuilder(const std::function<int(const std::string &)> &ct, bool e) : col_tag(ct), emit(e) {}

    void push_node(const ExprNode &nd) {
        if (prog.n_nodes >= kMaxExprNodes) expr_unsupported("more than 31 nodes");
        if (nd.kind == EX_COL || nd.kind == EX_LIT) {
            ++depth;
        } else {
            depth -= 1;  // two operands out, one result in
        }
        if (depth > kExprStack) expr_unsupported("tree deeper than the evaluation stack");
        max_depth = std::max(max_depth, depth);
        prog.nodes[prog.n_nodes++] = nd;
    }

    // Types the subtree; with emit, appends its postfix nodes except an
    // open literal's (its parent places it once the type is known).
    ETy walk(const PlanExpr &e) {
        ETy t;
        if (e.kind == PlanExpr::COLUMN) {
            t.tag = col_tag(e.name);
            if (!ex_is_int(t.tag) && t.tag != PMH_F32 && t.tag != PMH_F64)
                expr_unsupported("column " + e.name + " of dtype " + tag_name(t.tag) + " in arithmetic");
            if (emit) {
                int ci = -1;
                for (size_t j = 0; j < names.size(); ++j)
                    if (names[j] == e.name) ci = (int)j;
                if (ci < 0) {
                    if ((int)names.size() >= kMaxExprCols) expr_unsupported("more than 16 columns");
                    ci = (int)names.size();
                    names.push_back(e.name);
                }
                ExprNode nd{};
                nd.kind = EX_COL;
                nd.tag = nd.ctag = nd.ta = nd.tb = t.tag;
                nd.col = ci;
                push_node(nd);
            }
            return t;
        }
        if (e.kind == PlanExpr::LIT_I64) { t.ival = e.ival; return t; }
        if (e.kind == PlanExpr::LIT_F64) { t.lit_float = true; t.fval = e.fval; return t; }
        // binary: type both sides first (without emitting) to place literals
        ETy l = type_of(*e.lhs), r = type_of(*e.rhs);
        const int op = e.op;
        if (l.tag < 0 && r.tag < 0) {  // fold two open literals
            ETy f;
            if (!l.lit_float && !r.lit_float && op != EX_DIV) {
                uint64_t a = (uint64_t)l.ival, b = (uint64_t)r.ival;
                f.ival = (int64_t)(op == EX_ADD ? a + b : op == EX_SUB ? a - b : a * b);
            } else {
                double a = l.lit_float ? l.fval : (double)l.ival, b = r.lit_float ? r.fval : (double)r.ival;
                f.lit_float = true;
                f.fval = op == EX_ADD ? a + b : op == EX_SUB ? a - b : op == EX_MUL ? a * b : a / b;
            }
            return f;
        }
        // concrete type of each side
        auto side = [&](const ETy &me, const ETy &other) -> int {
            if (me.tag >= 0) return me.tag;
            if (!me.lit_float) return other.tag;                           // int literal: the other's type
            return other.tag == PMH_F32 ? PMH_F32 : PMH_F64;               // float literal
        };
        const int tl = side(l, r), tr = side(r, l);
        int ctag;
        if (op == EX_DIV && ex_is_int(tl) && ex_is_int(tr)) ctag = PMH_F64;
        else ctag = supertype(tl, tr);
        // a literal placed as an operand takes the operand type it has
        // beside the other side (an int literal must fit it when integer)
        if (emit) {
            emit_side(*e.lhs, l, tl, ctag);
            emit_side(*e.rhs, r, tr, ctag);
            ExprNode nd{};
            nd.kind = op;
            nd.ctag = ctag;
            nd.tag = ctag;
            nd.ta = l.tag >= 0 ? l.tag : ctag;
            nd.tb = r.tag >= 0 ? r.tag : ctag;
            push_node(nd);
        } else {
            if (l.tag < 0) (void)lit_bits(l, tl);
            if (r.tag < 0) (void)lit_bits(r, tr);
        }
        ETy out;
        out.tag = ctag;
        return out;
    }

    ETy type_of(const PlanExpr &e) {
        ExprBuilder sub(col_tag, false);
        return sub.walk(e);
    }

    void emit_side(const PlanExpr &e, const ETy &t, int side_tag, int ctag) {
        if (t.tag >= 0) {
            walk(e);
            return;
        }
        // an open literal: it must hold its operand type (an integer
        // literal beside an integer column), then it is placed in ctag
        (void)lit_bits(t, side_tag);
        ExprNode nd{};
        nd.kind = EX_LIT;
        nd.tag = nd.ctag = nd.ta = nd.tb = ctag;
        nd.lit = lit_bits(t, ctag);
        push_node(nd);
    }
};

}  // namespace

int expr_result_tag(const PlanExpr &e, const std::function<int(const std::string &)> &col_tag) {
    ExprBuilder b(col_tag, false);
    ETy t = b.walk(e);
    if (t.tag < 0) expr_unsupported("a tree of literals only");
    return t.tag;
}

Column materialize_expr(const PlanExpr &e, const std::vector<Column> &cols, uint64_t n_rows,
                        const std::string &slab_name) {
    auto find = [&](const std::string &nm) -> const Column & {
        for (auto &c : cols)
            if (c.name == nm) return c;
        fail(PMH_E_KEY, "polars_metal: column not found: " + nm);
    };
    std::function<int(const std::string &)> col_tag = [&](const std::string &nm) { return find(nm).tag; };
    ExprBuilder b(col_tag, true);
    ETy t = b.walk(e);
    if (t.tag < 0) expr_unsupported("a tree of literals only");
    ExprProg &P = b.prog;
    P.n_cols = (int)b.names.size();
    bool any_valid = false;
    for (int j = 0; j < P.n_cols; ++j) {
        const Column &c = find(b.names[j]);
        P.cols[j].data = c.d_data;
        P.cols[j].valid = c.d_valid;
        P.cols[j].tag = c.tag;
        any_valid |= c.d_valid != nullptr;
    }
    if (b.depth != 1 || P.n_nodes < 1) fail(PMH_E_VALUE, "polars_metal: malformed expression program");
    Column out;
    out.name = slab_name;
    out.tag = t.tag;
    out.n_rows = n_rows;
    const size_t bytes = std::max<size_t>(8, (size_t)n_rows * tag_size(t.tag) + 8);
    void *d = slab(slab_name, bytes);
    uint64_t *dv = any_valid ? (uint64_t *)slab(slab_name + ":valid", std::max<uint64_t>(8, ceil_div(n_rows, 64) * 8))
                             : nullptr;
    out.d_data = d;
    out.d_valid = (const uint8_t *)dv;
    if (n_rows) {
        Runtime &R = rt();
        const uint64_t blocks = std::min<uint64_t>(ceil_div(n_rows, kExprBlock), 8192);
        KernelTimer kt("ex_eval");
        hipLaunchKernelGGL(ex_eval, dim3((unsigned)blocks), dim3(kExprBlock), 0, R.stream, P, n_rows, d, dv);
        PMH_HIP(hipGetLastError());
    }
    return out;
}

}  // namespace pmh
