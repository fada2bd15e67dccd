// pmh_table.h -- the global open-addressing table and the per-row /
// per-entry aggregation helpers shared by the group-by kernels.
#pragma once

#include "pmh_device.h"

namespace pmh {

// ---------------------------------------------------------------------------
// global hash table
// ---------------------------------------------------------------------------
struct GTable {
    uint64_t *klo;
    uint64_t *khi;      // KM_LANE128 only
    uint32_t *kstate;   // KM_LANE128 only: 0 empty, 1 being written, 2 published
    uint32_t *list;     // claimed slots in claim order
    uint32_t *ctrl;     // [0] groups claimed [1] overflow [2..3] side-slot claims [4] rows selected
    uint64_t mask;      // capacity - 1 (capacity is a power of two)
    uint64_t limit;     // distinct keys allowed before the pass overflows
    uint64_t list_cap;
    void *arr[kMaxArrs];
};

constexpr int kLdsProbe = 16;

// ---------------------------------------------------------------------------
// global table
// ---------------------------------------------------------------------------
__device__ __forceinline__ void g_register(const GTable &t, uint64_t slot) {
    uint32_t idx = __hip_atomic_fetch_add(gp<uint32_t>(t.ctrl), 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (idx < t.list_cap) gst<uint32_t>(t.list, idx, (uint32_t)slot);
    if (idx >= t.limit) __hip_atomic_fetch_or(gp<uint32_t>(t.ctrl) + 1, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ int64_t g_side_slot(const GTable &t, int special) {
    uint32_t s = (uint32_t)special - 1;
    uint64_t slot = t.mask + 1 + s;
    PMH_G uint32_t *claim = gp<uint32_t>(t.ctrl) + 2 + s;
    if (__hip_atomic_load(claim, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0) {
        uint32_t old = 0;
        __hip_atomic_compare_exchange_strong(claim, &old, 1u, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_AGENT);
        if (old == 0) g_register(t, slot);
    }
    return (int64_t)slot;
}

// 64-bit keys: the key word itself is claimed with one CAS, so no other
// field needs publishing (the state arrays are initialised before the pass).
__device__ __forceinline__ int64_t g_find64(const GTable &t, const Key &k, uint64_t h) {
    if (k.special) return g_side_slot(t, k.special);
    uint64_t s = h & t.mask;
    for (uint64_t p = 0; p <= t.mask; ++p) {
        uint64_t cur = __hip_atomic_load(gp<uint64_t>(t.klo) + s, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (cur == k.lo) return (int64_t)s;
        if (cur == kEmpty) {
            uint64_t old = kEmpty;
            __hip_atomic_compare_exchange_strong(gp<uint64_t>(t.klo) + s, &old, k.lo, __ATOMIC_RELAXED,
                                                 __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (old == kEmpty) { g_register(t, s); return (int64_t)s; }
            if (old == k.lo) return (int64_t)s;
        }
        s = (s + 1) & t.mask;
        if ((p & 63) == 63 && __hip_atomic_load(gp<uint32_t>(t.ctrl) + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
            return -1;
    }
    __hip_atomic_fetch_or(gp<uint32_t>(t.ctrl) + 1, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return -1;
}

// 128-bit lanes: claim word 0 -> 1 by CAS, write both key words, publish 2
// with release; readers acquire before comparing.  The claimer finishes in
// the same loop iteration, so a wave never waits on one of its own lanes.
__device__ __forceinline__ int64_t g_find128(const GTable &t, const Key &k, uint64_t h) {
    if (k.special) return g_side_slot(t, k.special);
    uint64_t s = h & t.mask;
    uint64_t p = 0;
    while (p <= t.mask) {
        PMH_G uint32_t *state = gp<uint32_t>(t.kstate) + s;
        uint32_t st = __hip_atomic_load(state, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
        if (st == 2u) {
            if (gld<uint64_t>(t.klo, s) == k.lo && gld<uint64_t>(t.khi, s) == k.hi) return (int64_t)s;
        } else if (st == 0u) {
            uint32_t old = 0;
            __hip_atomic_compare_exchange_strong(state, &old, 1u, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_AGENT);
            if (old == 0u) {
                gst<uint64_t>(t.klo, s, k.lo);
                gst<uint64_t>(t.khi, s, k.hi);
                __hip_atomic_store(state, 2u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
                g_register(t, s);
                return (int64_t)s;
            }
            continue;  // lost the race: re-read this slot
        } else {
            continue;  // being written by another wave: re-read
        }
        s = (s + 1) & t.mask;
        ++p;
        if ((p & 63) == 0 && __hip_atomic_load(gp<uint32_t>(t.ctrl) + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
            return -1;
    }
    __hip_atomic_fetch_or(gp<uint32_t>(t.ctrl) + 1, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return -1;
}

// ---------------------------------------------------------------------------
// per-row contribution
// ---------------------------------------------------------------------------
template <class S, class Slot, class Row>
__device__ __forceinline__ void add_row(S &s, const QueryDesc *__restrict__ q, Slot slot, const Row &row) {
    s.add32(q->layout.a_len, slot, 1u);
    const int nv = q->n_vals;
    for (int j = 0; j < nv; ++j) {
        const ValSpec &vs = q->vals[j];
        if (!row.valid(vs.col)) continue;
        s.add32(vs.a_cnt, slot, 1u);
        if (vs.kind == VK_F) {
            double x = row.f64(vs.col);
            if (vs.need & NEED_SUM) add_f64_exact(s, vs.a_limb, vs.a_res, slot, x);
            if (vs.need & NEED_MINMAX) {
                if (x != x) {
                    s.or32(vs.a_nan, slot, 1u);
                } else {
                    int64_t kk = f64_order_key(x);
                    s.min64(vs.a_min, slot, kk);
                    s.max64(vs.a_max, slot, kk);
                }
            }
        } else if (vs.kind == VK_I64) {
            int64_t x = row.i64(vs.col);
            if (vs.need & NEED_SUM) s.add64(vs.a_sum, slot, x);
            if (vs.need & NEED_MINMAX) { s.min64(vs.a_min, slot, x); s.max64(vs.a_max, slot, x); }
        } else {
            int32_t x = (int32_t)row.i64(vs.col);
            if (vs.need & NEED_SUM) s.add64(vs.a_sum, slot, (int64_t)x);
            if (vs.need & NEED_MINMAX) { s.min32(vs.a_min, slot, x); s.max32(vs.a_max, slot, x); }
        }
    }
}

// add_row into an LDS table, for queries with at most two value columns,
// with the value specs resolved once into LDS pointers: per row, no field
// of the query descriptor is read from memory.  A record's validity bit of
// value j is bit j of its flags; its value bits are raw (width of the tag).
struct LdsAgg2 {
    int nv;
    uint32_t nslots;  // slots per state array (the limb stride)
    uint32_t *len_p;
    int kind[2], need[2], tag[2];
    uint32_t *cnt_p[2], *nan_p[2];
    unsigned long long *limb_p[2], *sum_p[2];
    double *res_p[2];
    void *min_p[2], *max_p[2];
};

__device__ __forceinline__ LdsAgg2 lds_agg2(const QueryDesc *__restrict__ q, uint8_t *state, const uint32_t *lds_off,
                                            uint32_t nslots) {
    LdsAgg2 A;
    A.nv = q->n_vals;
    A.nslots = nslots;
    auto at = [&](int a) -> uint8_t * { return a >= 0 ? state + lds_off[a] : state; };
    A.len_p = (uint32_t *)at(q->layout.a_len);
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        const bool in = j < A.nv;
        const ValSpec &S = q->vals[in ? j : 0];
        A.kind[j] = in ? S.kind : VK_I64;
        A.need[j] = in ? S.need : 0;
        A.tag[j] = in ? S.tag : PMH_I64;
        const bool sum_f = (A.need[j] & NEED_SUM) && A.kind[j] == VK_F;
        const bool sum_i = (A.need[j] & NEED_SUM) && A.kind[j] != VK_F;
        const bool mm = (A.need[j] & NEED_MINMAX) != 0;
        A.cnt_p[j] = (uint32_t *)at(in ? S.a_cnt : -1);
        A.limb_p[j] = (unsigned long long *)at(sum_f ? S.a_limb : -1);
        A.res_p[j] = (double *)at(sum_f ? S.a_res : -1);
        A.sum_p[j] = (unsigned long long *)at(sum_i ? S.a_sum : -1);
        A.min_p[j] = at(mm ? S.a_min : -1);
        A.max_p[j] = at(mm ? S.a_max : -1);
        A.nan_p[j] = (uint32_t *)at(mm && A.kind[j] == VK_F ? S.a_nan : -1);
    }
    return A;
}

__device__ __forceinline__ void lds_agg2_add(const LdsAgg2 &A, uint32_t sl, uint32_t flags, uint64_t v0, uint64_t v1) {
    atomicAdd(A.len_p + sl, 1u);
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        if (j >= A.nv) break;
        if (!((flags >> j) & 1u)) continue;
        const uint64_t raw = j ? v1 : v0;
        atomicAdd(A.cnt_p[j] + sl, 1u);
        if (A.kind[j] == VK_F) {
            const double x = raw_to_f64(A.tag[j], raw);
            if (A.need[j] & NEED_SUM) lds_add_f64(A.limb_p[j], A.res_p[j], A.nslots, sl, x);
            if (A.need[j] & NEED_MINMAX) {
                if (x != x) {
                    atomicOr(A.nan_p[j] + sl, 1u);
                } else {
                    const long long kk = (long long)f64_order_key(x);
                    atomicMin((long long *)A.min_p[j] + sl, kk);
                    atomicMax((long long *)A.max_p[j] + sl, kk);
                }
            }
        } else if (A.kind[j] == VK_I64) {
            const int64_t x = raw_to_i64(A.tag[j], raw);
            if (A.need[j] & NEED_SUM) atomicAdd(A.sum_p[j] + sl, (unsigned long long)x);
            if (A.need[j] & NEED_MINMAX) {
                atomicMin((long long *)A.min_p[j] + sl, (long long)x);
                atomicMax((long long *)A.max_p[j] + sl, (long long)x);
            }
        } else {
            const int32_t x = (int32_t)raw_to_i64(A.tag[j], raw);
            if (A.need[j] & NEED_SUM) atomicAdd(A.sum_p[j] + sl, (unsigned long long)(int64_t)x);
            if (A.need[j] & NEED_MINMAX) {
                atomicMin((int *)A.min_p[j] + sl, x);
                atomicMax((int *)A.max_p[j] + sl, x);
            }
        }
    }
}

// Compile-time aggregate signatures of the common query shapes, so the
// per-row fold of the hottest kernels carries no branch on the plan:
//   1: sum/mean of an f64 column and min/max of an i32 column
//   2: sum of an i64 column
//   3: sum, min and max of one f64 column
//   4: sum of one f64 column
// (counts and lens ride along with every signature).  0: any shape with at
// most two value columns, resolved at run time.
struct AggSig {
    int nv, kind0, need0, tag0, kind1, need1, tag1;
};
constexpr AggSig kAggSigs[] = {
    {0, 0, 0, 0, 0, 0, 0},
    {2, VK_F, NEED_SUM, PMH_F64, VK_I32, NEED_MINMAX, PMH_I32},
    {1, VK_I64, NEED_SUM, PMH_I64, 0, 0, 0},
    {1, VK_F, NEED_SUM | NEED_MINMAX, PMH_F64, 0, 0, 0},
    {1, VK_F, NEED_SUM, PMH_F64, 0, 0, 0},
};
constexpr int kNumAggSigs = sizeof(kAggSigs) / sizeof(kAggSigs[0]);

// The signature of q (0 when none matches).
inline int agg_signature(const QueryDesc &q) {
    for (int s = 1; s < kNumAggSigs; ++s) {
        const AggSig &g = kAggSigs[s];
        if (q.n_vals != g.nv) continue;
        bool ok = true;
        for (int j = 0; j < g.nv; ++j) {
            const ValSpec &v = q.vals[j];
            const int kind = j ? g.kind1 : g.kind0, need = j ? g.need1 : g.need0, tag = j ? g.tag1 : g.tag0;
            ok = ok && v.kind == kind && v.need == need && v.tag == tag;
        }
        if (ok) return s;
    }
    return 0;
}

template <int KIND, int NEED, int TAG>
__device__ __forceinline__ void lds_fold_value(const LdsAgg2 &A, int j, uint32_t sl, bool valid, uint64_t raw) {
    if (!valid) return;
    atomicAdd(A.cnt_p[j] + sl, 1u);
    if constexpr (KIND == VK_F) {
        const double x = TAG == PMH_F64 ? __longlong_as_double((long long)raw) : raw_to_f64(TAG, raw);
        if constexpr ((NEED & NEED_SUM) != 0) lds_add_f64(A.limb_p[j], A.res_p[j], A.nslots, sl, x);
        if constexpr ((NEED & NEED_MINMAX) != 0) {
            if (x != x) {
                atomicOr(A.nan_p[j] + sl, 1u);
            } else {
                const long long kk = (long long)f64_order_key(x);
                atomicMin((long long *)A.min_p[j] + sl, kk);
                atomicMax((long long *)A.max_p[j] + sl, kk);
            }
        }
    } else if constexpr (KIND == VK_I64) {
        const int64_t x = raw_to_i64(TAG, raw);
        if constexpr ((NEED & NEED_SUM) != 0) atomicAdd(A.sum_p[j] + sl, (unsigned long long)x);
        if constexpr ((NEED & NEED_MINMAX) != 0) {
            atomicMin((long long *)A.min_p[j] + sl, (long long)x);
            atomicMax((long long *)A.max_p[j] + sl, (long long)x);
        }
    } else {
        const int32_t x = (int32_t)raw_to_i64(TAG, raw);
        if constexpr ((NEED & NEED_SUM) != 0) atomicAdd(A.sum_p[j] + sl, (unsigned long long)(int64_t)x);
        if constexpr ((NEED & NEED_MINMAX) != 0) {
            atomicMin((int *)A.min_p[j] + sl, x);
            atomicMax((int *)A.max_p[j] + sl, x);
        }
    }
}

// lds_agg2_add for signature SIG (0: resolved at run time).
template <int SIG>
__device__ __forceinline__ void lds_agg2_add_sig(const LdsAgg2 &A, uint32_t sl, uint32_t flags, uint64_t v0,
                                                 uint64_t v1) {
    if constexpr (SIG == 0) {
        lds_agg2_add(A, sl, flags, v0, v1);
    } else {
        constexpr AggSig G = kAggSigs[SIG];
        atomicAdd(A.len_p + sl, 1u);
        lds_fold_value<G.kind0, G.need0, G.tag0>(A, 0, sl, flags & 1u, v0);
        if constexpr (G.nv > 1) lds_fold_value<G.kind1, G.need1, G.tag1>(A, 1, sl, (flags >> 1) & 1u, v1);
    }
}

// Merge LDS slot ls into global slot gs: one atomic per touched array.
__device__ __forceinline__ void merge_lds_entry(GSink &g, uint64_t gs, LSink &l, uint32_t ls,
                                                const Layout &lay, const int32_t *ops) {
    for (int a = 0; a < lay.n_arrs; ++a) {
        const int op = ops[a];
        if (mo_size(op) == 4) {
            uint32_t v = *l.p<uint32_t>(a, ls);
            if (v == (uint32_t)mo_identity(op)) continue;
            if (op == MO_ADD32) g.add32(a, gs, v);
            else if (op == MO_MIN32) g.min32(a, gs, (int32_t)v);
            else if (op == MO_MAX32) g.max32(a, gs, (int32_t)v);
            else g.or32(a, gs, v);
        } else {
            uint64_t v = *l.p<uint64_t>(a, ls);
            if (v == mo_identity(op)) continue;
            if (op == MO_ADD64) g.add64(a, gs, (int64_t)v);
            else if (op == MO_ADDF64) g.addf(a, gs, __longlong_as_double((long long)v));
            else if (op == MO_MIN64) g.min64(a, gs, (int64_t)v);
            else g.max64(a, gs, (int64_t)v);
        }
    }
}


// ---------------------------------------------------------------------------
// output columns and group emission (shared by every finalising kernel)
// ---------------------------------------------------------------------------
constexpr int kOutMax = 64;
enum OutSrc : int32_t { OS_KEY = 0, OS_AGG = 1 };

struct OutCol {
    int32_t src, idx, op, tag;  // OS_KEY: idx = key field; OS_AGG: idx = value spec (-1: len)
    void *data;                 // element width of tag (Bool keys: one byte per row, packed later)
    uint64_t *valid;            // ballot-written bitmap words (hash path v1), or nullptr
    uint8_t *vbytes;            // one validity byte per row (partitioned path), or nullptr
    uint8_t *data_bits;         // Bool keys on the partitioned path: packed data
};

// Reads of per-group state arrays, from global memory or from LDS.
struct GReader {
    void *const *arr;
    template <class T> __device__ __forceinline__ T get(int a, uint64_t s) const { return gld<T>(arr[a], s); }
};
struct LReader {
    uint8_t *base;
    const uint32_t *off;
    template <class T> __device__ __forceinline__ T get(int a, uint32_t s) const {
        return ((const T *)(base + off[a]))[s];
    }
};

// Key field fi of a group whose key is (lo, hi, special) in the lane layout.
__device__ __forceinline__ void key_field(const QueryDesc *__restrict__ q, int fi, uint64_t lo, uint64_t hi,
                                          int special, uint64_t &bits, bool &valid) {
    const int mode = q->key_mode;
    if (mode == KM_RAW64) {
        if (special == 2) { bits = 0; valid = false; }
        else { bits = special == 1 ? kEmpty : lo; valid = true; }
        return;
    }
    if (special == 1) { lo = kEmpty; hi = mode == KM_LANE128 ? kEmpty : 0; }
    const KeyField &f = q->keys[fi];
    const int no = f.null_off;
    valid = no < 0 || ((no < 64 ? (lo >> no) : (hi >> (no - 64))) & 1) == 0;
    const int off = f.data_off;
    uint64_t v;
    if (off == 0) v = lo;
    else if (off < 64) v = (lo >> off) | (hi << (64 - off));
    else v = hi >> (off - 64);
    if (f.bits < 64) v &= (1ull << f.bits) - 1;
    if (f.xform != KX_RAW) {  // a narrowed field: back to the raw bits
        v += f.bias;
        if (f.xform == KX_SIGNED) v ^= 1ull << 63;
    }
    bits = valid ? v : 0;
}

// Exact f64 sums already converted for one group: the Sum and the Mean
// output of a column share one conversion of its limbs.
struct SumCache {
    int idx[2] = {-1, -1};
    double d[2] = {0.0, 0.0};
};

// Value of aggregate output oc for the group in slot s of reader rd (sc:
// the group's converted sums, or null).
template <class Rd, class Slot>
__device__ __forceinline__ void agg_value(const QueryDesc *__restrict__ q, const OutCol &oc, const Rd &rd, Slot s,
                                          uint64_t &bits, bool &valid, SumCache *sc = nullptr) {
    const Layout &lay = q->layout;
    if (oc.idx < 0) { bits = rd.template get<uint32_t>(lay.a_len, s); valid = true; return; }
    const ValSpec &vs = q->vals[oc.idx];
    const uint32_t cnt = rd.template get<uint32_t>(vs.a_cnt, s);
    bits = 0;
    valid = false;
    switch (oc.op) {
    case AGG_COUNT: bits = cnt; valid = true; return;
    case AGG_SUM:
    case AGG_MEAN: {
        double d = 0.0;
        int64_t si = 0;
        if (vs.kind == VK_F) {
            if (sc && sc->idx[0] == oc.idx) {
                d = sc->d[0];
            } else if (sc && sc->idx[1] == oc.idx) {
                d = sc->d[1];
            } else {
                int64_t limb[kLimbs];
#pragma unroll
                for (int j = 0; j < kLimbs; ++j) limb[j] = rd.template get<int64_t>(vs.a_limb + j, s);
                d = limbs_to_double(limb, rd.template get<double>(vs.a_res, s));
                if (sc) {
                    const int e = sc->idx[0] < 0 ? 0 : 1;
                    sc->idx[e] = oc.idx;
                    sc->d[e] = d;
                }
            }
        } else {
            si = rd.template get<int64_t>(vs.a_sum, s);
        }
        if (oc.op == AGG_SUM) {
            valid = true;
            if (vs.kind == VK_F) {
                if (oc.tag == PMH_F32) bits = (uint32_t)__float_as_uint((float)d);
                else bits = (uint64_t)__double_as_longlong(d);
            } else {
                bits = (uint64_t)si;  // an I32 output keeps the low 32 bits: wrapping sum
            }
        } else {
            valid = cnt > 0;
            double m = 0.0;
            if (valid) m = (vs.kind == VK_F ? d : (double)si) / (double)cnt;
            if (oc.tag == PMH_F32) bits = (uint32_t)__float_as_uint((float)m);
            else bits = (uint64_t)__double_as_longlong(m);
        }
        return;
    }
    default: {  // AGG_MIN / AGG_MAX
        valid = cnt > 0;
        if (!valid) return;
        const int a = oc.op == AGG_MIN ? vs.a_min : vs.a_max;
        if (vs.kind == VK_F) {
            double d = rd.template get<uint32_t>(vs.a_nan, s) ? __longlong_as_double(0x7FF8000000000000ll)
                                                               : f64_from_order_key(rd.template get<int64_t>(a, s));
            if (oc.tag == PMH_F32) bits = (uint32_t)__float_as_uint((float)d);
            else bits = (uint64_t)__double_as_longlong(d);
        } else if (vs.kind == VK_I64) {
            bits = (uint64_t)rd.template get<int64_t>(a, s);
        } else {
            bits = (uint32_t)rd.template get<int32_t>(a, s);
        }
        return;
    }
    }
}

__device__ __forceinline__ void store_tag(void *data, uint64_t i, int tag, uint64_t bits) {
    switch (tag_size(tag)) {
    case 8: gst<uint64_t>(data, i, bits); break;
    case 4: gst<uint32_t>(data, i, (uint32_t)bits); break;
    case 2: gst<uint16_t>(data, i, (uint16_t)bits); break;
    case 1: gst<uint8_t>(data, i, (uint8_t)bits); break;
    default: gst<uint8_t>(data, i, (uint8_t)(bits & 1)); break;  // Bool: a byte, packed later
    }
}

// Writes every output column of one group at output row o (byte validity).
template <class Rd, class Slot>
__device__ __forceinline__ void emit_group(const QueryDesc *__restrict__ q, const OutCol *outs, int n_out,
                                           const Rd &rd, Slot s, uint64_t lo, uint64_t hi, int special,
                                           uint64_t o) {
    SumCache sc;
    for (int i = 0; i < n_out; ++i) {
        const OutCol &oc = outs[i];
        uint64_t bits;
        bool valid;
        if (oc.src == OS_KEY) key_field(q, oc.idx, lo, hi, special, bits, valid);
        else agg_value(q, oc, rd, s, bits, valid, &sc);
        store_tag(oc.data, o, oc.tag, bits);
        if (oc.vbytes) gst<uint8_t>(oc.vbytes, o, valid ? 1 : 0);
    }
}

// Host-lowered form of a conjunctive predicate for the specialised scan: a
// row is selected iff, for every r < nr, column rcol[r] is valid and its
// int64 image (the value itself, or its TotalOrd image when rflt[r]) lies in
// [rlo[r], rhi[r]].
struct FastScan {
    int32_t enabled, nr;
    int32_t rcol[2], rflt[2];
    int64_t rlo[2], rhi[2];
    uint32_t ring_off;       // byte offset of the per-wave row rings in LDS
    int32_t rows;            // rows per lane per step (4 or 8)
    int32_t pad;
};

// Lowers q's predicate into fs (false: another shape).
bool lower_fast_scan(const QueryDesc &q, FastScan &fs);

// ---------------------------------------------------------------------------
// host: a query prepared once, run by one of the group-by engines
// ---------------------------------------------------------------------------
struct Prep {
    QueryDesc q;
    OutCol outs[kOutMax];  // src / idx / op / tag; engines fill the pointers
    int n_out = 0;
    std::vector<std::string> names;
    std::vector<int> limb_k, limb_base;
    std::string sig;
    bool wide = false;
    uint64_t n_rows = 0;
    std::vector<Column> dcols;    // the query's columns, in QueryDesc.cols order
    float feat[PMH_MLP_IN] = {};  // the learned sizer's features, once the sketch ran
    pmh_sketch_info *sketch_out = nullptr;  // sketch-only runs: filled with the whole sketch
    int route = 0;                // 1: the learned sizer sent the query to the sort engine
    int sizer_choice = 0;         // pmh_exec_stats::sizer_choice
    bool overflowed = false;      // the partitioned engine's fallback table overflowed
    bool narrow_miss = false;     // a key fell outside the sampled narrowing ranges: run again, narrowed exactly
    bool narrow_exact = false;    // narrow from a pass over every row, not a sample
    double d_used = 0;            // the distinct estimate that run was sized for
    double d_retry = 0;           // > 0: size the partitioned engine for this many keys
};

// Sample sketch of the partitioned engine.  The estimators and features
// below are shared by the host (the batch plans from them) and the device
// (the single-query sketch block evaluates them next to its counts).
struct Sketch {
    uint32_t sample_rows = 0, sample_size = 0, distinct = 0, f1 = 0, f2 = 0, dropped = 0;
    uint64_t n_rows = 0;
    double n_sel = 0, F2 = 0, shl_num = 0, shl_den = 0;
};
Sketch sketch_from(uint32_t s_sel, uint32_t n_sample, uint64_t n, uint32_t d, uint32_t f1, uint32_t f2,
                   uint32_t dropped, const std::vector<uint32_t> &hist);

// The sketch fields that do not need the count histogram.  A dropped row's
// key found no slot within the probe cap, so none of that key's rows were
// counted.  Keys arriving late meet a fuller table, so dropped rows are
// mostly rare keys: each counts as a singleton.
__host__ __device__ inline Sketch sketch_head(uint32_t s_sel, uint32_t n_sample, uint64_t n, uint32_t d,
                                              uint32_t f1, uint32_t f2, uint32_t dropped) {
    Sketch k;
    k.sample_rows = s_sel;
    k.sample_size = n_sample;
    k.n_rows = n;
    k.distinct = d + dropped;
    k.f1 = f1 + dropped;
    k.f2 = f2;
    k.dropped = dropped;
    k.n_sel = n_sample ? (double)n * s_sel / n_sample : 0.0;
    return k;
}
// Fraction of the selection the sample covers.
__host__ __device__ inline double sketch_qf(const Sketch &k) {
    return k.n_sel > 0 ? fmin(1.0, k.sample_rows / k.n_sel) : 1.0;
}
// Bucket c of the count histogram (h keys seen c times, dropped rows
// included in c = 1) adds to the F2 moment and the Shlosser sums.
// (1 - q)^c by squaring: c < 1024 takes ten steps, against a general pow's
// range reduction and polynomial on the sketch block's critical path.
__host__ __device__ inline double sketch_ipow(double b, int e) {
    double r = 1.0;
    for (; e; e >>= 1, b *= b)
        if (e & 1) r *= b;
    return r;
}
__host__ __device__ inline void sketch_bucket(int c, double h, double qf, double &F2, double &num, double &den) {
    if (h == 0) return;
    const double t = sketch_ipow(1.0 - qf, c - 1);  // (1 - q)^(c - 1)
    F2 += (double)c * c * h;
    num += t * (1.0 - qf) * h;
    den += c * qf * t * h;
}

// Distinct keys among the selected rows: the geometric mean of GEE
// (Charikar et al., 2000; it undershoots on skewed keys) and Shlosser's
// estimator (Haas et al., 1995; it overshoots), floored at the sample's own
// distinct count and at Chao's lower bound (Chao, 1984).  A sample of
// singletons with no key seen twice (a near-unique selection), or a
// fingerprint table that still dropped most of it, extrapolates linearly; a
// few doubletons already say that keys repeat.
__host__ __device__ inline double gee_estimate(const Sketch &k) {
    if (k.sample_rows == 0) return 0.0;
    const double s = k.sample_rows, d = k.distinct, f1 = k.f1;
    return sqrt(fmax(k.n_sel, s) / s) * f1 + (d - f1);
}
__host__ __device__ inline double shlosser_estimate(const Sketch &k) {
    if (k.sample_rows == 0) return 0.0;
    return k.shl_den > 0 ? k.distinct + k.f1 * k.shl_num / k.shl_den : gee_estimate(k);
}
// The estimate from GEE and Shlosser's values (gee, shl) of the same sketch.
__host__ __device__ inline double plain_from(const Sketch &k, double gee, double shl) {
    if (k.sample_rows == 0) return 0.0;
    const double s = k.sample_rows, d = k.distinct;
    if (2.0 * k.dropped > s || (k.f2 == 0 && k.f1 >= 0.95 * s)) return fmax(d, k.n_sel * d / s);
    if (k.n_sel <= s * 1.0000001) return d;  // the sample is the whole selection
    // Chao's lower bound (bias-corrected Chao1) lifts the mean where keys
    // repeat a few times each across a selection the sample barely covers
    // (there GEE undershoots far more than Shlosser overshoots)
    const double chao = fmin(k.n_sel, d + (double)k.f1 * (k.f1 > 0 ? k.f1 - 1.0 : 0.0) / (2.0 * (k.f2 + 1.0)));
    return fmax(fmax(d, sqrt(fmax(gee, 1.0) * fmax(shl, 1.0))), chao);
}
__host__ __device__ inline double plain_distinct_estimate(const Sketch &k) {
    return plain_from(k, gee_estimate(k), shlosser_estimate(k));
}

// The learned sizer's PMH_MLP_IN features of a sketch: scale-free
// quantities, each roughly in [0, 1].  Feature i is lg(a) = log2(a + 1) / 32,
// a ratio a / b, or a constant a; sizer_feature_arg gives its kind and
// operands, so the device evaluates all of them at once, one per lane.
enum SizerFeatureKind : int { SF_LG = 0, SF_RATIO = 1, SF_CONST = 2 };
__host__ __device__ inline int sizer_feature_arg(int i, const Sketch &k, uint32_t null_keys, int lane_bits,
                                                 double gee, double shl, double plain, double &a, double &b) {
    const double s = fmax(1.0, (double)k.sample_rows);
    b = s;
    switch (i) {
    case 0: a = (double)k.n_rows; return SF_LG;
    case 1:
        a = (double)k.sample_rows;
        b = (double)k.sample_size;
        if (!k.sample_size) { a = 0.0; return SF_CONST; }
        return SF_RATIO;
    case 2: a = k.distinct; return SF_LG;
    case 3: a = k.f1; return SF_RATIO;
    case 4: a = k.f2; return SF_RATIO;
    case 5: a = k.distinct; return SF_RATIO;
    case 6: a = gee; return SF_LG;
    case 7: a = shl; return SF_LG;
    case 8: a = plain; return SF_LG;
    case 9: a = k.F2; b = s * s; return SF_RATIO;
    case 10: a = null_keys; return SF_RATIO;
    case 11: a = k.dropped; return SF_RATIO;
    case 12: a = lane_bits <= 33 ? 1.0 : 0.0; return SF_CONST;
    case 13: a = lane_bits > 33 && lane_bits <= 65 ? 1.0 : 0.0; return SF_CONST;
    case 14: a = lane_bits > 65 ? 1.0 : 0.0; return SF_CONST;
    case 15: a = k.n_sel; return SF_LG;
    default: a = 0.0; return SF_CONST;
    }
}
__host__ __device__ inline float sizer_lg(double x) { return (float)(log2(fmax(0.0, x) + 1.0) / 32.0); }
__host__ __device__ inline float sizer_feature_value(int kind, double a, double b) {
    return kind == SF_LG ? sizer_lg(a) : kind == SF_RATIO ? (float)(a / b) : (float)a;
}
inline void sizer_features(const Sketch &k, uint32_t null_keys, int lane_bits, float *f) {
    const double gee = gee_estimate(k), shl = shlosser_estimate(k), plain = plain_from(k, gee, shl);
    for (int i = 0; i < PMH_MLP_IN; ++i) {
        double a, b;
        const int kind = sizer_feature_arg(i, k, null_keys, lane_bits, gee, shl, plain, a, b);
        f[i] = sizer_feature_value(kind, a, b);
    }
}

void prepare_query(Prep &P, const GroupByPlan &plan, uint64_t n_rows, const pmh_column *cols, uint32_t n_cols);
pmh_result *run_hash_global(Prep &P, const GroupByPlan &plan, uint32_t flags);
// Lowers plan's predicate over P's columns into q (replacing q's predicate).
void lower_predicate_into(QueryDesc &q, const GroupByPlan &plan, const Prep &P);

// One query of a batch (pmh_batch.hip).
struct BatchQueryIn {
    uint64_t n_rows;
    const pmh_column *cols;
    uint32_t n_cols;
    const char *predicate;  // PredicateAst JSON, or nullptr
};
void run_batch(const GroupByPlan &tmpl, const std::vector<BatchQueryIn> &qs, uint32_t flags,
               std::vector<pmh_result *> &results, pmh_batch_stats *stats);

// Sort engine: radix sort of the selected rows' key lanes, then runs fold.
pmh_result *run_sort(Prep &P, const GroupByPlan &plan, uint32_t flags);
// Per-group state arrays of nslots entries carved from one slab (limb k > 0
// of an exact accumulator sits nslots entries after limb k-1).
void carve_state(const Prep &P, uint8_t *base, size_t &o, uint64_t nslots, void **arr);
size_t state_bytes(const Prep &P, uint64_t nslots);
// Returns nullptr when the query shape does not fit the partitioned engine.
pmh_result *run_hash_partitioned(Prep &P, const GroupByPlan &plan, uint32_t flags);

}  // namespace pmh
