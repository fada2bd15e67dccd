// pmh_scanfast.h -- specialised phase-1 scan of the partitioned hash engine.
//
// Included by pmh_hashagg.hip after HLaunch, the spill helpers and the
// generic scan.  It serves the common query shape: the predicate is a
// conjunction of column-vs-literal comparisons (lowered on the host to at
// most two closed intervals on an order-preserving int64 image of a column),
// the key is one 64-bit column or a lane of at most 64 bits over at most four
// columns, and at most two columns are aggregated.
//
// What makes it cheaper than the generic scan:
//  * Every descriptor field, LDS array address and spill pointer is read
//    into registers once per block.
//  * Column loads have one form for every width: the aligned 8-byte word
//    holding the value (or validity bit) is loaded, and the value is cut out
//    with a shift and mask whose parameters are wave-uniform.  A naturally
//    aligned value never straddles words, and an aligned word never crosses
//    a page, so the load is always mapped.  With no per-width switch and no
//    lane predicate around a load, nothing forces the compiler to wait for
//    one load before issuing the next: a step's loads are all in flight
//    together and one wait covers them.
//  * Rows are compacted before the per-row work.  Each wave evaluates the
//    predicate on coalesced loads of the predicate columns only (issued one
//    step ahead), queues the offsets of the selected rows in a per-wave LDS
//    ring, and whenever 64 are queued runs one dense round over them: gather
//    the key and value words for those rows, hash, probe the hot table, fold
//    hot rows into LDS and append the rest to the block's spill region.  At
//    30% selectivity that removes three quarters of the idle lanes from the
//    expensive part, and key/value cache lines holding no selected row are
//    never fetched.
#pragma once

#include "pmh_words.h"

namespace pmh {

// Per-wave ring of selected-row offsets: at most 63 carried rows plus one
// row step of 64 new ones are queued at any time.
constexpr uint32_t kRing = 127;


// SIG: the query's compile-time aggregate signature (0: resolved at run
// time), so the fold of hot rows carries no branch on the plan.  NR2: the
// predicate may test two columns (else at most one, and the second
// column's descriptor takes no registers).
// KM: the key: kScanLane (a lane of at most 64 bits), kScanRaw64 (one
// 64-bit column) or kScanLaneChecked (a lane narrowed from a sample: keys
// outside its ranges are flagged).
constexpr int kScanLane = 0, kScanRaw64 = 1, kScanLaneChecked = 2;
template <int KM, int U, int SIG, bool NR2>
__global__ __launch_bounds__(1024) void hs_scan_fast(const HLaunch *__restrict__ L) {
    constexpr bool RAW64 = KM == kScanRaw64;
    constexpr bool nchk = KM == kScanLaneChecked;
    extern __shared__ __align__(16) uint8_t smem[];
    __shared__ uint32_t s_cursor, s_sel;
    const QueryDesc *__restrict__ q = &L->q;
    const uint32_t CL = L->hot_slots;
    // launched with 1024 threads
    constexpr uint32_t bs = 1024;
    const uint32_t tid = threadIdx.x;
    const uint32_t lane = tid & 63;
    const unsigned long long below = (1ull << lane) - 1;
    uint64_t *lkeys = (uint64_t *)smem;
    uint8_t *state = (uint8_t *)(lkeys + CL);
    uint32_t *ring = (uint32_t *)(smem + L->fs.ring_off) + (tid >> 6) * kRing;

    // The block's rows [r_begin, r_end), addressed from r_begin in 32 bits
    // (a block spans fewer than 2^25 rows); its spill region starts at
    // r_begin too.
    const uint64_t n = q->n_rows;
    const uint64_t r_begin = (uint64_t)blockIdx.x * L->chunk;
    const uint64_t r_end = min(n, r_begin + L->chunk);
    const uint32_t span = r_end > r_begin ? (uint32_t)(r_end - r_begin) : 0;
    const uint32_t r_last = span ? span - 1 : 0;  // in-range stand-in for rows past the end

    // ---- descriptor, once per block
    const int nr = NR2 ? L->fs.nr : min(L->fs.nr, 1);
    BlockCol pc[2];
    int pflt[2];
    int64_t plo[2], phi[2];
#pragma unroll
    for (int r = 0; r < 2; ++r) {
        pc[r] = block_col(role_col(q, r < nr ? L->fs.rcol[r] : 0), r_begin);
        pflt[r] = L->fs.rflt[r];
        plo[r] = L->fs.rlo[r];
        phi[r] = L->fs.rhi[r];
    }
    const int nk = RAW64 ? 1 : (q->key_mode == KM_NONE ? 0 : q->n_keys);
    const bool allones = q->allones_possible != 0;
    BlockCol kc[4];
    int knull[4], kdata[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        kc[i] = block_col(role_col(q, i < nk ? q->keys[i].col : 0), r_begin);
        knull[i] = i < nk ? q->keys[i].null_off : 0;
        kdata[i] = i < nk ? q->keys[i].data_off : 0;
    }
    const int nv = SIG > 0 ? kAggSigs[SIG].nv : q->n_vals;
    const LdsAgg2 A = lds_agg2(q, state, L->lds1_off, CL);
    BlockCol vc[2];
    int vwid[2];
    void *sval[2];
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        const bool in = j < nv;
        vc[j] = block_col(role_col(q, in ? q->vals[j].col : 0), r_begin);
        vwid[j] = in ? L->vw[j] : 8;
        sval[j] = in ? (void *)((uint8_t *)L->spill.val[j] + r_begin * (uint64_t)L->vw[j]) : nullptr;
    }
    // a signature fixes the value columns' tags and a raw key is 8 bytes, so
    // their widths, shifts and masks are constants rather than live registers
    if constexpr (SIG > 0) {
        constexpr AggSig G = kAggSigs[SIG];
#pragma unroll
        for (int j = 0; j < G.nv; ++j) {
            const int t = j ? G.tag1 : G.tag0;
            const int w = tag_size(t);
            vc[j].tag = t;
            vc[j].bshift = w == 8 ? 6 : w == 4 ? 5 : w == 2 ? 4 : 3;
            vc[j].mask = w == 8 ? ~0ull : (1ull << (8 * w)) - 1;
            vwid[j] = w;
        }
    }
    if constexpr (RAW64) {
        kc[0].bshift = 6;
        kc[0].mask = ~0ull;
    }
    uint64_t *sklo = L->spill.klo + r_begin;
    uint16_t *sflags = L->spill.flags + r_begin;

    // ---- LDS: hot keys and state identities
    if (tid == 0) { s_cursor = 0; s_sel = 0; }
    for (uint32_t i = tid; i < CL; i += bs) lkeys[i] = L->hot_lo[i];
    LSink ls{state, L->lds1_off, CL};
    for (int a = 0; a < q->layout.n_arrs; ++a) {
        const int op = q->layout.init[a];
        if (mo_size(op) == 4) {
            const uint32_t id = (uint32_t)mo_identity(op);
            for (uint32_t i = tid; i < CL; i += bs) *ls.p<uint32_t>(a, i) = id;
        } else {
            const uint64_t id = mo_identity(op);
            for (uint32_t i = tid; i < CL; i += bs) *ls.p<uint64_t>(a, i) = id;
        }
    }
    __syncthreads();

    // every lane of a wave runs the same steps and rounds, so ballots see
    // whole waves
    const uint32_t steps = (span + U * bs - 1) / (U * bs);
    uint32_t qhead = 0, qn = 0;  // wave-uniform ring state
    uint32_t wave_sel = 0;
    uint64_t nmiss = 0;  // nchk: bits of selected keys outside the sampled ranges

    // A step covers U * 1024 rows; wave w takes the U * 64 contiguous rows
    // from w * U * 64, 64 per sub-step, so a wave's cache lines are its own.
    const uint32_t wslice = (tid >> 6) * (U * 64);
    // Predicate words run one step ahead: the loads for step it + 1 are in
    // flight while step it's rounds run.
    uint64_t pw[2][U], pv[2][U];
    auto load_pred = [&](uint32_t step) __attribute__((always_inline)) {
        const uint32_t rs = step * (U * bs) + wslice + lane;
#pragma unroll
        for (int r = 0; r < 2; ++r) {
            if (r >= nr) continue;
#pragma unroll
            for (int u = 0; u < U; ++u) pw[r][u] = word_at(pc[r].base, vpos(pc[r], min(rs + u * 64, r_last)));
            // columns without a validity bitmap load no validity words
            if (pc[r].vstride) {
#pragma unroll
                for (int u = 0; u < U; ++u) pv[r][u] = word_at(pc[r].vbase, bpos(pc[r], min(rs + u * 64, r_last)));
            }
        }
    };
    if (steps) load_pred(0);

    // one dense round: the rows (one per active lane) gathered, hashed,
    // folded into the hot table or appended to the tail spill
    auto dense = [&](const uint32_t row, const bool act) __attribute__((always_inline)) {
        // gathers: every word first, then the cuts
        // (columns without a validity bitmap load no validity word)
        uint64_t kw[4], kv[4], vw_[2], vv_[2];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            kw[i] = 0;
            kv[i] = ~0ull;
            if (i >= nk) continue;
            kw[i] = word_at_nt(kc[i].base, vpos(kc[i], row));
            if (kc[i].vstride) kv[i] = word_at(kc[i].vbase, bpos(kc[i], row));
        }
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            vw_[j] = 0;
            vv_[j] = ~0ull;
            if (j >= nv) continue;
            vw_[j] = word_at_nt(vc[j].base, vpos(vc[j], row));
            if (vc[j].vstride) vv_[j] = word_at(vc[j].vbase, bpos(vc[j], row));
        }
        uint64_t key = 0;
        int special = 0;
        if (RAW64) {
            if (kc[0].vstride && !cut(kv[0], bpos(kc[0], row), 1)) special = 2;
            else {
                key = kw[0];
                special = key == kEmpty ? 1 : 0;
            }
        } else {
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                if (i >= nk) continue;
                const bool ok = !kc[i].vstride || cut(kv[i], bpos(kc[i], row), 1);
                const uint64_t raw = cut(kw[i], vpos(kc[i], row), kc[i].mask);
                if constexpr (nchk) {
                    uint64_t over;
                    const uint64_t kd = key_data_over(q->keys[i], raw, over);
                    key |= ok ? kd << kdata[i] : (knull[i] >= 0 ? 1ull << knull[i] : 0);
                    nmiss |= (ok ? over : (uint64_t)(knull[i] < 0)) & (act ? ~0ull : 0ull);
                } else {
                    key |= ok ? key_data(q->keys[i], raw) << kdata[i] : (knull[i] >= 0 ? 1ull << knull[i] : 0);
                }
            }
            special = (allones && key == kEmpty) ? 1 : 0;
        }
        uint64_t vr[2] = {0, 0};
        uint32_t vv = 0;  // bit j: value j valid
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            if (j >= nv) continue;
            vr[j] = cut(vw_[j], vpos(vc[j], row), vc[j].mask);
            const bool ok = !vc[j].vstride || cut(vv_[j], bpos(vc[j], row), 1);
            vv |= (uint32_t)ok << j;
        }
        // hot-table probe
        // the all-ones and null keys fold into the hot table's side
        // slots instead of spilling (they would otherwise pile up on
        // one partition slot of the tail)
        int32_t slot = -1;
        if (act && special && CL) slot = (int32_t)hot_side_slot(special, CL);
        if (act && special == 0 && CL) slot = hot_find(lkeys, key, hash_key(key, 0), CL);
        // hot rows: fold into LDS
        if (slot >= 0) lds_agg2_add_sig<SIG>(A, (uint32_t)slot, vv, vr[0], vr[1]);
        // tail rows: wave-contiguous append
        const bool tail = act && slot < 0;
        const unsigned long long m = __ballot(tail);
        if (m) {
            const int leader = __ffsll((long long)__ballot(1)) - 1;
            uint32_t base = 0;
            if ((int)lane == leader) base = atomicAdd(&s_cursor, (uint32_t)__popcll(m));
            base = __shfl(base, leader);
            if (tail) {
                const uint32_t pos = base + (uint32_t)__popcll(m & below);
                uint16_t fl = special == 1 ? kFlagAllOnes : (special == 2 ? kFlagNull : 0);
                fl |= (uint16_t)vv;
#pragma unroll
                for (int j = 0; j < 2; ++j)
                    if (j < nv) gstore_w(sval[j], vwid[j], pos, vr[j]);
                gst<uint64_t>(sklo, pos, key);
                gst<uint16_t>(sflags, pos, fl);
            }
        }
    };
    for (uint32_t it = 0; it <= steps; ++it) {
        const bool drain = it == steps;
        uint32_t selbits = 0;  // bit u: row u of this lane passes
        const uint32_t r0 = it * (U * bs) + wslice + lane;
        if (!drain) {
#pragma unroll
            for (int u = 0; u < U; ++u) selbits |= (uint32_t)(r0 + u * 64 < span) << u;
#pragma unroll
            for (int r = 0; r < 2; ++r) {
                if (r >= nr) continue;
                uint64_t raw[U];
                uint32_t vb = (1u << U) - 1;
#pragma unroll
                for (int u = 0; u < U; ++u) raw[u] = cut(pw[r][u], vpos(pc[r], min(r0 + u * 64, r_last)), pc[r].mask);
                if (pc[r].vstride) {
                    vb = 0;
#pragma unroll
                    for (int u = 0; u < U; ++u)
                        vb |= (uint32_t)cut(pv[r][u], bpos(pc[r], min(r0 + u * 64, r_last)), 1) << u;
                }
                int64_t img[U];
                pred_images<U>(pc[r].tag, pflt[r] != 0, raw, img);
#pragma unroll
                for (int u = 0; u < U; ++u)
                    if (img[u] < plo[r] || img[u] > phi[r]) vb &= ~(1u << u);
                selbits &= vb;
            }
            if (it + 1 < steps) load_pred(it + 1);
        }
#pragma unroll 1
        for (int u = 0; u < (drain ? 1 : U); ++u) {
            if constexpr (!RAW64) {
                if (nr == 0) {
                    // no predicate: every row is selected, so each sub-step's
                    // 64 rows are a dense round as they stand (no ring)
                    if (drain) break;
                    const bool s = (selbits >> u) & 1;
                    wave_sel += (uint32_t)__popcll(__ballot(s));
                    dense(s ? r0 + u * 64 : 0, s);
                    continue;
                }
            }
            if (!drain) {
                const bool s = (selbits >> u) & 1;
                const unsigned long long M = __ballot(s);
                if (s) {
                    uint32_t pos = qhead + qn + (uint32_t)__popcll(M & below);
                    pos = pos >= kRing ? pos - kRing : pos;
                    ring[pos] = r0 + u * 64;
                }
                const uint32_t c = (uint32_t)__popcll(M);
                qn += c;
                wave_sel += c;
            }
#pragma unroll 1
            while (qn >= 64 || (drain && qn > 0)) {
                // ---- one dense round over the oldest min(qn, 64) queued rows
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                const uint32_t cnt = qn < 64 ? qn : 64;
                const bool act = lane < cnt;
                uint32_t rp = qhead + lane;
                rp = rp >= kRing ? rp - kRing : rp;
                const uint32_t off = ring[rp];
                const uint32_t row = act ? off : 0;
                qhead = qhead + cnt >= kRing ? qhead + cnt - kRing : qhead + cnt;
                qn -= cnt;
                dense(row, act);
            }
        }
    }
    if (lane == 0) atomicAdd(&s_sel, wave_sel);
    // a selected key outside the sampled narrowing ranges: the query runs again
    if constexpr (nchk) {
        if (__ballot(nmiss != 0) && lane == 0) atomicOr(&L->stats[ST_NARROW_MISS], 1u);
    }
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
        if (*ls.p<uint32_t>(a_len, i) == 0) continue;
        merge_lds_entry(g, i, ls, i, q->layout, q->layout.init);
    }
}

}  // namespace pmh
