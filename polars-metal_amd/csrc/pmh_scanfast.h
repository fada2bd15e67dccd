// pmh_scanfast.h -- specialised phase-1 scan of the partitioned hash engine.
//
// Included by pmh_hashagg.hip after HLaunch, ScanCtx and the spill helpers.
// It serves the common query shape: the predicate is a conjunction of
// column-vs-literal comparisons (lowered on the host to at most two closed
// intervals on an order-preserving int64 image of the column), the key is
// one 64-bit column or a lane of at most 64 bits over at most four columns,
// and at most two columns are aggregated.  The generic scan re-reads the
// query descriptor with scalar loads for every row; each of those waits
// (s_waitcnt lgkmcnt) also drains the block's outstanding LDS atomics.  Here
// every descriptor field, LDS array address and spill pointer is read once
// into registers before the row loop, rows are handled four per lane per
// step, and one LDS cursor reservation covers the tail rows of all four.
#pragma once

namespace pmh {

// Order-preserving int64 image of a double under Polars' TotalOrd (the order
// of tot_cmp): -0.0 and +0.0 map together, every NaN maps above +inf.
__host__ __device__ inline int64_t tot_key_bits(int64_t b, bool is_nan) {
    if (is_nan) return 0x7FFFFFFFFFFFFFFFll;
    if ((b & 0x7FFFFFFFFFFFFFFFll) == 0) return 0;
    return b >= 0 ? b : (b ^ 0x7FFFFFFFFFFFFFFFll);
}
__device__ __forceinline__ int64_t tot_key(double x) { return tot_key_bits(__double_as_longlong(x), x != x); }

// Exact-sum sink over two LDS arrays addressed directly (limb k of slot s is
// limb[k * CL + s]).
struct FSink {
    unsigned long long *limb;
    double *res;
    uint32_t CL;
    __device__ __forceinline__ void addf(int, uint32_t s, double v) { unsafeAtomicAdd(res + s, v); }
    __device__ __forceinline__ void addlimb(int, int k, uint32_t s, int64_t v) {
        atomicAdd(limb + (uint32_t)k * CL + s, (unsigned long long)v);
    }
};

__device__ __forceinline__ uint64_t load_bits_w(const void *p, int w, uint64_t r) {
    switch (w) {
    case 8: return gld<uint64_t>(p, r);
    case 4: return gld<uint32_t>(p, r);
    case 2: return gld<uint16_t>(p, r);
    case 1: return gld<uint8_t>(p, r);
    default: return (gld<uint8_t>(p, r >> 3) >> (r & 7)) & 1;  // bit-packed Boolean
    }
}

template <bool RAW64>
__global__ __launch_bounds__(1024) void hs_scan_fast(const HLaunch *__restrict__ L) {
    extern __shared__ __align__(16) uint8_t smem[];
    __shared__ uint32_t s_cursor, s_sel;
    const QueryDesc *__restrict__ q = &L->q;
    const uint32_t CL = L->hot_slots;
    const uint32_t tid = threadIdx.x, bs = blockDim.x;
    const int lane = (int)(tid & 63);
    uint64_t *lkeys = (uint64_t *)smem;
    uint8_t *state = (uint8_t *)(lkeys + CL);

    // ---- descriptor, once per block
    const int ncols = q->n_cols;
    const void *cdat[4];
    const uint8_t *cval[4];
    int ctag[4], cw[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const bool in = k < ncols;
        cdat[k] = in ? q->cols[k].data : nullptr;
        cval[k] = in ? q->cols[k].valid : nullptr;
        ctag[k] = in ? q->cols[k].tag : PMH_I64;
        cw[k] = ctag[k] == PMH_BOOL ? 0 : tag_size(ctag[k]);
    }
    const int nr = L->fs.nr;
    int rcol[2], rflt[2], rtag[2];
    int64_t rlo[2], rhi[2];
#pragma unroll
    for (int r = 0; r < 2; ++r) {
        rcol[r] = L->fs.rcol[r];
        rflt[r] = L->fs.rflt[r];
        rlo[r] = L->fs.rlo[r];
        rhi[r] = L->fs.rhi[r];
        rtag[r] = rcol[r] == 0 ? ctag[0] : rcol[r] == 1 ? ctag[1] : rcol[r] == 2 ? ctag[2] : ctag[3];
    }
    // key: one raw column, or up to four lane fields
    const int nk = RAW64 ? 1 : (q->key_mode == KM_NONE ? 0 : q->n_keys);
    const bool allones = q->allones_possible != 0;
    int kcol[4], knull[4], kdata[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        kcol[i] = i < nk ? q->keys[i].col : 0;
        knull[i] = i < nk ? q->keys[i].null_off : 0;
        kdata[i] = i < nk ? q->keys[i].data_off : 0;
    }
    // aggregated columns and their LDS arrays
    const int nv = q->n_vals;
    auto lds_at = [&](int a) -> uint8_t * { return a >= 0 ? state + L->lds1_off[a] : state; };
    uint32_t *len_p = (uint32_t *)lds_at(q->layout.a_len);
    int vcol[2], vtag[2], vkind[2], vneed[2], vwid[2];
    uint32_t *cnt_p[2], *nan_p[2];
    unsigned long long *limb_p[2], *sum_p[2];
    double *res_p[2];
    void *min_p[2], *max_p[2];
    void *sval[2];
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        const bool in = j < nv;
        const ValSpec &S = q->vals[in ? j : 0];
        vcol[j] = in ? S.col : 0;
        vtag[j] = in ? S.tag : PMH_I64;
        vkind[j] = in ? S.kind : VK_I64;
        vneed[j] = in ? S.need : 0;
        vwid[j] = in ? L->vw[j] : 8;
        const bool sum_f = (vneed[j] & NEED_SUM) && vkind[j] == VK_F;
        const bool sum_i = (vneed[j] & NEED_SUM) && vkind[j] != VK_F;
        const bool mm = (vneed[j] & NEED_MINMAX) != 0;
        cnt_p[j] = (uint32_t *)lds_at(in ? S.a_cnt : -1);
        limb_p[j] = (unsigned long long *)lds_at(sum_f ? S.a_limb : -1);
        res_p[j] = (double *)lds_at(sum_f ? S.a_res : -1);
        sum_p[j] = (unsigned long long *)lds_at(sum_i ? S.a_sum : -1);
        min_p[j] = lds_at(mm ? S.a_min : -1);
        max_p[j] = lds_at(mm ? S.a_max : -1);
        nan_p[j] = (uint32_t *)lds_at(mm && vkind[j] == VK_F ? S.a_nan : -1);
        sval[j] = in ? L->spill.val[j] : nullptr;
    }
    uint64_t *sklo = L->spill.klo;
    uint16_t *sflags = L->spill.flags;
    const uint64_t region = (uint64_t)blockIdx.x * L->chunk;

    // ---- LDS: hot keys and state identities
    if (tid == 0) { s_cursor = 0; s_sel = 0; }
    for (uint32_t i = tid; i < CL; i += bs) lkeys[i] = L->hot_lo[i];
    {
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
    }
    __syncthreads();

    const uint64_t n = q->n_rows;
    const uint64_t r_begin = region;
    const uint64_t r_end = min(n, r_begin + L->chunk);
    const uint64_t span = r_end > r_begin ? r_end - r_begin : 0;
    // every lane runs the same number of steps: the ballots below always see
    // whole waves
    const uint64_t steps = (span + 4ull * bs - 1) / (4ull * bs);
    uint32_t sel_count = 0;

    for (uint64_t it = 0; it < steps; ++it) {
        const uint64_t r0 = r_begin + it * 4ull * bs + tid;
        bool live[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) live[u] = r0 + (uint64_t)u * bs < r_end;
        // ---- loads, column by column so each width switch is paid once
        uint64_t v[4][4];
        uint32_t vm[4] = {0, 0, 0, 0};
#pragma unroll
        for (int k = 0; k < 4; ++k) {
#pragma unroll
            for (int u = 0; u < 4; ++u) v[u][k] = 0;
            if (k >= ncols) continue;
            const int w = cw[k];
#pragma unroll
            for (int u = 0; u < 4; ++u)
                if (live[u]) v[u][k] = load_bits_w(cdat[k], w, r0 + (uint64_t)u * bs);
            if (cval[k]) {
#pragma unroll
                for (int u = 0; u < 4; ++u)
                    if (live[u]) vm[u] |= (uint32_t)dvalid(cval[k], r0 + (uint64_t)u * bs) << k;
            } else {
#pragma unroll
                for (int u = 0; u < 4; ++u) vm[u] |= 1u << k;
            }
        }
        auto pick = [&](int u, int k) -> uint64_t {
            return k == 0 ? v[u][0] : (k == 1 ? v[u][1] : (k == 2 ? v[u][2] : v[u][3]));
        };
        // ---- predicate: interval tests
        bool sel[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            bool ok = live[u];
#pragma unroll
            for (int r = 0; r < 2; ++r) {
                if (r >= nr) continue;
                const uint64_t raw = pick(u, rcol[r]);
                const int64_t key = rflt[r] ? tot_key(raw_to_f64(rtag[r], raw)) : raw_to_i64(rtag[r], raw);
                ok = ok && ((vm[u] >> rcol[r]) & 1) && key >= rlo[r] && key <= rhi[r];
            }
            sel[u] = ok;
            sel_count += ok ? 1u : 0u;
        }
        // ---- keys and hot-table probes
        uint64_t key[4];
        int special[4];
        int32_t slot[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            key[u] = 0;
            special[u] = 0;
            if (RAW64) {
                if (!((vm[u] >> kcol[0]) & 1)) special[u] = 2;
                else {
                    key[u] = pick(u, kcol[0]);
                    special[u] = key[u] == kEmpty ? 1 : 0;
                }
            } else {
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    if (i >= nk) continue;
                    if ((vm[u] >> kcol[i]) & 1) key[u] |= pick(u, kcol[i]) << kdata[i];
                    else key[u] |= 1ull << knull[i];
                }
                special[u] = (allones && key[u] == kEmpty) ? 1 : 0;
            }
            slot[u] = -1;
            if (sel[u] && special[u] == 0 && CL) {
                uint32_t s = (uint32_t)(hash_key(key[u], 0) >> 40) & (CL - 1);
                for (int p = 0; p < kProbe; ++p) {
                    const uint64_t cur = lkeys[s];
                    if (cur == key[u]) { slot[u] = (int32_t)s; break; }
                    if (cur == kEmpty) break;
                    s = (s + 1) & (CL - 1);
                }
            }
        }
        // ---- hot rows: LDS state updates
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if (slot[u] >= 0) atomicAdd(len_p + slot[u], 1u);
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            if (j >= nv) continue;
            const int vc = vcol[j], tg = vtag[j], need = vneed[j];
            if (vkind[j] == VK_F) {
                FSink fs{limb_p[j], res_p[j], CL};
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    if (slot[u] < 0 || !((vm[u] >> vc) & 1)) continue;
                    const uint32_t sl = (uint32_t)slot[u];
                    const double x = raw_to_f64(tg, pick(u, vc));
                    atomicAdd(cnt_p[j] + sl, 1u);
                    if (need & NEED_SUM) add_f64_exact(fs, 0, 0, sl, x);
                    if (need & NEED_MINMAX) {
                        if (x != x) atomicOr(nan_p[j] + sl, 1u);
                        else {
                            const long long kk = (long long)f64_order_key(x);
                            atomicMin((long long *)min_p[j] + sl, kk);
                            atomicMax((long long *)max_p[j] + sl, kk);
                        }
                    }
                }
            } else if (vkind[j] == VK_I64) {
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    if (slot[u] < 0 || !((vm[u] >> vc) & 1)) continue;
                    const uint32_t sl = (uint32_t)slot[u];
                    const int64_t x = raw_to_i64(tg, pick(u, vc));
                    atomicAdd(cnt_p[j] + sl, 1u);
                    if (need & NEED_SUM) atomicAdd(sum_p[j] + sl, (unsigned long long)x);
                    if (need & NEED_MINMAX) {
                        atomicMin((long long *)min_p[j] + sl, (long long)x);
                        atomicMax((long long *)max_p[j] + sl, (long long)x);
                    }
                }
            } else {
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    if (slot[u] < 0 || !((vm[u] >> vc) & 1)) continue;
                    const uint32_t sl = (uint32_t)slot[u];
                    const int32_t x = (int32_t)raw_to_i64(tg, pick(u, vc));
                    atomicAdd(cnt_p[j] + sl, 1u);
                    if (need & NEED_SUM) atomicAdd(sum_p[j] + sl, (unsigned long long)(int64_t)x);
                    if (need & NEED_MINMAX) {
                        atomicMin((int *)min_p[j] + sl, x);
                        atomicMax((int *)max_p[j] + sl, x);
                    }
                }
            }
        }
        // ---- tail rows: one cursor reservation per wave for all four rows
        unsigned long long m[4];
        uint32_t tot = 0;
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            m[u] = __ballot(sel[u] && slot[u] < 0);
            tot += (uint32_t)__popcll(m[u]);
        }
        if (tot) {
            const int leader = __ffsll((long long)__ballot(1)) - 1;
            uint32_t base = 0;
            if (lane == leader) base = atomicAdd(&s_cursor, tot);
            base = __shfl(base, leader);
            const unsigned long long below = (1ull << lane) - 1;
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                if ((m[u] >> lane) & 1) {
                    const uint64_t pos = region + base + (uint32_t)__popcll(m[u] & below);
                    uint16_t fl = special[u] == 1 ? kFlagAllOnes : (special[u] == 2 ? kFlagNull : 0);
#pragma unroll
                    for (int j = 0; j < 2; ++j) {
                        if (j >= nv) continue;
                        fl |= (uint16_t)(((vm[u] >> vcol[j]) & 1) << j);
                        gstore_w(sval[j], vwid[j], pos, pick(u, vcol[j]));
                    }
                    gst<uint64_t>(sklo, pos, key[u]);
                    gst<uint16_t>(sflags, pos, fl);
                }
                base += (uint32_t)__popcll(m[u]);
            }
        }
    }
    atomicAdd(&s_sel, sel_count);
    __syncthreads();
    if (tid == 0) {
        L->tail_n[blockIdx.x] = s_cursor;
        atomicAdd(&L->stats[ST_SELECTED], s_sel);
        atomicAdd(&L->stats[ST_TAIL], s_cursor);
    }
    // hot partials: block slot s -> global slot s (coalesced atomics)
    GSink g{L->hot_arr, CL};
    LSink ls{state, L->lds1_off, CL};
    const int a_len = q->layout.a_len;
    for (uint32_t i = tid; i < CL; i += bs) {
        if (*ls.p<uint32_t>(a_len, i) == 0) continue;
        merge_lds_entry(g, i, ls, i, q->layout, q->layout.init);
    }
}

}  // namespace pmh
