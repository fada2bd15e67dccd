// pmh_filter.hip -- standalone filter path for gfx950: stable stream
// compaction, column-vs-scalar / column-vs-column comparisons and Kleene
// AND / OR on bit-packed (data, valid) pairs.
//
// Replaces crates/polars-metal-kernels/src/{pipeline.rs:87-270,
// filter.rs:134-649, cmp.rs:180-641, logical.rs:98-233} and the MLX cumsum
// (crates/polars-metal-mlx-sys/cxx/mlx_bridge.cc:59-92).
//
// Compaction is three launches with no host round trip in between:
//   cmp_count   one 64-bit ballot per wave gives each block's kept-row count
//   cmp_scan    one block turns the block counts into exclusive offsets
//   cmp_scatter every wave ranks its kept rows with the ballot mask
//               (popcount of the lanes below) and writes them in source order
// Validity and Bool data are written as one byte per output row, then packed
// into bitmaps by pack_bits (one thread per output byte).
#include <cstring>

#include "pmh_device.h"

namespace pmh {

constexpr int kCompactBlock = 256;
constexpr int kCompactRows = kCompactBlock * 16;  // rows per block

__device__ __forceinline__ bool keep_bit(const uint8_t *d, const uint8_t *v, uint64_t r) {
    return ((d[r >> 3] & v[r >> 3]) >> (r & 7)) & 1;
}

__global__ __launch_bounds__(kCompactBlock) void cmp_count(const uint8_t *__restrict__ d, const uint8_t *__restrict__ v,
                                                          uint64_t n, uint32_t *__restrict__ counts) {
    __shared__ uint32_t s_cnt;
    if (threadIdx.x == 0) s_cnt = 0;
    __syncthreads();
    const uint64_t base = (uint64_t)blockIdx.x * kCompactRows;
    uint32_t c = 0;
    for (int k = 0; k < kCompactRows / kCompactBlock; ++k) {
        uint64_t r = base + (uint64_t)k * kCompactBlock + threadIdx.x;
        bool kb = r < n && keep_bit(d, v, r);
        unsigned long long m = __ballot(kb);
        if ((threadIdx.x & 63) == 0) c += (uint32_t)__popcll(m);
    }
    if ((threadIdx.x & 63) == 0) atomicAdd(&s_cnt, c);
    __syncthreads();
    if (threadIdx.x == 0) counts[blockIdx.x] = s_cnt;
}

// Exclusive scan of nb block counts by one block; offsets[nb] = total.
__global__ __launch_bounds__(1024) void cmp_scan(const uint32_t *__restrict__ counts, uint64_t nb,
                                                uint64_t *__restrict__ offsets) {
    __shared__ uint64_t s[1024];
    __shared__ uint64_t s_carry;
    if (threadIdx.x == 0) s_carry = 0;
    __syncthreads();
    for (uint64_t b0 = 0; b0 < nb; b0 += 1024) {
        uint64_t i = b0 + threadIdx.x;
        uint64_t x = i < nb ? counts[i] : 0;
        s[threadIdx.x] = x;
        __syncthreads();
        for (int off = 1; off < 1024; off <<= 1) {
            uint64_t y = threadIdx.x >= (unsigned)off ? s[threadIdx.x - off] : 0;
            __syncthreads();
            s[threadIdx.x] += y;
            __syncthreads();
        }
        uint64_t incl = s[threadIdx.x] + s_carry;
        if (i < nb) offsets[i] = incl - x;
        __syncthreads();
        if (threadIdx.x == 1023) s_carry = incl;
        __syncthreads();
    }
    if (threadIdx.x == 0) offsets[nb] = s_carry;
}

struct CompactCol {
    const void *data;     // Bool: bit-packed
    const uint8_t *valid; // nullptr: all valid
    void *out;            // element width (Bool: one byte per row)
    uint8_t *out_valid;   // one byte per row
    int32_t esz;          // 0 for Bool
    int32_t pad;
};

__global__ __launch_bounds__(kCompactBlock) void cmp_scatter(const uint8_t *__restrict__ d, const uint8_t *__restrict__ v,
                                                            uint64_t n, const uint64_t *__restrict__ offsets,
                                                            const CompactCol *__restrict__ cols, int ncols) {
    __shared__ uint32_t s_wave[kCompactBlock / 64];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const uint64_t base = (uint64_t)blockIdx.x * kCompactRows;
    uint64_t out0 = offsets[blockIdx.x];
    for (int k = 0; k < kCompactRows / kCompactBlock; ++k) {
        uint64_t r = base + (uint64_t)k * kCompactBlock + threadIdx.x;
        bool kb = r < n && keep_bit(d, v, r);
        unsigned long long m = __ballot(kb);
        if (lane == 0) s_wave[wave] = (uint32_t)__popcll(m);
        __syncthreads();
        uint32_t before = 0, total = 0;
        for (int w = 0; w < kCompactBlock / 64; ++w) {
            if (w < wave) before += s_wave[w];
            total += s_wave[w];
        }
        if (kb) {
            uint64_t o = out0 + before + (uint32_t)__popcll(m & ((1ull << lane) - 1));
            for (int c = 0; c < ncols; ++c) {
                const CompactCol &cc = cols[c];
                switch (cc.esz) {
                case 8: ((uint64_t *)cc.out)[o] = ((const uint64_t *)cc.data)[r]; break;
                case 4: ((uint32_t *)cc.out)[o] = ((const uint32_t *)cc.data)[r]; break;
                case 2: ((uint16_t *)cc.out)[o] = ((const uint16_t *)cc.data)[r]; break;
                case 1: ((uint8_t *)cc.out)[o] = ((const uint8_t *)cc.data)[r]; break;
                default: ((uint8_t *)cc.out)[o] = (((const uint8_t *)cc.data)[r >> 3] >> (r & 7)) & 1; break;
                }
                cc.out_valid[o] = cc.valid ? ((cc.valid[r >> 3] >> (r & 7)) & 1) : 1;
            }
        }
        out0 += total;
        __syncthreads();
    }
}

// bytes (0/1 per row) -> LSB-first bitmap of nbytes_out bytes (zero padded)
__global__ void pack_bits(const uint8_t *__restrict__ in, uint64_t n, uint8_t *__restrict__ out, uint64_t nbytes_out) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nbytes_out) return;
    uint8_t b = 0;
    for (int k = 0; k < 8; ++k) {
        uint64_t r = i * 8 + k;
        if (r < n && in[r]) b |= (uint8_t)(1u << k);
    }
    out[i] = b;
}

// ---- comparisons: one thread per 32 rows -> one word of data and valid
__global__ void cmp_words(int is_f64, int col_col, const void *__restrict__ lhs, const uint8_t *__restrict__ lv,
                          const void *__restrict__ rhs, const uint8_t *__restrict__ rv, int64_t rhs_i, double rhs_f,
                          int op, uint64_t n, uint32_t *__restrict__ od, uint32_t *__restrict__ ov, uint64_t nwords) {
    uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= nwords) return;
    uint32_t D = 0, V = 0;
    for (int k = 0; k < 32; ++k) {
        uint64_t r = w * 32 + k;
        if (r >= n) break;
        bool valid = dvalid(lv, r) && (!col_col || dvalid(rv, r));
        if (!valid) continue;
        int c;
        if (is_f64) {
            double a = ((const double *)lhs)[r];
            double b = col_col ? ((const double *)rhs)[r] : rhs_f;
            c = tot_cmp(a, b);
        } else {
            int64_t a = ((const int64_t *)lhs)[r];
            int64_t b = col_col ? ((const int64_t *)rhs)[r] : rhs_i;
            c = (a < b) ? -1 : ((a > b) ? 1 : 0);
        }
        V |= 1u << k;
        if (apply_cmp(op, c)) D |= 1u << k;
    }
    od[w] = D;
    ov[w] = V;
}

// ---- Kleene AND / OR on whole words (shaders/logical_bool.metal:1-31)
__global__ void bool_words(int is_or, const uint8_t *__restrict__ ld, const uint8_t *__restrict__ lv,
                           const uint8_t *__restrict__ rd, const uint8_t *__restrict__ rv, uint64_t n,
                           uint32_t *__restrict__ od, uint32_t *__restrict__ ov, uint64_t nwords) {
    uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= nwords) return;
    uint32_t a = 0, av = 0, b = 0, bv = 0;
    for (int k = 0; k < 4; ++k) {
        uint64_t byte = w * 4 + k;
        if (byte * 8 >= n) break;
        a |= (uint32_t)ld[byte] << (8 * k);
        av |= (uint32_t)(lv ? lv[byte] : 0xFF) << (8 * k);
        b |= (uint32_t)rd[byte] << (8 * k);
        bv |= (uint32_t)(rv ? rv[byte] : 0xFF) << (8 * k);
    }
    uint64_t rem = n - w * 32;
    uint32_t live = rem >= 32 ? 0xFFFFFFFFu : ((1u << rem) - 1);
    uint32_t D, V;
    if (!is_or) {
        uint32_t lf = av & ~a, rf = bv & ~b;  // known false dominates
        uint32_t both = av & bv;
        V = (lf | rf | both) & live;
        D = (a & b & both) & ~(lf | rf) & live;
    } else {
        uint32_t lt = av & a, rt = bv & b;    // known true dominates
        uint32_t both = av & bv;
        V = (lt | rt | both) & live;
        D = (lt | rt) & live;
    }
    od[w] = D;
    ov[w] = V;
}

// ---------------------------------------------------------------------------
// host
// ---------------------------------------------------------------------------
pmh_result *run_filter_compact(const uint8_t *pd, const uint8_t *pv, uint64_t n_rows, const pmh_column *cols,
                               uint32_t n_cols) {
    if (n_rows >= (1ull << 31)) fail(PMH_E_VALUE, "polars_metal: row count must be below 2^31");
    if (n_rows && (!pd || !pv)) fail(PMH_E_RUNTIME, "polars_metal: predicate buffers missing");
    for (uint32_t i = 0; i < n_cols; ++i) {
        int t = cols[i].tag;
        if (t < 0 || t > PMH_U64 || t == PMH_UTF8)
            fail(PMH_E_RUNTIME, std::string("polars_metal: execute_filter_compact: unsupported dtype ") + tag_name(t));
    }
    Runtime &R = rt();
    std::vector<std::string> names;
    for (uint32_t i = 0; i < n_cols; ++i) {
        if (!cols[i].name) fail(PMH_E_VALUE, "polars_metal: column without a name");
        names.push_back(cols[i].name);
    }
    std::vector<Column> dc = stage_columns(cols, n_cols, n_rows, names);
    const uint64_t vb = (n_rows + 7) / 8;
    uint8_t *dpd = (uint8_t *)slab("fc:pd", vb + 16);
    uint8_t *dpv = (uint8_t *)slab("fc:pv", vb + 16);
    if (vb) {
        PMH_HIP(hipMemcpyAsync(dpd, pd, vb, hipMemcpyHostToDevice, R.stream));
        PMH_HIP(hipMemcpyAsync(dpv, pv, vb, hipMemcpyHostToDevice, R.stream));
    }
    const uint64_t nb = std::max<uint64_t>(1, ceil_div(n_rows, kCompactRows));
    uint32_t *counts = (uint32_t *)slab("fc:counts", nb * 4);
    uint64_t *offs = (uint64_t *)slab("fc:offs", (nb + 1) * 8);
    {
        KernelTimer kt("cmp_count");
        hipLaunchKernelGGL(cmp_count, dim3(nb), dim3(kCompactBlock), 0, R.stream, dpd, dpv, n_rows, counts);
        PMH_HIP(hipGetLastError());
    }
    {
        KernelTimer kt("cmp_scan");
        hipLaunchKernelGGL(cmp_scan, dim3(1), dim3(1024), 0, R.stream, counts, nb, offs);
        PMH_HIP(hipGetLastError());
    }
    uint64_t n_out = 0;
    PMH_HIP(hipMemcpyAsync(&n_out, offs + nb, 8, hipMemcpyDeviceToHost, R.stream));
    PMH_HIP(hipStreamSynchronize(R.stream));
    // name-matched columns are de-duplicated by stage_columns; map back
    std::vector<CompactCol> cc(n_cols);
    std::vector<uint8_t *> obytes(n_cols), ovbytes(n_cols);
    for (uint32_t i = 0; i < n_cols; ++i) {
        const Column *src = nullptr;
        for (auto &c : dc)
            if (c.name == names[i]) src = &c;
        int esz = tag_size(cols[i].tag);
        cc[i].data = src->d_data;
        cc[i].valid = src->d_valid;
        cc[i].esz = esz;
        cc[i].out = slab("fc:o" + std::to_string(i), (n_out + 1) * (esz ? esz : 1) + 16);
        cc[i].out_valid = (uint8_t *)slab("fc:ov" + std::to_string(i), n_out + 16);
    }
    CompactCol *dcc = (CompactCol *)slab("fc:cols", sizeof(CompactCol) * (n_cols ? n_cols : 1));
    if (n_cols) PMH_HIP(hipMemcpyAsync(dcc, cc.data(), sizeof(CompactCol) * n_cols, hipMemcpyHostToDevice, R.stream));
    if (n_rows && n_cols) {
        KernelTimer kt("cmp_scatter");
        hipLaunchKernelGGL(cmp_scatter, dim3(nb), dim3(kCompactBlock), 0, R.stream, dpd, dpv, n_rows, offs, dcc,
                           (int)n_cols);
        PMH_HIP(hipGetLastError());
    }
    pmh_result *res = new_result(n_cols);
    res->n_rows = n_out;
    res->stats.n_selected = n_out;
    const uint64_t vbytes = bitmap_bytes_padded(n_out);
    for (uint32_t i = 0; i < n_cols; ++i) {
        pmh_out_column &c = res->columns[i];
        set_out_name(c, names[i]);
        c.tag = cols[i].tag;
        c.location = PMH_HOST;
        int esz = cc[i].esz;
        c.valid_len = n_out ? vbytes : 0;
        c.valid = (uint8_t *)malloc(vbytes);
        uint8_t *pk = (uint8_t *)slab("fc:pk" + std::to_string(i), vbytes + 16);
        if (n_out) {
            hipLaunchKernelGGL(pack_bits, dim3(ceil_div(vbytes, 256)), dim3(256), 0, R.stream, cc[i].out_valid, n_out, pk,
                               vbytes);
            PMH_HIP(hipMemcpyAsync(c.valid, pk, vbytes, hipMemcpyDeviceToHost, R.stream));
        }
        if (esz == 0) {
            c.data_len = n_out ? vbytes : 0;
            c.data = malloc(vbytes);
            uint8_t *pk2 = (uint8_t *)slab("fc:pkd" + std::to_string(i), vbytes + 16);
            if (n_out) {
                hipLaunchKernelGGL(pack_bits, dim3(ceil_div(vbytes, 256)), dim3(256), 0, R.stream, (uint8_t *)cc[i].out,
                                   n_out, pk2, vbytes);
                PMH_HIP(hipMemcpyAsync(c.data, pk2, vbytes, hipMemcpyDeviceToHost, R.stream));
            }
        } else {
            c.data_len = n_out * esz;
            c.data = malloc(c.data_len ? c.data_len : 1);
            if (c.data_len) PMH_HIP(hipMemcpyAsync(c.data, cc[i].out, c.data_len, hipMemcpyDeviceToHost, R.stream));
        }
    }
    PMH_HIP(hipStreamSynchronize(R.stream));
    collect_kernel_times();
    return res;
}

void run_cmp(int is_f64, int col_col, const void *lhs, const uint8_t *lv, const void *rhs, const uint8_t *rv,
             int64_t rhs_i, double rhs_f, int op, uint64_t n_rows, uint8_t *out_d, uint8_t *out_v) {
    if (op < 0 || op > PMH_GE) fail(PMH_E_VALUE, "polars_metal: unknown compare op");
    if (n_rows >= (1ull << 31)) fail(PMH_E_VALUE, "polars_metal: row count must be below 2^31");
    Runtime &R = rt();
    const uint64_t vb = (n_rows + 7) / 8;
    const uint64_t out_bytes = bitmap_bytes_padded(n_rows);
    const uint64_t nwords = out_bytes / 4;
    void *dl = slab("cmp:l", n_rows * 8 + 16);
    void *dr = slab("cmp:r", n_rows * 8 + 16);
    uint8_t *dlv = lv ? (uint8_t *)slab("cmp:lv", vb + 16) : nullptr;
    uint8_t *drv = rv ? (uint8_t *)slab("cmp:rv", vb + 16) : nullptr;
    if (n_rows) {
        PMH_HIP(hipMemcpyAsync(dl, lhs, n_rows * 8, hipMemcpyHostToDevice, R.stream));
        if (col_col) PMH_HIP(hipMemcpyAsync(dr, rhs, n_rows * 8, hipMemcpyHostToDevice, R.stream));
        if (lv) PMH_HIP(hipMemcpyAsync(dlv, lv, vb, hipMemcpyHostToDevice, R.stream));
        if (rv && col_col) PMH_HIP(hipMemcpyAsync(drv, rv, vb, hipMemcpyHostToDevice, R.stream));
    }
    uint32_t *od = (uint32_t *)slab("cmp:od", out_bytes);
    uint32_t *ov = (uint32_t *)slab("cmp:ov", out_bytes);
    {
        KernelTimer kt("cmp_words");
        hipLaunchKernelGGL(cmp_words, dim3(ceil_div(nwords, 256)), dim3(256), 0, R.stream, is_f64, col_col, dl, dlv, dr,
                           col_col ? drv : nullptr, rhs_i, rhs_f, op, n_rows, od, ov, nwords);
        PMH_HIP(hipGetLastError());
    }
    PMH_HIP(hipMemcpyAsync(out_d, od, out_bytes, hipMemcpyDeviceToHost, R.stream));
    PMH_HIP(hipMemcpyAsync(out_v, ov, out_bytes, hipMemcpyDeviceToHost, R.stream));
    PMH_HIP(hipStreamSynchronize(R.stream));
    collect_kernel_times();
}

void run_bool(int is_or, const uint8_t *ld, const uint8_t *lv, const uint8_t *rd, const uint8_t *rv, uint64_t n_rows,
              uint8_t *out_d, uint8_t *out_v) {
    if (n_rows >= (1ull << 31)) fail(PMH_E_VALUE, "polars_metal: row count must be below 2^31");
    Runtime &R = rt();
    const uint64_t out_bytes = bitmap_bytes_padded(n_rows);
    const uint64_t nwords = out_bytes / 4;
    const uint64_t vb = (n_rows + 7) / 8;
    uint8_t *b[4];
    const uint8_t *src[4] = {ld, lv, rd, rv};
    for (int i = 0; i < 4; ++i) {
        b[i] = (uint8_t *)slab("bool:" + std::to_string(i), out_bytes + 16);
        PMH_HIP(hipMemsetAsync(b[i], 0, out_bytes, R.stream));
        if (src[i] && vb) PMH_HIP(hipMemcpyAsync(b[i], src[i], vb, hipMemcpyHostToDevice, R.stream));
    }
    uint32_t *od = (uint32_t *)slab("bool:od", out_bytes);
    uint32_t *ov = (uint32_t *)slab("bool:ov", out_bytes);
    {
        KernelTimer kt("bool_words");
        hipLaunchKernelGGL(bool_words, dim3(ceil_div(nwords, 256)), dim3(256), 0, R.stream, is_or, b[0],
                           lv ? b[1] : nullptr, b[2], rv ? b[3] : nullptr, n_rows, od, ov, nwords);
        PMH_HIP(hipGetLastError());
    }
    PMH_HIP(hipMemcpyAsync(out_d, od, out_bytes, hipMemcpyDeviceToHost, R.stream));
    PMH_HIP(hipMemcpyAsync(out_v, ov, out_bytes, hipMemcpyDeviceToHost, R.stream));
    PMH_HIP(hipStreamSynchronize(R.stream));
    collect_kernel_times();
}

}  // namespace pmh
