// pmh_runtime.cpp -- device runtime of the backend: one HIP device and one
// stream per process, grow-only named device slabs (the analogue of the
// reference's persistent BuildScratch / StagingPool,
// crates/polars-metal-kernels/src/groupby_build_partitioned/scratch.rs:29-145,
// crates/polars-metal-buffer/src/staging.rs:24-58), column staging and the
// optional per-kernel HIP-event timers (host columns upload through the
// staging ring of pmh_ingest.hip).
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <set>

#include "pmh_internal.h"

namespace pmh {

static thread_local std::string g_last_error;
void set_last_error(const std::string &msg) { g_last_error = msg; }
const char *last_error() { return g_last_error.c_str(); }

static Runtime *g_rt = nullptr;

// Device memory of the allocation-free entry point: one array per device,
// held until the process ends.  Host code never uses this name as an
// address; Runtime::scratch holds the device address hipGetSymbolAddress
// returns for it.
static __device__ __align__(256) uint8_t scratch[kScratchBytes];

// Pinned mailbox bytes and pooled events made at set-up, enough for every
// single-query engine, so an allocation-free call needs no more.
constexpr size_t kSetupPinnedBytes = size_t(1) << 20;
constexpr int kSetupEvents = 8;

Runtime &rt() {
    if (g_rt) return *g_rt;
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n <= 0)
        fail(PMH_E_COMPUTE, "polars_metal: no HIP device available (the MI355X backend has no CPU fallback)");
    auto *r = new Runtime();
    PMH_HIP(hipGetDevice(&r->device));
    PMH_HIP(hipGetDeviceProperties(&r->prop, r->device));
    // Default priority: streams of different priorities are not used.
    PMH_HIP(hipStreamCreateWithFlags(&r->stream, hipStreamNonBlocking));
    // the scratch array's device address on this (the current) device
    void *sp = nullptr;
    const hipError_t se = hipGetSymbolAddress(&sp, HIP_SYMBOL(scratch));
    if (se != hipSuccess || !sp) {
        (void)hipGetLastError();
        fail(PMH_E_COMPUTE, std::string("polars_metal: hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(se));
    }
    r->scratch = (uint8_t *)sp;
    r->scratch_cap = kScratchBytes;
    for (int i = 0; i < kSetupEvents; ++i) {
        hipEvent_t ev;
        PMH_HIP(hipEventCreate(&ev));
        r->event_pool.push_back(ev);
    }
    g_rt = r;
    (void)pinned_host(kSetupPinnedBytes);
    return *r;
}

ScratchScope::ScratchScope() {
    Runtime &r = rt();
    r.scratch_mode = true;
    r.scratch_used = 0;
    r.scratch_slabs.clear();
}

ScratchScope::~ScratchScope() {
    Runtime &r = *g_rt;
    r.scratch_mode = false;
    r.scratch_slabs.clear();
}

PooledEvent::PooledEvent() {
    Runtime &r = rt();
    if (!r.event_pool.empty()) {
        ev = r.event_pool.back();
        r.event_pool.pop_back();
        return;
    }
    if (r.scratch_mode) fail(PMH_E_RUNTIME, "polars_metal: allocation-free call ran out of pooled events");
    PMH_HIP(hipEventCreate(&ev));
}

PooledEvent::~PooledEvent() {
    if (ev) g_rt->event_pool.push_back(ev);
}

void *slab(const std::string &name, size_t bytes) {
    Runtime &r = rt();
    if (bytes == 0) bytes = 256;
    if (r.scratch_mode) {
        auto it = r.scratch_slabs.find(name);
        if (it != r.scratch_slabs.end() && it->second.second >= bytes) return r.scratch + it->second.first;
        const size_t off = (r.scratch_used + 255) & ~size_t(255);
        if (off + bytes > r.scratch_cap)
            fail(PMH_E_VALUE, "polars_metal: allocation-free call needs more than its " +
                                  std::to_string(r.scratch_cap) + "-byte scratch buffer (at " + name + ")");
        r.scratch_used = off + bytes;
        r.scratch_slabs[name] = {off, bytes};
        return r.scratch + off;
    }
    auto it = r.slabs.find(name);
    if (it != r.slabs.end() && it->second.second >= bytes) return it->second.first;
    if (it != r.slabs.end()) {
        PMH_HIP(hipStreamSynchronize(r.stream));
        PMH_HIP(hipFree(it->second.first));
        r.slabs.erase(it);
    }
    size_t sz = bytes + bytes / 4;
    sz = (sz + (2u << 20) - 1) & ~size_t((2u << 20) - 1);
    void *p = nullptr;
    PMH_HIP(hipMalloc(&p, sz));
    r.slabs[name] = {p, sz};
    return p;
}

void *pinned_host(size_t bytes) {
    Runtime &r = rt();
    if (r.pinned_size >= bytes && r.pinned) return r.pinned;
    if (r.scratch_mode) fail(PMH_E_RUNTIME, "polars_metal: allocation-free call needs a larger pinned mailbox");
    if (r.pinned) {
        PMH_HIP(hipStreamSynchronize(r.stream));
        PMH_HIP(hipHostFree(r.pinned));
        r.pinned = nullptr;
    }
    size_t sz = bytes + bytes / 4 + 4096;
    PMH_HIP(hipHostMalloc(&r.pinned, sz, hipHostMallocDefault));
    r.pinned_size = sz;
    return r.pinned;
}

KernelTimer::KernelTimer(const char *n) : name(n) {
    Runtime &r = rt();
    if (!r.profiling || r.scratch_mode) return;
    auto take = [&]() {
        hipEvent_t ev;
        if (!r.event_pool.empty()) { ev = r.event_pool.back(); r.event_pool.pop_back(); }
        else PMH_HIP(hipEventCreate(&ev));
        return ev;
    };
    a = take();
    b = take();
    PMH_HIP(hipEventRecord(a, r.stream));
}

KernelTimer::~KernelTimer() {
    if (!a) return;
    Runtime &r = rt();
    if (hipEventRecord(b, r.stream) == hipSuccess) r.pending.push_back({name, {a, b}});
}

void collect_kernel_times() {
    Runtime &r = rt();
    if (r.pending.empty()) return;
    for (auto &p : r.pending) {
        (void)hipEventSynchronize(p.second.second);
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, p.second.first, p.second.second) == hipSuccess) {
            auto &slot = r.ktimes[p.first];
            slot.first += 1;
            slot.second += ms;
        }
        r.event_pool.push_back(p.second.first);
        r.event_pool.push_back(p.second.second);
    }
    r.pending.clear();
}

pmh_result *new_result(size_t n_cols) {
    auto *res = (pmh_result *)calloc(1, sizeof(pmh_result));
    res->n_columns = (uint32_t)n_cols;
    res->columns = (pmh_out_column *)calloc(n_cols ? n_cols : 1, sizeof(pmh_out_column));
    return res;
}

void set_out_name(pmh_out_column &c, const std::string &name) {
    char *s = (char *)malloc(name.size() + 1);
    memcpy(s, name.c_str(), name.size() + 1);
    c.name = s;
}

void check_column(const pmh_column &c, uint64_t n_rows, const std::string &nm) {
    if (c.tag < 0 || c.tag > PMH_UTF8) fail(PMH_E_VALUE, "polars_metal: unknown dtype tag for column " + nm);
    const int esz = tag_size(c.tag);
    const uint64_t need = esz == 0 ? (n_rows + 7) / 8 : n_rows * (uint64_t)esz;
    if (c.data_len < need || (!c.data && need))
        fail(PMH_E_VALUE, "polars_metal: " + std::string(tag_name(c.tag)) + " value buffer too short for " + nm +
                              ": " + std::to_string(c.data_len) + " < " + std::to_string(need));
    const uint64_t vneed = (n_rows + 7) / 8;
    if (c.valid && c.valid_len < vneed)
        fail(PMH_E_VALUE, "polars_metal: validity buffer too short for " + nm + ": " + std::to_string(c.valid_len) +
                              " < " + std::to_string(vneed));
}

std::vector<Column> stage_columns(const pmh_column *cols, uint32_t n_cols, uint64_t n_rows,
                                  const std::vector<std::string> &needed) {
    rt();  // device set up before any staging
    std::vector<Column> out;
    std::set<std::string> seen;
    int staged = 0;
    for (const auto &nm : needed) {
        if (seen.count(nm)) continue;
        seen.insert(nm);
        const pmh_column *c = nullptr;
        for (uint32_t i = 0; i < n_cols; ++i)
            if (cols[i].name && nm == cols[i].name) { c = &cols[i]; break; }
        if (!c) fail(PMH_E_KEY, "polars_metal: column not found: " + nm);
        check_column(*c, n_rows, nm);
        const int esz = tag_size(c->tag);
        const uint64_t need = esz == 0 ? (n_rows + 7) / 8 : n_rows * (uint64_t)esz;
        const uint64_t vneed = (n_rows + 7) / 8;
        Column col;
        col.name = nm;
        col.tag = c->tag;
        col.n_rows = n_rows;
        if (c->location == PMH_DEVICE) {
            col.d_data = c->data;
            col.d_valid = c->valid;
        } else {
            // Device copies are padded to 16 bytes so vector loads never run
            // past the allocation.
            void *d = slab("in:" + std::to_string(staged), need + 16);
            upload_host(d, c->data, need);
            col.d_data = d;
            col.d_valid = nullptr;
            if (c->valid) {
                void *dv = slab("inv:" + std::to_string(staged), vneed + 16);
                upload_host(dv, c->valid, vneed);
                col.d_valid = (const uint8_t *)dv;
            }
            ++staged;
        }
        out.push_back(col);
    }
    return out;
}

// Bitmap bytes [0, vbytes) of a group count ng: all ones for rows < ng
// (ones), or the copied bits kept; every bit from row ng on cleared.
__global__ void tidy_bitmap(uint8_t *bm, uint64_t ng, uint64_t vbytes, int ones) {
    const uint64_t full = ng >> 3;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < vbytes; i += (uint64_t)gridDim.x * blockDim.x) {
        if (i < full) {
            if (ones) bm[i] = 0xFF;
        } else if (i == full && (ng & 7)) {
            const uint8_t keep = (uint8_t)((1u << (ng & 7)) - 1);
            bm[i] = ones ? keep : (uint8_t)(bm[i] & keep);
        } else {
            bm[i] = 0;
        }
    }
}

void copy_outputs_to_caller(const pmh_result &r, pmh_out_column *outs, uint32_t n_outs) {
    Runtime &R = rt();
    const uint64_t ng = r.n_rows;
    const uint64_t vbytes = bitmap_bytes_padded(ng);
    const uint64_t packed = (ng + 7) / 8;
    for (uint32_t i = 0; i < n_outs; ++i) {
        const pmh_out_column &src = r.columns[i];
        pmh_out_column &dst = outs[i];
        if (src.location != PMH_DEVICE) fail(PMH_E_RUNTIME, "polars_metal: allocation-free call got host outputs");
        const bool bits = src.tag == PMH_BOOL;
        const uint64_t need = bits ? vbytes : ng * (uint64_t)tag_size(src.tag);
        if (need && src.data_len < (bits ? packed : need))
            fail(PMH_E_RUNTIME, "polars_metal: allocation-free call: engine output shorter than its group count");
        if (need) PMH_HIP(hipMemcpyAsync(dst.data, src.data, bits ? packed : need, hipMemcpyDeviceToDevice, R.stream));
        if (bits && need) hipLaunchKernelGGL(tidy_bitmap, dim3(1), dim3(256), 0, R.stream, (uint8_t *)dst.data, ng, vbytes, 0);
        if (dst.valid) {
            if (src.valid && packed)
                PMH_HIP(hipMemcpyAsync(dst.valid, src.valid, packed, hipMemcpyDeviceToDevice, R.stream));
            hipLaunchKernelGGL(tidy_bitmap, dim3((uint32_t)std::min<uint64_t>(1024, vbytes / 256 + 1)), dim3(256), 0,
                               R.stream, dst.valid, ng, vbytes, src.valid ? 0 : 1);
        }
        PMH_HIP(hipGetLastError());
        dst.tag = src.tag;
        dst.location = PMH_DEVICE;
        dst.data_len = need;
        dst.valid_len = dst.valid ? vbytes : 0;
    }
    PMH_HIP(hipStreamSynchronize(R.stream));
}

}  // namespace pmh
