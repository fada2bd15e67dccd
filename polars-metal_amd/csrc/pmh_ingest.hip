// pmh_ingest.hip -- host columns onto the device: the pinned staging ring
// every host-to-device copy of the library goes through, and the device
// column cache that keeps uploaded (and chunk-assembled) columns resident
// across queries.
//
// The reference ingests through a reusable page-aligned Shared staging
// buffer (crates/polars-metal-buffer/src/staging.rs:24-58) and wraps
// page-aligned Arrow buffers without a copy (bridge.rs:56-64); on Apple's
// unified memory that second path reads host bytes in place.  A discrete
// MI355X reads host memory only over PCIe, so here:
//
//   * staging ring: kRingSlots pinned host slots of kRingSlot bytes, made once.
//     A copy is cut into slot-sized pieces; a pool of host threads copies a
//     piece into the next free slot (waiting for that slot's previous DMA)
//     and the DMA engine moves it to the device on the library's stream, so
//     the host copy of piece i + 1 overlaps the DMA of piece i.  Pageable
//     memory is never registered (hipHostRegister is not used).
//   * column cache: the zero-copy analogue.  A column's device copy is found
//     again by its tag, its chunks' host addresses, row counts and bit
//     offsets, and a caller stamp; a second query on the same buffers copies
//     nothing.  The library cannot see host writes, so a caller whose bytes
//     at an address change (or whose buffer is freed and the address reused)
//     passes a new stamp.  Entries are held while a query uses them and
//     evicted least recently used beyond a byte capacity.
//   * chunked columns (the reference concatenates Arrow chunks on the host,
//     python/polars_metal/_udf.py:1441-1456) are assembled on the device:
//     fixed-width values DMA straight into place; validity bitmaps (and
//     Bool values) whose chunks start at any bit go through assemble_bits.
#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <cstring>
#include <mutex>
#include <sched.h>
#include <thread>

#include "pmh_internal.h"

namespace pmh {

namespace {

// ---------------------------------------------------------------------------
// host copy pool: persistent threads that split one memcpy
// ---------------------------------------------------------------------------
class CopyPool {
public:
    explicit CopyPool(int n_threads) {
        for (int i = 0; i + 1 < n_threads; ++i) workers_.emplace_back([this] { work(); });
    }
    // dst[0, n) = src[0, n), split across the pool and the calling thread.
    void copy(void *dst, const void *src, size_t n) {
        if (workers_.empty() || n < (size_t(1) << 20)) {
            memcpy(dst, src, n);
            return;
        }
        {
            std::lock_guard<std::mutex> lk(m_);
            dst_ = (uint8_t *)dst;
            src_ = (const uint8_t *)src;
            n_ = n;
            const size_t parts = (workers_.size() + 1) * 2;
            piece_ = std::max<size_t>(size_t(256) << 10, ((n + parts - 1) / parts + 4095) & ~size_t(4095));
            next_.store(0);
            active_ = (int)workers_.size();
            ++gen_;
        }
        cv_.notify_all();
        pieces();
        std::unique_lock<std::mutex> lk(m_);
        done_.wait(lk, [&] { return active_ == 0; });
    }

private:
    void pieces() {
        for (;;) {
            const size_t off = next_.fetch_add(1) * piece_;
            if (off >= n_) return;
            memcpy(dst_ + off, src_ + off, std::min(piece_, n_ - off));
        }
    }
    void work() {
        uint64_t seen = 0;
        for (;;) {
            {
                std::unique_lock<std::mutex> lk(m_);
                cv_.wait(lk, [&] { return gen_ != seen; });
                seen = gen_;
            }
            pieces();
            std::lock_guard<std::mutex> lk(m_);
            if (--active_ == 0) done_.notify_one();
        }
    }
    std::vector<std::thread> workers_;
    std::mutex m_;
    std::condition_variable cv_, done_;
    uint8_t *dst_ = nullptr;
    const uint8_t *src_ = nullptr;
    size_t n_ = 0, piece_ = 0;
    std::atomic<size_t> next_{0};
    int active_ = 0;
    uint64_t gen_ = 0;
};

int copy_threads() {
    if (const char *e = getenv("PMH_COPY_THREADS")) {
        const int v = atoi(e);
        if (v >= 1) return std::min(v, 64);
    }
    cpu_set_t s;
    int n = 8;
    if (sched_getaffinity(0, sizeof(s), &s) == 0) n = CPU_COUNT(&s);
    return std::max(1, std::min(n, 8));
}

// The pool lives until the process ends (its threads are never joined).
CopyPool &pool() {
    static CopyPool *p = new CopyPool(copy_threads());
    return *p;
}

// ---------------------------------------------------------------------------
// staging ring
// ---------------------------------------------------------------------------
constexpr size_t kRingSlot = size_t(32) << 20;
constexpr int kRingSlots = 4;

struct Ring {
    uint8_t *slot[kRingSlots] = {};
    hipEvent_t done[kRingSlots] = {};
    bool pending[kRingSlots] = {};
    int next = 0;
};
Ring *g_ring = nullptr;
uint64_t g_h2d_bytes = 0;

Ring &ring() {
    if (g_ring) return *g_ring;
    rt();
    auto *r = new Ring();
    for (int i = 0; i < kRingSlots; ++i) {
        PMH_HIP(hipHostMalloc((void **)&r->slot[i], kRingSlot, hipHostMallocDefault));
        PMH_HIP(hipEventCreateWithFlags(&r->done[i], hipEventDisableTiming));
    }
    g_ring = r;
    return *r;
}

}  // namespace

void upload_host(void *dst, const void *src, size_t bytes) {
    if (!bytes) return;
    Runtime &R = rt();
    if (R.scratch_mode) fail(PMH_E_RUNTIME, "polars_metal: allocation-free call got host columns");
    Ring &g = ring();
    for (size_t off = 0; off < bytes; off += kRingSlot) {
        const size_t n = std::min(kRingSlot, bytes - off);
        const int i = g.next;
        g.next = (i + 1) % kRingSlots;
        if (g.pending[i]) PMH_HIP(hipEventSynchronize(g.done[i]));
        pool().copy(g.slot[i], (const uint8_t *)src + off, n);
        PMH_HIP(hipMemcpyAsync((uint8_t *)dst + off, g.slot[i], n, hipMemcpyHostToDevice, R.stream));
        PMH_HIP(hipEventRecord(g.done[i], R.stream));
        g.pending[i] = true;
    }
    g_h2d_bytes += bytes;
}

uint64_t h2d_bytes() { return g_h2d_bytes; }

void download_host(void *dst, const void *src, size_t bytes) {
    if (!bytes) return;
    Runtime &R = rt();
    if (bytes < (size_t(1) << 20)) {  // small: one pageable copy costs less than the ring's round trips
        PMH_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, R.stream));
        return;
    }
    Ring &g = ring();
    const size_t pieces = ceil_div(bytes, kRingSlot);
    // piece k uses slot k % kRingSlots: up to kRingSlots DMAs in flight while
    // the host copies the oldest one out
    auto issue = [&](size_t k) {
        const int i = (int)(k % kRingSlots);
        if (g.pending[i]) PMH_HIP(hipEventSynchronize(g.done[i]));
        const size_t off = k * kRingSlot, n = std::min(kRingSlot, bytes - off);
        PMH_HIP(hipMemcpyAsync(g.slot[i], (const uint8_t *)src + off, n, hipMemcpyDeviceToHost, R.stream));
        PMH_HIP(hipEventRecord(g.done[i], R.stream));
        g.pending[i] = true;
    };
    for (size_t k = 0; k < std::min<size_t>(pieces, kRingSlots); ++k) issue(k);
    for (size_t k = 0; k < pieces; ++k) {
        const int i = (int)(k % kRingSlots);
        PMH_HIP(hipEventSynchronize(g.done[i]));
        g.pending[i] = false;
        const size_t off = k * kRingSlot, n = std::min(kRingSlot, bytes - off);
        pool().copy((uint8_t *)dst + off, g.slot[i], n);
        if (k + kRingSlots < pieces) issue(k + kRingSlots);
    }
}

// ---------------------------------------------------------------------------
// bit assembly
// ---------------------------------------------------------------------------
struct BitPart {
    uint64_t row0, rows;  // output rows [row0, row0 + rows)
    uint64_t src;         // byte offset of the part's first bitmap byte in the staged bits
    uint32_t bit0;        // bit of that byte holding the part's first row
    uint32_t present;     // 0: no bitmap, every row set
};

// dst byte b holds output rows 8b .. 8b + 7; bits at rows >= n_rows and the
// padding bytes up to dst_bytes are zero.  Parts are sorted by row0 and
// cover [0, n_rows).
__global__ __launch_bounds__(256) void assemble_bits(uint8_t *__restrict__ dst, uint64_t n_rows, uint64_t dst_bytes,
                                                     const BitPart *__restrict__ parts, uint32_t n_parts,
                                                     const uint8_t *__restrict__ bits) {
    for (uint64_t b = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; b < dst_bytes;
         b += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t out = 0;
        const uint64_t r0 = b * 8;
        if (r0 < n_rows) {
            // the part holding row r0: the last with row0 <= r0
            uint32_t lo = 0, hi = n_parts - 1;
            while (lo < hi) {
                const uint32_t mid = (lo + hi + 1) / 2;
                if (parts[mid].row0 <= r0) lo = mid;
                else hi = mid - 1;
            }
            uint32_t p = lo;
            for (int j = 0; j < 8; ++j) {
                const uint64_t r = r0 + j;
                if (r >= n_rows) break;
                while (r >= parts[p].row0 + parts[p].rows) ++p;
                const BitPart &q = parts[p];
                uint32_t v = 1;
                if (q.present) {
                    const uint64_t k = q.bit0 + (r - q.row0);
                    v = (bits[q.src + (k >> 3)] >> (k & 7)) & 1u;
                }
                out |= v << j;
            }
        }
        dst[b] = (uint8_t)out;
    }
}

namespace {

// Stages each chunk's bitmap bytes and assembles them into dst (padded to
// dst_bytes).  which: 0 validity, 1 Bool values.
void assemble(uint8_t *dst, uint64_t n_rows, uint64_t dst_bytes, const pmh_chunk *chunks, uint32_t n_chunks,
              int which) {
    Runtime &R = rt();
    std::vector<BitPart> parts;
    uint64_t row0 = 0, staged = 0;
    for (uint32_t i = 0; i < n_chunks; ++i) {
        const pmh_chunk &c = chunks[i];
        if (!c.n_rows) continue;
        const uint8_t *src = which ? (const uint8_t *)c.data : c.valid;
        const uint32_t bit0 = which ? c.data_bit_offset : c.valid_bit_offset;
        BitPart p{row0, c.n_rows, staged, bit0, src ? 1u : 0u};
        if (src) staged += (bit0 + c.n_rows + 7) / 8;
        parts.push_back(p);
        row0 += c.n_rows;
    }
    if (parts.empty()) {
        PMH_HIP(hipMemsetAsync(dst, 0, dst_bytes, R.stream));
        return;
    }
    uint8_t *bits = (uint8_t *)slab("cc:bits", staged + 16);
    for (size_t i = 0, k = 0; i < n_chunks; ++i) {
        const pmh_chunk &c = chunks[i];
        if (!c.n_rows) continue;
        const BitPart &p = parts[k++];
        if (p.present) upload_host(bits + p.src, which ? c.data : (const void *)c.valid, (p.bit0 + p.rows + 7) / 8);
    }
    auto *dparts = (BitPart *)slab("cc:parts", parts.size() * sizeof(BitPart));
    upload_host(dparts, parts.data(), parts.size() * sizeof(BitPart));
    const uint32_t grid = (uint32_t)std::min<uint64_t>(ceil_div(dst_bytes, 256), 4096);
    hipLaunchKernelGGL(assemble_bits, dim3(grid), dim3(256), 0, R.stream, dst, n_rows, dst_bytes, dparts,
                       (uint32_t)parts.size(), bits);
    PMH_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// column cache
// ---------------------------------------------------------------------------
struct CcKey {
    int32_t tag;
    uint64_t stamp;
    std::vector<uint64_t> parts;  // per chunk: rows, data, data bit offset, valid, valid bit offset
    bool operator<(const CcKey &o) const {
        if (tag != o.tag) return tag < o.tag;
        if (stamp != o.stamp) return stamp < o.stamp;
        return parts < o.parts;
    }
};
struct CcEntry {
    uint64_t id;
    void *data;
    uint8_t *valid;
    uint64_t n_rows, data_len, valid_len;
    size_t bytes;
    int holds;
    uint64_t tick;
};
struct ColCache {
    std::map<CcKey, CcEntry> map;
    std::map<uint64_t, CcKey> by_id;
    uint64_t next_id = 1, tick = 0, capacity = 0, resident = 0, hits = 0, misses = 0, uploaded = 0;
};
ColCache *g_cc = nullptr;

ColCache &cc() {
    if (g_cc) return *g_cc;
    Runtime &R = rt();
    g_cc = new ColCache();
    // a quarter of device memory by default; pmh_colcache_set_capacity changes it
    g_cc->capacity = R.prop.totalGlobalMem / 4;
    return *g_cc;
}

void drop_entry(ColCache &c, std::map<CcKey, CcEntry>::iterator it) {
    Runtime &R = rt();
    PMH_HIP(hipStreamSynchronize(R.stream));
    if (it->second.data) PMH_HIP(hipFree(it->second.data));
    if (it->second.valid) PMH_HIP(hipFree(it->second.valid));
    c.resident -= it->second.bytes;
    c.by_id.erase(it->second.id);
    c.map.erase(it);
}

// Evicts unheld entries, least recently used first, until `need` more bytes
// fit the capacity (or nothing unheld is left).
void make_room(ColCache &c, size_t need) {
    while (c.resident + need > c.capacity) {
        auto victim = c.map.end();
        for (auto it = c.map.begin(); it != c.map.end(); ++it)
            if (it->second.holds == 0 && (victim == c.map.end() || it->second.tick < victim->second.tick)) victim = it;
        if (victim == c.map.end()) return;
        drop_entry(c, victim);
    }
}

void fill_out(const CcEntry &e, int hit, pmh_resident_column *out) {
    out->entry = e.id;
    out->n_rows = e.n_rows;
    out->data = e.data;
    out->data_len = e.data_len;
    out->valid = e.valid;
    out->valid_len = e.valid_len;
    out->hit = hit;
    out->pad = 0;
}

}  // namespace

void column_resident(int32_t tag, const pmh_chunk *chunks, uint32_t n_chunks, uint64_t stamp,
                     pmh_resident_column *out) {
    const int esz = tag_size(tag);
    if (esz < 0) fail(PMH_E_VALUE, "polars_metal: unknown column tag " + std::to_string(tag));
    if (n_chunks && !chunks) fail(PMH_E_VALUE, "polars_metal: chunks is NULL");
    ColCache &c = cc();
    CcKey key{tag, stamp, {}};
    uint64_t n_rows = 0;
    bool any_valid = false;
    for (uint32_t i = 0; i < n_chunks; ++i) {
        const pmh_chunk &k = chunks[i];
        if (k.n_rows && !k.data) fail(PMH_E_VALUE, "polars_metal: chunk " + std::to_string(i) + " has no values");
        if (k.data_bit_offset > 7 || k.valid_bit_offset > 7)
            fail(PMH_E_VALUE, "polars_metal: chunk bit offsets are 0..7 (point data/valid at the first row's byte)");
        if (k.data_bit_offset && esz != 0)
            fail(PMH_E_VALUE, "polars_metal: a data bit offset applies to Bool columns only");
        key.parts.insert(key.parts.end(), {k.n_rows, (uint64_t)(uintptr_t)k.data, k.data_bit_offset,
                                           (uint64_t)(uintptr_t)k.valid, k.valid_bit_offset});
        n_rows += k.n_rows;
        any_valid |= k.n_rows && k.valid;
    }
    if (n_rows >= (uint64_t(1) << 40)) fail(PMH_E_VALUE, "polars_metal: column too long");
    ++c.tick;
    auto it = c.map.find(key);
    if (it != c.map.end()) {
        ++it->second.holds;
        it->second.tick = c.tick;
        ++c.hits;
        fill_out(it->second, 1, out);
        return;
    }
    ++c.misses;
    const uint64_t data_len = esz ? n_rows * (uint64_t)esz : (n_rows + 7) / 8;
    const uint64_t valid_len = any_valid ? (n_rows + 7) / 8 : 0;
    // device copies padded to 16 bytes, so vector loads never run past them
    const size_t bytes = (data_len + 16) + (any_valid ? valid_len + 16 : 0);
    make_room(c, bytes);
    CcEntry e{c.next_id++, nullptr, nullptr, n_rows, data_len, valid_len, bytes, 1, c.tick};
    PMH_HIP(hipMalloc(&e.data, data_len + 16));
    if (any_valid) {
        const hipError_t err = hipMalloc((void **)&e.valid, valid_len + 16);
        if (err != hipSuccess) {
            (void)hipFree(e.data);
            PMH_HIP(err);
        }
    }
    try {
        const uint64_t before = h2d_bytes();
        if (esz) {
            uint64_t row0 = 0;
            for (uint32_t i = 0; i < n_chunks; ++i) {
                upload_host((uint8_t *)e.data + row0 * esz, chunks[i].data, chunks[i].n_rows * (uint64_t)esz);
                row0 += chunks[i].n_rows;
            }
        } else {
            assemble((uint8_t *)e.data, n_rows, data_len + 16, chunks, n_chunks, 1);
        }
        if (any_valid) assemble(e.valid, n_rows, valid_len + 16, chunks, n_chunks, 0);
        c.uploaded += h2d_bytes() - before;
    } catch (...) {
        Runtime &R = rt();
        (void)hipStreamSynchronize(R.stream);
        (void)hipFree(e.data);
        if (e.valid) (void)hipFree(e.valid);
        throw;
    }
    c.resident += bytes;
    c.by_id[e.id] = key;
    auto ins = c.map.emplace(key, e).first;
    fill_out(ins->second, 0, out);
}

void column_release(uint64_t entry) {
    ColCache &c = cc();
    auto k = c.by_id.find(entry);
    if (k == c.by_id.end()) fail(PMH_E_KEY, "polars_metal: no resident column " + std::to_string(entry));
    CcEntry &e = c.map.at(k->second);
    if (e.holds <= 0) fail(PMH_E_VALUE, "polars_metal: resident column " + std::to_string(entry) + " is not held");
    --e.holds;
}

void column_evict(uint64_t entry) {
    ColCache &c = cc();
    auto k = c.by_id.find(entry);
    if (k == c.by_id.end()) return;  // already evicted
    auto it = c.map.find(k->second);
    if (it->second.holds > 0)
        fail(PMH_E_VALUE, "polars_metal: resident column " + std::to_string(entry) + " is held by a query");
    drop_entry(c, it);
}

void colcache_clear() {
    ColCache &c = cc();
    for (auto it = c.map.begin(); it != c.map.end();) {
        auto cur = it++;
        if (cur->second.holds == 0) drop_entry(c, cur);
    }
}

void colcache_set_capacity(uint64_t bytes) {
    ColCache &c = cc();
    c.capacity = bytes;
    make_room(c, 0);
}

void colcache_stats(pmh_colcache_info *out) {
    ColCache &c = cc();
    out->entries = c.map.size();
    out->bytes_resident = c.resident;
    out->capacity = c.capacity;
    out->hits = c.hits;
    out->misses = c.misses;
    out->bytes_uploaded = c.uploaded;
    out->h2d_bytes = h2d_bytes();
}

}  // namespace pmh
