// Chunked device memory of the batched engine (see devmem.h).
#include "devmem.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <iterator>
#include <map>
#include <mutex>

namespace edg {

namespace {

constexpr size_t kAlign = size_t(2) << 20;   // region granularity (2 MiB)
// smallest chunk taken from the driver: chunks left by small problems must
// still hold the large ones (an Ns=14 problem with its Krylov basis is ~10 GB)
constexpr size_t kMinChunk = size_t(16) << 30;
constexpr int kMaxDevices = 64;

size_t round_up(size_t v) { return (v + kAlign - 1) / kAlign * kAlign; }

}  // namespace

struct DevChunk {
    char *base = nullptr;
    size_t size = 0;
    std::map<size_t, size_t> free;  // offset -> length, coalesced
    std::map<size_t, size_t> used;  // offset -> length

    void reset()
    {
        free.clear();
        used.clear();
        free.emplace(0, size);
    }
    // Smallest free range of at least n bytes: offset, or size when none.
    size_t best_fit(size_t n, size_t &len) const
    {
        size_t best = size;
        len = 0;
        for (const auto &f : free)
            if (f.second >= n && (best == size || f.second < len)) {
                best = f.first;
                len = f.second;
            }
        return best;
    }
    char *carve(size_t off, size_t n)
    {
        const size_t len = free.at(off);
        free.erase(off);
        if (len > n) free.emplace(off + n, len - n);
        used.emplace(off, n);
        return base + off;
    }
    void release(size_t off)
    {
        auto u = used.find(off);
        if (u == used.end()) return;
        size_t start = off, len = u->second;
        used.erase(u);
        auto next = free.lower_bound(start);
        if (next != free.end() && next->first == start + len) {
            len += next->second;
            next = free.erase(next);
        }
        if (next != free.begin()) {
            auto prev = std::prev(next);
            if (prev->first + prev->second == start) {
                start = prev->first;
                len += prev->second;
                free.erase(prev);
            }
        }
        free.emplace(start, len);
    }
};

namespace {

std::mutex g_mu;
std::vector<DevChunk *> g_cache[kMaxDevices];

DevChunk *new_chunk(size_t size)
{
    void *p = nullptr;
    const hipError_t e = hipMalloc(&p, size);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return nullptr;
    }
    DevChunk *c = new DevChunk;
    c->base = static_cast<char *>(p);
    c->size = size;
    c->reset();
    return c;
}

}  // namespace

int DevHeap::alloc(size_t bytes, void **p)
{
    *p = nullptr;
    if (device_ < 0 || device_ >= kMaxDevices) return -1;
    const size_t n = round_up(std::max<size_t>(bytes, 1));
    // 1. best fit over the chunks this heap owns
    DevChunk *bc = nullptr;
    size_t boff = 0, blen = 0;
    for (DevChunk *c : chunks_) {
        size_t len;
        const size_t off = c->best_fit(n, len);
        if (off < c->size && (!bc || len < blen)) {
            bc = c;
            boff = off;
            blen = len;
        }
    }
    if (bc) {
        *p = bc->carve(boff, n);
        return 0;
    }
    // 2. the smallest idle chunk of the device cache that holds it
    {
        std::lock_guard<std::mutex> lk(g_mu);
        auto &cache = g_cache[device_];
        auto best = cache.end();
        for (auto it = cache.begin(); it != cache.end(); ++it)
            if ((*it)->size >= n && (best == cache.end() || (*it)->size < (*best)->size)) best = it;
        if (best != cache.end()) {
            DevChunk *c = *best;
            cache.erase(best);
            chunks_.push_back(c);
            *p = c->carve(0, n);
            return 0;
        }
    }
    // 3. a new chunk from the driver: kMinChunk or the request, else exactly
    //    the request after returning idle cached chunks (largest first, one at
    //    a time: unmapping costs about 4 ms per GB)
    DevChunk *c = new_chunk(std::max(n, kMinChunk));
    if (!c && n < kMinChunk) c = new_chunk(n);
    while (!c) {
        DevChunk *drop = nullptr;
        {
            std::lock_guard<std::mutex> lk(g_mu);
            auto &cache = g_cache[device_];
            auto big = std::max_element(cache.begin(), cache.end(),
                                        [](const DevChunk *a, const DevChunk *b) { return a->size < b->size; });
            if (big != cache.end()) {
                drop = *big;
                cache.erase(big);
            }
        }
        if (!drop) break;
        (void)hipFree(drop->base);
        delete drop;
        c = new_chunk(n);
    }
    if (!c) return 1;
    chunks_.push_back(c);
    *p = c->carve(0, n);
    return 0;
}

int DevHeap::reserve(size_t bytes)
{
    if (device_ < 0 || device_ >= kMaxDevices) return -1;
    size_t have = 0;
    for (const DevChunk *c : chunks_)
        for (const auto &f : c->free) have += f.second;
    if (have >= bytes) return 0;
    {
        // idle cached chunks first: the smallest that covers the rest, else
        // the largest, until the rest is covered or the cache is empty
        std::lock_guard<std::mutex> lk(g_mu);
        auto &cache = g_cache[device_];
        while (have < bytes && !cache.empty()) {
            const size_t rest = bytes - have;
            auto pick = cache.end();
            for (auto it = cache.begin(); it != cache.end(); ++it)
                if ((*it)->size >= rest && (pick == cache.end() || (*it)->size < (*pick)->size)) pick = it;
            if (pick == cache.end())
                pick = std::max_element(cache.begin(), cache.end(),
                                        [](const DevChunk *a, const DevChunk *b) { return a->size < b->size; });
            have += (*pick)->size;
            chunks_.push_back(*pick);
            cache.erase(pick);
        }
    }
    if (have >= bytes) return 0;
    DevChunk *c = new_chunk(std::max(round_up(bytes - have), kMinChunk));
    if (c) chunks_.push_back(c);
    return 0;
}

void DevHeap::free(void *p)
{
    if (!p) return;
    char *q = static_cast<char *>(p);
    for (DevChunk *c : chunks_)
        if (q >= c->base && q < c->base + c->size) {
            c->release((size_t)(q - c->base));
            return;
        }
}

size_t DevHeap::owned_bytes() const
{
    size_t s = 0;
    for (const DevChunk *c : chunks_) s += c->size;
    return s;
}

DevHeap::~DevHeap()
{
    if (device_ < 0 || device_ >= kMaxDevices) return;
    std::lock_guard<std::mutex> lk(g_mu);
    for (DevChunk *c : chunks_) {
        c->reset();
        g_cache[device_].push_back(c);
    }
    chunks_.clear();
}

size_t devmem_cached_bytes(int device)
{
    if (device < 0 || device >= kMaxDevices) return 0;
    std::lock_guard<std::mutex> lk(g_mu);
    size_t s = 0;
    for (const DevChunk *c : g_cache[device]) s += c->size;
    return s;
}

int devmem_release_cache(int device)
{
    if (device < 0 || device >= kMaxDevices) return 0;
    std::vector<DevChunk *> drop;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        drop.swap(g_cache[device]);
    }
    int failed = 0;
    for (DevChunk *c : drop) {
        if (hipFree(c->base) != hipSuccess) {
            (void)hipGetLastError();
            ++failed;
        }
        delete c;
    }
    return failed;
}

}  // namespace edg
