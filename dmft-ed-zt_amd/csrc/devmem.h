// Device memory of the batched engine.
//
// Problems come and go while a queue runs, each needing one contiguous
// region (tables, Krylov basis, rings) of a few MB to tens of GB.  Returning
// device memory to the driver is slow on MI355X (unmapping costs about 4 ms
// per GB, and the stream-ordered pool releases freed blocks instead of
// keeping them), so the engine never does it while it runs: it takes large
// chunks with hipMalloc and sub-allocates them itself.  A queue owns the
// chunks it took; when it is destroyed they go to a per-device cache from
// which later queues of the process take them first.  Chunks are returned
// to the driver only when a request cannot be met otherwise, or by
// devmem_release_cache().
#pragma once
#include <cstddef>
#include <vector>

namespace edg {

struct DevChunk;

class DevHeap {
public:
    explicit DevHeap(int device) : device_(device) {}
    ~DevHeap();  // chunks go to the device cache; the caller has synchronized their users
    DevHeap(const DevHeap &) = delete;
    DevHeap &operator=(const DevHeap &) = delete;
    // 0: *p is a region of at least `bytes` (2 MiB aligned); 1: the device
    // cannot hold it now; -1: HIP error.  Must be called with this heap's
    // device current.
    int alloc(size_t bytes, void **p);
    // Return a region from alloc (stream order: the next user of the region
    // must be ordered after the last user, e.g. on the same stream).
    void free(void *p);
    size_t owned_bytes() const;
    // Make the free bytes of the owned chunks at least `bytes`, with one
    // chunk (from the device cache, else the driver) when they fall short.
    // 0: done or not possible now (alloc then grows as it can); -1: HIP error.
    int reserve(size_t bytes);

private:
    int device_;
    std::vector<DevChunk *> chunks_;
};

// Bytes of idle chunks in the device's cache (free for the next queue).
size_t devmem_cached_bytes(int device);
// Return the device's idle cached chunks to the driver; returns the number
// of chunks whose hipFree failed.
int devmem_release_cache(int device);

}  // namespace edg
