// Host ring allocator for finished proofs (pure host logic, no device code).
#pragma once
#include <stdint.h>

#include <deque>

namespace vx {

// Pinned host memory for finished proofs, allocated in batch-sized regions in
// ring order; a region is reused once every proof in it has been released.
struct ResultRing {
    uint8_t* base = nullptr;
    uint64_t size = 0;
    struct Region {
        uint64_t off, bytes;
        uint32_t live;
    };
    std::deque<Region> regions;  // allocation order
    uint64_t first_id = 0;       // id of regions.front()

    // returns the region id, or -1 when there is no contiguous room now
    int64_t alloc(uint64_t bytes, uint32_t live) {
        if (!bytes || bytes > size) return -1;
        uint64_t off = 0;
        if (!regions.empty()) {
            const uint64_t tail = regions.front().off, head = regions.back().off + regions.back().bytes;
            if (head > tail) {  // live data in [tail, head)
                if (head + bytes <= size)
                    off = head;
                else if (bytes <= tail)
                    off = 0;
                else
                    return -1;
            } else {  // wrapped: live data in [tail, size) and [0, head)
                if (head + bytes <= tail)
                    off = head;
                else
                    return -1;
            }
        }
        regions.push_back({off, bytes, live});
        return (int64_t)(first_id + regions.size() - 1);
    }
    uint8_t* at(int64_t id) const { return base + regions[id - first_id].off; }
    void release(int64_t id) {
        if (id < (int64_t)first_id || id >= (int64_t)(first_id + regions.size())) return;
        Region& r = regions[id - first_id];
        if (r.live) r.live--;
        while (!regions.empty() && regions.front().live == 0) {
            regions.pop_front();
            first_id++;
        }
    }
};

}  // namespace vx
