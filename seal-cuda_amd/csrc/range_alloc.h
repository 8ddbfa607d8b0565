// range_alloc.h -- first-fit allocator over one contiguous range (the server's device
// arena and pinned staging arena).  Offsets and sizes in bytes, in 256-byte granules;
// released blocks coalesce with free neighbours.  Not thread safe (the server holds its
// lock around it).  Header only and free of HIP, so host tests build it alone.
#pragma once
#include <algorithm>
#include <cstddef>
#include <cstdint>
#include <iterator>
#include <map>
#include <stdexcept>

namespace shc {

class RangeAllocator {
public:
    static constexpr size_t GRAIN = 256;
    explicit RangeAllocator(size_t bytes = 0) { reset(bytes); }

    void reset(size_t bytes) {
        free_.clear();
        taken_.clear();
        total_ = bytes / GRAIN * GRAIN;
        used_ = 0;
        if (total_) free_[0] = total_;
    }

    // Offset of a free block of at least `bytes` (the lowest that fits), or SIZE_MAX.
    size_t take(size_t bytes) {
        bytes = round_up(std::max<size_t>(bytes, 1));
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < bytes) continue;
            const size_t off = it->first, rest = it->second - bytes;
            free_.erase(it);
            if (rest) free_[off + bytes] = rest;
            taken_[off] = bytes;
            used_ += bytes;
            return off;
        }
        return SIZE_MAX;
    }

    void give(size_t off) {
        auto t = taken_.find(off);
        if (t == taken_.end()) throw std::logic_error("release of a block that is not taken");
        size_t size = t->second;
        taken_.erase(t);
        used_ -= size;
        auto next = free_.lower_bound(off);
        if (next != free_.end() && off + size == next->first) {
            size += next->second;
            next = free_.erase(next);
        }
        if (next != free_.begin()) {
            auto prev = std::prev(next);
            if (prev->first + prev->second == off) {
                prev->second += size;
                return;
            }
        }
        free_[off] = size;
    }

    size_t largest_free() const {
        size_t best = 0;
        for (const auto &f : free_) best = std::max(best, f.second);
        return best;
    }
    size_t in_use() const { return used_; }
    size_t capacity() const { return total_; }
    size_t free_blocks() const { return free_.size(); }

private:
    static size_t round_up(size_t x) { return (x + GRAIN - 1) / GRAIN * GRAIN; }
    std::map<size_t, size_t> free_;   // offset -> size
    std::map<size_t, size_t> taken_;  // offset -> size
    size_t total_ = 0, used_ = 0;
};

}  // namespace shc
