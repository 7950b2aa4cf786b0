// arena.h -- the batch's device-memory arena allocator (host side): first
// fit over [0, size) with coalescing of freed blocks.  Offsets only; the
// batch owns the device allocation they index.
#pragma once

#include <cstddef>
#include <iterator>
#include <map>

namespace cfs {

// First-fit allocator over [0, size) with coalescing of freed blocks.
class Arena {
public:
    void reset(size_t size) {
        free_.clear();
        if (size) free_[0] = size;
    }
    bool alloc(size_t bytes, size_t* off) {
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < bytes) continue;
            *off = it->first;
            const size_t rest = it->second - bytes;
            const size_t next = it->first + bytes;
            free_.erase(it);
            if (rest) free_[next] = rest;
            return true;
        }
        return false;
    }
    void release(size_t off, size_t bytes) {
        auto it = free_.emplace(off, bytes).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
    size_t free_bytes() const {
        size_t s = 0;
        for (auto& kv : free_) s += kv.second;
        return s;
    }

private:
    std::map<size_t, size_t> free_;   // offset -> size
};

}  // namespace cfs
