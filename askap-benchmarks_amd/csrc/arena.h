// First-fit sub-allocator over one device block (host-side bookkeeping only):
// the serving loop's HBM budget.  Offsets and sizes are rounded up to 256 B;
// freed ranges coalesce with their neighbours.  Not thread-safe (the server
// holds its lock around every call).
#pragma once

#include <cstddef>
#include <iterator>
#include <map>

namespace tcv_detail {

class Arena {
public:
    static size_t round(size_t x) { return (x + 255) & ~size_t(255); }

    void init(size_t bytes)
    {
        free_.clear();
        cap_ = bytes & ~size_t(255);
        used_ = 0;
        if (cap_) free_[0] = cap_;
    }

    // Reserves `bytes`; returns false (and changes nothing) when no free range fits.
    bool alloc(size_t bytes, size_t *off)
    {
        bytes = round(bytes ? bytes : 1);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < bytes) continue;
            *off = it->first;
            const size_t rest = it->second - bytes, start = it->first + bytes;
            free_.erase(it);
            if (rest) free_[start] = rest;
            used_ += bytes;
            return true;
        }
        return false;
    }

    void release(size_t off, size_t bytes)
    {
        bytes = round(bytes ? bytes : 1);
        used_ -= bytes;
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

    bool fits_ever(size_t bytes) const { return round(bytes ? bytes : 1) <= cap_; }
    size_t capacity() const { return cap_; }
    size_t used() const { return used_; }
    size_t free_ranges() const { return free_.size(); }

private:
    std::map<size_t, size_t> free_;   // offset -> length
    size_t cap_ = 0, used_ = 0;
};

}  // namespace tcv_detail
