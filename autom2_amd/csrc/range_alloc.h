// range_alloc.h -- first-fit allocator of byte ranges [0, size), the
// bookkeeping of the server's device arena.  Host code only (no HIP), so it
// is unit-tested on the CPU under the address and undefined-behaviour
// sanitizers.
#pragma once

#include <cstddef>
#include <cstdint>
#include <iterator>
#include <map>

namespace trec {

class RangeAllocator {
  public:
    static constexpr size_t NONE = SIZE_MAX;
    static constexpr size_t ALIGN = 256;
    static size_t align_up(size_t v) { return (v + ALIGN - 1) & ~(ALIGN - 1); }

    void reset(size_t size)
    {
        free_.clear();
        used_.clear();
        size_ = size;
        if (size) free_[0] = size;
    }
    size_t size() const { return size_; }
    // Offset of the first free range of align_up(bytes) bytes, or NONE.
    size_t alloc(size_t bytes)
    {
        bytes = align_up(bytes);
        if (bytes == 0) bytes = ALIGN;
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second >= bytes) {
                const size_t off = it->first, len = it->second;
                free_.erase(it);
                if (len > bytes) free_[off + bytes] = len - bytes;
                used_[off] = bytes;
                return off;
            }
        }
        return NONE;
    }
    // Returns a range from alloc to the free list, merging it with its free
    // neighbours; offsets that are not allocated are ignored.
    void release(size_t off)
    {
        auto u = used_.find(off);
        if (u == used_.end()) return;
        size_t len = u->second;
        used_.erase(u);
        auto next = free_.lower_bound(off);
        if (next != free_.end() && next->first == off + len) {
            len += next->second;
            free_.erase(next);
        }
        auto it = free_.emplace(off, len).first;
        if (it != free_.begin()) {
            auto prev = std::prev(it);
            if (prev->first + prev->second == off) {
                prev->second += len;
                free_.erase(it);
            }
        }
    }
    bool fits(size_t bytes) const
    {
        bytes = align_up(bytes);
        for (const auto &f : free_)
            if (f.second >= bytes) return true;
        return false;
    }
    // Invariants: free and used ranges tile [0, size) without overlap, and
    // no two free ranges touch.
    bool consistent() const
    {
        std::map<size_t, std::pair<size_t, bool>> all;
        for (const auto &f : free_) all[f.first] = {f.second, true};
        for (const auto &u : used_) {
            if (all.count(u.first)) return false;
            all[u.first] = {u.second, false};
        }
        size_t at = 0;
        bool prev_free = false;
        for (const auto &a : all) {
            if (a.first != at || a.second.first == 0) return false;
            if (prev_free && a.second.second) return false;
            prev_free = a.second.second;
            at += a.second.first;
        }
        return at == size_;
    }

  private:
    size_t size_ = 0;
    std::map<size_t, size_t> free_, used_;
};

}  // namespace trec
