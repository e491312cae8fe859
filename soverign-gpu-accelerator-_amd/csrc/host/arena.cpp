// hipMalloc-backed slab allocator with the reference arena's contract
// (src/gpu/vram_arena.cu:27-153): first fit, 256-byte granules, zero fill on allocation,
// coalescing free list, std::bad_alloc / std::invalid_argument errors, mutex-serialised.
#include <algorithm>
#include <cmath>
#include <stdexcept>
#include <string>
#include "gpu/vram_arena.hpp"
#include "lakshya_device.h"

namespace sankhya {
namespace gpu {

VRAMArena::VRAMArena(std::size_t capacity_bytes) : capacity_(capacity_bytes) {
    void* p = nullptr;
    const int e = lkd_malloc(&p, capacity_bytes);
    if (e != 0)
        throw std::runtime_error(std::string("VRAMArena: hipMalloc of the slab failed: ") +
                                 lkd_error_string(e));
    base_ = static_cast<char*>(p);
    if (capacity_ > 0) free_spans_.push_back({0, capacity_});
}

VRAMArena::~VRAMArena() {
    if (base_) lkd_free(base_);
}

void* VRAMArena::allocate(std::size_t bytes) { return allocate_impl(bytes, true); }
void* VRAMArena::allocate_uninitialized(std::size_t bytes) { return allocate_impl(bytes, false); }

void* VRAMArena::allocate_impl(std::size_t bytes, bool zero) {
    if (bytes == 0) return nullptr;
    std::lock_guard<std::mutex> lock(mu_);
    if (bytes > std::numeric_limits<std::size_t>::max() - (kAlignment - 1)) throw std::bad_alloc();
    const std::size_t need = (bytes + kAlignment - 1) & ~(kAlignment - 1);
    if (in_use_ + need > capacity_) throw std::bad_alloc();
    for (auto it = free_spans_.begin(); it != free_spans_.end(); ++it) {
        if (it->size < need) continue;
        const std::size_t off = it->offset;
        if (it->size == need) free_spans_.erase(it);
        else { it->offset += need; it->size -= need; }
        in_use_ += need;
        live_[off] = need;
        char* p = base_ + off;
        if (zero) {
            // The fill is finished before the pointer is handed out: the callers' streams are
            // non-blocking, so a fill left queued on the null stream would be ordered with
            // none of their work and could land after (and wipe) their first upload.
            int e = lkd_memset(p, 0, need, nullptr);
            if (e == 0) e = lkd_stream_sync(nullptr);
            if (e != 0) throw std::runtime_error(std::string("VRAMArena: zero fill failed: ") +
                                                 lkd_error_string(e));
        }
        return p;
    }
    throw std::bad_alloc();
}

void VRAMArena::free(void* ptr) {
    if (ptr == nullptr) return;
    std::lock_guard<std::mutex> lock(mu_);
    const char* p = static_cast<const char*>(ptr);
    if (p < base_ || p >= base_ + capacity_)
        throw std::invalid_argument("VRAMArena::free: the pointer does not lie in this arena's slab");
    const std::size_t off = static_cast<std::size_t>(p - base_);
    auto it = live_.find(off);
    if (it == live_.end())
        throw std::invalid_argument("VRAMArena::free: no live allocation starts at this pointer (freed twice, or interior)");
    const std::size_t size = it->second;
    live_.erase(it);
    in_use_ -= size;
    auto pos = std::lower_bound(free_spans_.begin(), free_spans_.end(), off,
                                [](const Span& s, std::size_t o) { return s.offset < o; });
    pos = free_spans_.insert(pos, Span{off, size});
    auto next = pos + 1;
    if (next != free_spans_.end() && pos->offset + pos->size == next->offset) {
        pos->size += next->size;
        free_spans_.erase(next);
    }
    if (pos != free_spans_.begin()) {
        auto prev = pos - 1;
        if (prev->offset + prev->size == pos->offset) {
            prev->size += pos->size;
            free_spans_.erase(pos);
        }
    }
}

std::size_t VRAMArena::occupancy_percentage() const {
    std::lock_guard<std::mutex> lock(mu_);
    if (capacity_ == 0) return 0;
    return static_cast<std::size_t>(
        std::round(static_cast<double>(in_use_) / static_cast<double>(capacity_) * 100.0));
}

std::size_t VRAMArena::allocated_bytes() const {
    std::lock_guard<std::mutex> lock(mu_);
    return in_use_;
}

}  // namespace gpu
}  // namespace sankhya
