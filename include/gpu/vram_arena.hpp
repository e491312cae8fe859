// One persistent hipMalloc slab with a first-fit free list.
// Interface and semantics of src/gpu/vram_arena.cuh:21-75 / vram_arena.cu:27-153 (reference):
//  * 256-byte aligned first-fit allocation, zero-filled on allocation;
//  * allocate(0) -> nullptr; exhausted or fragmented -> std::bad_alloc;
//  * free(nullptr) is a no-op; foreign / double free -> std::invalid_argument;
//  * freed blocks are coalesced with both neighbours;
//  * occupancy_percentage() = round(100 * allocated / capacity);
//  * host calls are serialised by a mutex; the object is neither copyable nor movable.
#pragma once
#include <cstddef>
#include <limits>
#include <mutex>
#include <new>
#include <unordered_map>
#include <vector>

namespace sankhya {
namespace gpu {

class VRAMArena {
public:
    explicit VRAMArena(std::size_t capacity_bytes = 256ULL * 1024ULL * 1024ULL);
    ~VRAMArena();
    VRAMArena(const VRAMArena&) = delete;
    VRAMArena& operator=(const VRAMArena&) = delete;

    void* allocate(std::size_t bytes);
    void free(void* ptr);
    std::size_t occupancy_percentage() const;

    // Extensions: allocation without the zero fill (for buffers fully overwritten before use),
    // and capacity / usage queries for plan-sized arenas.
    void* allocate_uninitialized(std::size_t bytes);
    std::size_t capacity() const { return capacity_; }
    std::size_t allocated_bytes() const;

    static constexpr std::size_t kAlignment = 256;

private:
    struct Span { std::size_t offset, size; };
    void* allocate_impl(std::size_t bytes, bool zero);

    char* base_ = nullptr;
    std::size_t capacity_ = 0;
    std::size_t in_use_ = 0;
    std::vector<Span> free_spans_;                        // sorted by offset, never adjacent
    std::unordered_map<std::size_t, std::size_t> live_;   // offset -> aligned size
    mutable std::mutex mu_;
};

}  // namespace gpu
}  // namespace sankhya
