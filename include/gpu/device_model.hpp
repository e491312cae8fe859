// Device-resident copy of a core::Model: the same struct, members and upload function as
// src/gpu/device_model.cuh:26-58 (reference), so callers written against it (the crossover's)
// compile here.  A non-owning view: every array lies in the VRAMArena passed to
// upload_to_device, which must outlive it.
#pragma once

#include "core/problem.hpp"
#include "gpu/vram_arena.hpp"
#include "sankhya/types.hpp"

namespace sankhya {
namespace gpu {

struct DeviceModel {
    Index rows = 0;
    Index cols = 0;
    Index nnz = 0;
    Index* col_ptrs = nullptr;      // cols + 1
    Index* row_indices = nullptr;   // nnz
    Float* values = nullptr;        // nnz
    Float* obj = nullptr;           // cols
    Float* lb = nullptr;            // cols
    Float* ub = nullptr;            // cols
    Float* rhs = nullptr;           // rows
};

// Arrays allocated from `arena` and copied from the finalized host model (values bit-exact);
// an empty model gives an empty view.  Missing lb / ub entries upload as 0 / +inf.
DeviceModel upload_to_device(const core::Model& host_model, VRAMArena& arena);

}  // namespace gpu
}  // namespace sankhya
