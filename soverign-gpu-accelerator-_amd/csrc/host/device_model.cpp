// upload_to_device (src/gpu/device_model.cu:21-80 is the reference's): arena allocations and
// synchronous host-to-device copies.
#include <limits>
#include <stdexcept>
#include <string>
#include <vector>

#include "gpu/device_model.hpp"
#include "lakshya_device.h"

namespace sankhya {
namespace gpu {

namespace {
void upload(void* dst, const void* src, std::size_t bytes) {
    int e = lkd_memcpy_htod(dst, src, bytes, nullptr);
    if (!e) e = lkd_stream_sync(nullptr);
    if (e) throw std::runtime_error(std::string("upload_to_device: ") + lkd_error_string(e));
}
}  // namespace

DeviceModel upload_to_device(const core::Model& host_model, VRAMArena& arena) {
    const core::CSCMatrix& A = host_model.A;
    DeviceModel d;
    d.rows = A.rows;
    d.cols = A.cols;
    d.nnz = static_cast<Index>(A.values.size());
    if (d.rows == 0 && d.cols == 0) return d;
    if (static_cast<Index>(A.col_ptrs.size()) != d.cols + 1 ||
        static_cast<Index>(A.row_indices.size()) != d.nnz ||
        static_cast<Index>(host_model.obj.size()) != d.cols ||
        (!host_model.rhs.empty() && static_cast<Index>(host_model.rhs.size()) != d.rows))
        throw std::invalid_argument("upload_to_device: the model is not finalized");
    const std::size_t ncol = static_cast<std::size_t>(d.cols), nrow = static_cast<std::size_t>(d.rows);
    d.col_ptrs = static_cast<Index*>(arena.allocate_uninitialized(sizeof(Index) * (ncol + 1)));
    upload(d.col_ptrs, A.col_ptrs.data(), sizeof(Index) * (ncol + 1));
    if (d.nnz > 0) {
        d.row_indices = static_cast<Index*>(arena.allocate_uninitialized(sizeof(Index) * d.nnz));
        d.values = static_cast<Float*>(arena.allocate_uninitialized(sizeof(Float) * d.nnz));
        upload(d.row_indices, A.row_indices.data(), sizeof(Index) * d.nnz);
        upload(d.values, A.values.data(), sizeof(Float) * d.nnz);
    }
    if (d.cols > 0) {
        std::vector<Float> lb(ncol, 0.0), ub(ncol, std::numeric_limits<Float>::infinity());
        for (std::size_t j = 0; j < ncol && j < host_model.lb.size(); ++j) lb[j] = host_model.lb[j];
        for (std::size_t j = 0; j < ncol && j < host_model.ub.size(); ++j) ub[j] = host_model.ub[j];
        d.obj = static_cast<Float*>(arena.allocate_uninitialized(sizeof(Float) * ncol));
        d.lb = static_cast<Float*>(arena.allocate_uninitialized(sizeof(Float) * ncol));
        d.ub = static_cast<Float*>(arena.allocate_uninitialized(sizeof(Float) * ncol));
        upload(d.obj, host_model.obj.data(), sizeof(Float) * ncol);
        upload(d.lb, lb.data(), sizeof(Float) * ncol);
        upload(d.ub, ub.data(), sizeof(Float) * ncol);
    }
    if (d.rows > 0 && !host_model.rhs.empty()) {
        d.rhs = static_cast<Float*>(arena.allocate_uninitialized(sizeof(Float) * nrow));
        upload(d.rhs, host_model.rhs.data(), sizeof(Float) * nrow);
    }
    return d;
}

}  // namespace gpu
}  // namespace sankhya
