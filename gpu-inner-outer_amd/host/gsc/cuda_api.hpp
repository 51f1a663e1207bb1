// gsc/cuda_api.hpp — the GPU backend boundary with the reference's names and
// layouts (include/gsc/cuda_api.hpp:11-88, reachability.hpp:15-28,
// grid.hpp:12-79, abstraction.hpp:16-30), implemented by libgsc_hip.so over
// the C ABI of libgsc_amd.so (gsc_run_case_u32).
//
// Two ways to use it:
//   - binary drop-in: a program compiled against the reference's own headers
//     (which also carry its CPU solver) links libgsc_hip.so in place of the
//     reference's CUDA library; the structs here have the same members in the
//     same order, so both sides agree on every layout;
//   - source mirror: a program that only calls run_case_cuda_u32, the config
//     functions and the HDF5 writer compiles against these headers.
#pragma once

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "gsc/config.hpp"

namespace gsc {

template <typename StateIndex>
struct Grid4D {
    double lb[kStateDim]{};
    double ub[kStateDim]{};
    double eta[kStateDim]{};
    std::uint32_t shape[kStateDim]{};
    std::uint64_t strides[kStateDim]{};
    std::uint64_t total_size = 0;
    int wrap_dim = 2;
};

struct InputGrid2D {
    double lb[kInputDim]{};
    double ub[kInputDim]{};
    double eta[kInputDim]{};
    std::uint32_t shape[kInputDim]{};
    std::uint64_t strides[kInputDim]{};
    std::uint64_t total_size = 0;

    // Cell centre of input `flat` (grid.hpp:68-78).
    void center(InputIndex flat, double out[kInputDim]) const {
        std::uint64_t v = flat;
        std::uint32_t idx[kInputDim]{};
        for (int d = kInputDim - 1; d >= 0; --d) {
            idx[d] = static_cast<std::uint32_t>(v / strides[d]);
            v %= strides[d];
        }
        for (int d = 0; d < kInputDim; ++d) out[d] = lb[d] + (static_cast<double>(idx[d]) + 0.5) * eta[d];
    }
};

struct MemoryBudget {
    std::uint64_t pair_min_bytes = 0;
    std::uint64_t pair_max_bytes = 0;
    std::uint64_t pair_valid_bytes = 0;
    std::uint64_t pair_satisfied_bytes = 0;
    std::uint64_t reachable_prev_bytes = 0;
    std::uint64_t reachable_next_bytes = 0;
    std::uint64_t valid_mask_bytes = 0;
    std::uint64_t candidate_mask_bytes = 0;
    std::uint64_t controller_bytes = 0;
    std::uint64_t reach_step_bytes = 0;
    std::uint64_t prefix_reachable_bytes = 0;
    std::uint64_t prefix_valid_bytes = 0;
    std::uint64_t total_bytes = 0;
};

template <typename StateIndex>
struct SolveResult {
    std::vector<std::uint8_t> valid_mask;
    std::vector<std::uint8_t> candidate_mask;
    std::vector<std::uint8_t> reachable_mask;
    std::vector<InputIndex> controller;
    std::vector<ReachStep> reach_step;
    std::uint32_t iterations = 0;
    bool converged = false;
    std::string message;
    std::uint64_t reachable_states = 0;
    std::uint64_t certified_candidate_states = 0;
    double solve_ms = 0.0;
};

struct IterationStats {
    std::uint32_t iteration = 0;
    std::uint64_t newly_reachable = 0;
    std::uint64_t newly_certified = 0;
    std::uint64_t total_reachable = 0;
    std::uint64_t total_certified = 0;
    double iteration_ms = 0.0;
    double prefix_build_ms = 0.0;
    double satisfaction_check_ms = 0.0;
    double reduction_ms = 0.0;
};

struct KernelTimings {
    double abstraction_kernel_ms = 0.0;
    double abstraction_memcpy_h2d_ms = 0.0;
    double scatter_mask_ms = 0.0;
    double scan_axis0_ms = 0.0;
    double scan_axis1_ms = 0.0;
    double scan_axis2_ms = 0.0;
    double scan_axis3_ms = 0.0;
    double prefix_build_total_ms = 0.0;
    double pair_satisfaction_ms = 0.0;
    double reduce_inputs_ms = 0.0;
    double iteration_memcpy_ms = 0.0;
    double result_memcpy_d2h_ms = 0.0;
    double total_kernel_ms = 0.0;
    double total_memcpy_ms = 0.0;
    double total_compute_ms = 0.0;
};

struct GpuMemoryStats {
    std::uint64_t total_allocated_bytes = 0;
    std::uint64_t peak_usage_bytes = 0;
    std::uint64_t abstraction_bytes = 0;
    std::uint64_t prefix_bytes = 0;
    std::uint64_t iteration_bytes = 0;
    std::uint64_t result_bytes = 0;
};

struct GpuRunReport {
    bool compiled_with_cuda = false;  // name kept for drop-in callers; true for this HIP backend
    bool runtime_available = false;
    bool executed = false;
    std::string message;
    double abstraction_ms = 0.0;
    double solve_ms = 0.0;
    double total_ms = 0.0;
    SolveResult<std::uint32_t> result;
    Grid4D<std::uint32_t> state_grid;
    InputGrid2D input_grid;
    MemoryBudget budget;
    std::uint64_t pair_count = 0;
    std::vector<IterationStats> iteration_stats;
    KernelTimings kernel_timings;
    GpuMemoryStats memory_stats;
    std::vector<std::uint32_t> abstraction_min_flat;
    std::vector<std::uint32_t> abstraction_max_flat;
    std::vector<std::uint8_t> abstraction_valid;
};

// The two backend symbols (cuda_api.hpp:87-88), exported by libgsc_hip.so with
// the reference's exact signatures. Like the reference, preparation errors
// (grid shape, candidate outside the valid set, ...) are thrown as
// std::runtime_error, and device errors come back in report.message with
// executed == false. The report always carries the per-pair abstraction
// triplet, as the reference's does; gsc::run_case_hip_u32 (gsc/hip_api.hpp)
// can leave it out.
bool cuda_backend_compiled();
GpuRunReport run_case_cuda_u32(const CaseConfig& cfg);

// abstraction.hpp:52-63 (inline there as well).
inline std::string format_bytes(std::uint64_t bytes) {
    static const char* kUnits[] = {"B", "KiB", "MiB", "GiB", "TiB"};
    double value = static_cast<double>(bytes);
    int unit = 0;
    while (value >= 1024.0 && unit < 4) {
        value /= 1024.0;
        ++unit;
    }
    char buf[64];
    std::snprintf(buf, sizeof(buf), "%.2f %s", value, kUnits[unit]);
    return buf;
}

}  // namespace gsc
