// Prints the size and member offsets of every type that crosses the
// run_case_cuda_u32 boundary (config.hpp:41-67, cuda_api.hpp:11-85,
// reachability.hpp:15-28, grid.hpp:12-79, abstraction.hpp:16-30). Compiled
// once against the reference's headers and once against this backend's
// mirror headers; the two outputs must be identical for a program built
// against one to pass and receive these objects across libgsc_hip.so.
#include <cstddef>
#include <cstdio>

#include "gsc/cuda_api.hpp"

#define SZ(T) std::printf("sizeof %s %zu\n", #T, sizeof(T))
#define OFF(T, m) std::printf("offsetof %s.%s %zu\n", #T, #m, offsetof(T, m))

int main() {
    using namespace gsc;
    using Grid = Grid4D<std::uint32_t>;
    using Result = SolveResult<std::uint32_t>;
    SZ(HyperRect4D);
    SZ(CaseConfig);
    OFF(CaseConfig, name);
    OFF(CaseConfig, state_lb);
    OFF(CaseConfig, state_ub);
    OFF(CaseConfig, state_eta);
    OFF(CaseConfig, wrap_dim);
    OFF(CaseConfig, input_lb);
    OFF(CaseConfig, input_ub);
    OFF(CaseConfig, input_eta);
    OFF(CaseConfig, candidate);
    OFF(CaseConfig, map);
    OFF(CaseConfig, obstacles);
    OFF(CaseConfig, constraint_type);
    OFF(CaseConfig, hyperbolic_params);
    OFF(CaseConfig, elliptic_params);
    OFF(CaseConfig, custom_constraint);
    OFF(CaseConfig, disturbance_half_width);
    OFF(CaseConfig, sample_time);
    OFF(CaseConfig, integration_substeps);
    OFF(CaseConfig, max_iterations);
    OFF(CaseConfig, verbose);
    SZ(Grid);
    OFF(Grid, lb);
    OFF(Grid, ub);
    OFF(Grid, eta);
    OFF(Grid, shape);
    OFF(Grid, strides);
    OFF(Grid, total_size);
    OFF(Grid, wrap_dim);
    SZ(InputGrid2D);
    OFF(InputGrid2D, shape);
    OFF(InputGrid2D, strides);
    OFF(InputGrid2D, total_size);
    SZ(MemoryBudget);
    OFF(MemoryBudget, total_bytes);
    SZ(Result);
    OFF(Result, valid_mask);
    OFF(Result, candidate_mask);
    OFF(Result, reachable_mask);
    OFF(Result, controller);
    OFF(Result, reach_step);
    OFF(Result, iterations);
    OFF(Result, converged);
    OFF(Result, message);
    OFF(Result, reachable_states);
    OFF(Result, certified_candidate_states);
    OFF(Result, solve_ms);
    SZ(IterationStats);
    OFF(IterationStats, newly_reachable);
    OFF(IterationStats, reduction_ms);
    SZ(KernelTimings);
    OFF(KernelTimings, total_compute_ms);
    SZ(GpuMemoryStats);
    OFF(GpuMemoryStats, result_bytes);
    SZ(GpuRunReport);
    OFF(GpuRunReport, compiled_with_cuda);
    OFF(GpuRunReport, runtime_available);
    OFF(GpuRunReport, executed);
    OFF(GpuRunReport, message);
    OFF(GpuRunReport, abstraction_ms);
    OFF(GpuRunReport, solve_ms);
    OFF(GpuRunReport, total_ms);
    OFF(GpuRunReport, result);
    OFF(GpuRunReport, state_grid);
    OFF(GpuRunReport, input_grid);
    OFF(GpuRunReport, budget);
    OFF(GpuRunReport, pair_count);
    OFF(GpuRunReport, iteration_stats);
    OFF(GpuRunReport, kernel_timings);
    OFF(GpuRunReport, memory_stats);
    OFF(GpuRunReport, abstraction_min_flat);
    OFF(GpuRunReport, abstraction_max_flat);
    OFF(GpuRunReport, abstraction_valid);
    return 0;
}
