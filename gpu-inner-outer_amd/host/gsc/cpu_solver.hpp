// gsc/cpu_solver.hpp — the host solver behind the CLI's --cpu and --compare
// modes: the reference's prepare_case_cpu + solve_backward_cpu
// (abstraction.hpp:161-209, 287-306; reachability.hpp:31-155) with the same
// results, multithreaded over states. It is part of the CLI only, never of
// libgsc_amd.so or libgsc_hip.so: the GPU entry points have no host fallback.
#pragma once

#include <cstdint>

#include "gsc/hip_api.hpp"

namespace gsc {

struct CpuRunReport {
    Grid4D<std::uint32_t> state_grid;
    InputGrid2D input_grid;
    MemoryBudget budget;              // the reference's estimate (abstraction.hpp:211-245)
    std::uint64_t pair_count = 0;
    std::uint64_t valid_pairs = 0;    // pairs whose successor box is in the domain and passes the constraint
    double abstraction_ms = 0.0;      // interval tables and state hulls (the reference times build_abstraction_cpu)
    unsigned threads = 0;
    SolveResult<std::uint32_t> result;
};

// Host threads for run_case_cpu_u32 when none are asked for: GSC_CPU_THREADS,
// else OMP_NUM_THREADS, else the hardware's count.
unsigned default_cpu_threads();

// Solves the case on the host. Throws std::runtime_error for configuration
// errors, with the reference's messages where the reference has one. Prints
// the reference's progress lines ("CPU: ...") and, when cfg.verbose, its
// per-iteration lines (reachability.hpp:113-128). threads == 0 takes
// default_cpu_threads().
CpuRunReport run_case_cpu_u32(const CaseConfig& cfg, const CaseExtensions& ext, unsigned threads = 0);

}  // namespace gsc
