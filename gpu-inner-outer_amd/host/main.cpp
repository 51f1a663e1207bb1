// main.cpp — cuda_symbolic_control_cli drop-in (src/main.cpp:125-218) on the
// gfx950 backend: same arguments, modes, summary and timing printout, HDF5
// output and exit codes (0 ok, 1 error or usage, 2 GPU run unavailable,
// 3 --compare mismatch).
//
//   cuda_symbolic_control_cli <config.cfg> [--cpu|--gpu|--compare] [--output result.h5]
//                             [--no-abstraction-export]
//
// --gpu runs the gfx950 backend. --cpu (the default, as in the reference)
// runs the host solver of gsc/cpu_solver.hpp; --compare runs both and exits 3
// unless the five result arrays agree exactly (main.cpp:114-121, 189-211).
// The host solver's threads: GSC_CPU_THREADS, else OMP_NUM_THREADS, else every
// hardware thread. --no-abstraction-export (an addition) skips the per-pair
// /abstraction datasets of --gpu, whose device-to-host copy and write dominate
// the reference's wall time at 401M pairs; --compare never exports them, as
// its file has no /abstraction group.
#include <algorithm>
#include <iostream>
#include <stdexcept>
#include <string>
#include <vector>

#include "gsc/cpu_solver.hpp"
#include "gsc/hip_api.hpp"
#include "gsc/result_io.hpp"

namespace {

void print_summary(const gsc::Grid4D<std::uint32_t>& state_grid, const gsc::InputGrid2D& input_grid,
                   const gsc::MemoryBudget& budget, std::uint64_t pair_count,
                   const std::vector<std::uint8_t>& candidate_mask, const gsc::SolveResult<std::uint32_t>& r,
                   double abstraction_ms, const std::string& label) {
    // main.cpp:12-39
    std::uint64_t candidate_states = 0;
    for (auto v : candidate_mask) candidate_states += (v != 0);
    std::cout << "\n[" << label << "]\n";
    std::cout << "states        : " << state_grid.total_size << '\n';
    std::cout << "inputs        : " << input_grid.total_size << '\n';
    std::cout << "pairs         : " << pair_count << '\n';
    std::cout << "reachable     : " << r.reachable_states << '\n';
    std::cout << "candidate     : " << candidate_states << '\n';
    std::cout << "candidate ctl : " << r.certified_candidate_states << '\n';
    std::cout << "iterations    : " << r.iterations << '\n';
    std::cout << "converged     : " << (r.converged ? "true" : "false") << '\n';
    std::cout << "message       : " << r.message << '\n';
    std::cout << "abstraction   : " << abstraction_ms << " ms\n";
    std::cout << "solve         : " << r.solve_ms << " ms\n";
    std::cout << "memory total  : " << gsc::format_bytes(budget.total_bytes) << '\n';
}

void print_summary(const gsc::CpuRunReport& cpu, const gsc::SolveResult<std::uint32_t>& r, const std::string& label) {
    print_summary(cpu.state_grid, cpu.input_grid, cpu.budget, cpu.pair_count, cpu.result.candidate_mask, r,
                  cpu.abstraction_ms, label);
}

bool same_result(const gsc::SolveResult<std::uint32_t>& a, const gsc::SolveResult<std::uint32_t>& b) {
    // main.cpp:114-121
    return a.valid_mask == b.valid_mask && a.candidate_mask == b.candidate_mask &&
           a.reachable_mask == b.reachable_mask && a.controller == b.controller && a.reach_step == b.reach_step;
}

void print_iteration(const gsc::IterationStats& s) {
    std::cout << "  Iter " << s.iteration << ": " << s.iteration_ms << " ms total"
              << " [prefix: " << s.prefix_build_ms << " ms"
              << ", satisfaction: " << s.satisfaction_check_ms << " ms"
              << ", reduction: " << s.reduction_ms << " ms]"
              << " -> " << s.newly_reachable << " new reachable, " << s.newly_certified << " new certified\n";
}

void print_gpu_detailed_timing(const gsc::GpuRunReport& gpu) {
    // main.cpp:41-103
    const auto& k = gpu.kernel_timings;
    std::cout << "\n[GPU Detailed Timing]\n";
    std::cout << "=== Abstraction Phase ===\n";
    std::cout << "  Kernel execution      : " << k.abstraction_kernel_ms << " ms\n";
    std::cout << "  H2D memcpy            : " << k.abstraction_memcpy_h2d_ms << " ms\n";
    std::cout << "  Total abstraction     : " << gpu.abstraction_ms << " ms\n";
    std::cout << "\n=== Reachability Iteration Phase ===\n";
    std::cout << "  Prefix build (total)  : " << k.prefix_build_total_ms << " ms\n";
    std::cout << "  Pair satisfaction     : " << k.pair_satisfaction_ms << " ms\n";
    std::cout << "  Reduce inputs         : " << k.reduce_inputs_ms << " ms\n";
    std::cout << "  Iteration memcpy      : " << k.iteration_memcpy_ms << " ms\n";
    std::cout << "  Total solve           : " << gpu.solve_ms << " ms\n";
    std::cout << "\n=== Result Copy Phase ===\n";
    std::cout << "  D2H memcpy            : " << k.result_memcpy_d2h_ms << " ms\n";
    std::cout << "\n=== Summary ===\n";
    std::cout << "  Total kernel time     : " << k.total_kernel_ms << " ms\n";
    std::cout << "  Total memcpy time     : " << k.total_memcpy_ms << " ms\n";
    std::cout << "  Total compute time    : " << k.total_compute_ms << " ms\n";
    std::cout << "  Total elapsed time    : " << gpu.total_ms << " ms\n";
    const auto& m = gpu.memory_stats;
    std::cout << "\n=== GPU Memory Usage ===\n";
    std::cout << "  Abstraction data      : " << gsc::format_bytes(m.abstraction_bytes) << "\n";
    std::cout << "  Prefix sum data       : " << gsc::format_bytes(m.prefix_bytes) << "\n";
    std::cout << "  Iteration data        : " << gsc::format_bytes(m.iteration_bytes) << "\n";
    std::cout << "  Total allocated       : " << gsc::format_bytes(m.total_allocated_bytes) << "\n";
    std::cout << "  Peak usage            : " << gsc::format_bytes(m.peak_usage_bytes) << "\n";
    const auto& st = gpu.iteration_stats;
    if (st.empty()) return;
    std::cout << "\n=== Per-Iteration Breakdown (first 5 and last 5) ===\n";
    const int show = std::min(5, static_cast<int>(st.size()));
    for (int i = 0; i < show; ++i) print_iteration(st[i]);
    if (static_cast<int>(st.size()) > 2 * show) std::cout << "  ...\n";
    for (int i = std::max(show, static_cast<int>(st.size()) - show); i < static_cast<int>(st.size()); ++i)
        print_iteration(st[i]);
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 2) {
        std::cerr << "usage: gsc_cli <config.cfg> [--cpu|--gpu|--compare] [--output result.h5]\n";
        return 1;
    }
    try {
        const std::string config_path = argv[1];
        std::string mode = "--cpu";
        std::string output_path;
        bool export_abstraction = true;
        for (int i = 2; i < argc; ++i) {
            const std::string arg = argv[i];
            if (arg == "--cpu" || arg == "--gpu" || arg == "--compare") {
                mode = arg;
                continue;
            }
            if (arg == "--output") {
                if (i + 1 >= argc) throw std::runtime_error("--output requires a file path");
                output_path = argv[++i];
                continue;
            }
            if (arg == "--no-abstraction-export") {
                export_abstraction = false;
                continue;
            }
            throw std::runtime_error("unknown argument: " + arg);
        }
        gsc::CaseExtensions ext;
        const auto cfg = gsc::load_case_config_ext(config_path, &ext);
        const std::string mode_name = mode.substr(2);
        if (output_path.empty()) output_path = gsc::default_result_path(cfg, mode_name);

        if (mode == "--cpu") {
            const auto cpu = gsc::run_case_cpu_u32(cfg, ext);
            print_summary(cpu, cpu.result, "cpu");
            gsc::write_results_hdf5(output_path, cfg, cpu.state_grid, cpu.input_grid, cpu.abstraction_ms,
                                    {{"cpu", &cpu.result}});
            std::cout << "result file   : " << output_path << '\n';
            return 0;
        }

        if (mode == "--gpu") {
            gsc::RunOptions run_opts;
            run_opts.export_abstraction = export_abstraction;
            const auto gpu = gsc::run_case_hip_u32(cfg, ext, run_opts);
            if (!gpu.executed) {
                std::cerr << "gpu run unavailable: " << gpu.message << '\n';
                return 2;
            }
            print_summary(gpu.state_grid, gpu.input_grid, gpu.budget, gpu.pair_count, gpu.result.candidate_mask,
                          gpu.result, gpu.abstraction_ms, "gpu");
            print_gpu_detailed_timing(gpu);
            gsc::write_gpu_results_hdf5(output_path, cfg, gpu);
            std::cout << "\nresult file   : " << output_path << '\n';
            std::cout << "  - includes abstraction data ("
                      << (gpu.pair_count * (2 * sizeof(std::uint32_t) + sizeof(std::uint8_t)) / (1024.0 * 1024.0))
                      << " MB)\n";
            std::cout << "  - includes " << gpu.iteration_stats.size() << " iteration statistics\n";
            return 0;
        }

        if (mode == "--compare") {
            const auto cpu = gsc::run_case_cpu_u32(cfg, ext);
            gsc::RunOptions run_opts;
            run_opts.export_abstraction = false;
            const auto gpu = gsc::run_case_hip_u32(cfg, ext, run_opts);
            if (!gpu.executed) {
                std::cerr << "gpu compare unavailable: " << gpu.message << '\n';
                print_summary(cpu, cpu.result, "cpu");
                gsc::write_results_hdf5(output_path, cfg, cpu.state_grid, cpu.input_grid, cpu.abstraction_ms,
                                        {{"cpu", &cpu.result}});
                std::cout << "result file   : " << output_path << '\n';
                return 2;
            }
            // main.cpp:200-211: both summaries from the CPU preparation.
            print_summary(cpu, cpu.result, "cpu");
            print_summary(cpu, gpu.result, "gpu");
            print_gpu_detailed_timing(gpu);
            gsc::write_results_hdf5(output_path, cfg, cpu.state_grid, cpu.input_grid, cpu.abstraction_ms,
                                    {{"cpu", &cpu.result}, {"gpu", &gpu.result}});
            const bool match = same_result(cpu.result, gpu.result);
            std::cout << "\nresult file   : " << output_path << '\n';
            std::cout << "exact_match   : " << (match ? "true" : "false") << '\n';
            return match ? 0 : 3;
        }

        throw std::runtime_error("unknown mode: " + mode);
    } catch (const std::exception& ex) {
        std::cerr << "error: " << ex.what() << '\n';
        return 1;
    }
}
