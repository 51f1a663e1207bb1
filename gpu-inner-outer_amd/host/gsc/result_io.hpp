// gsc/result_io.hpp — HDF5 result files with the reference's layouts:
//   write_gpu_results_hdf5 (src/result_io.cpp:364-517): /meta, /state, /input,
//     /config, /summary/gpu, /result/gpu, /abstraction, /iterations, /timings;
//   write_results_hdf5 (src/result_io.cpp:238-362), for --cpu and --compare:
//     /meta, /state, /input, /config, and /summary/<run>, /result/<run> per run.
#pragma once

#include <string>
#include <vector>

#include "gsc/cuda_api.hpp"

namespace gsc {

// One solve to write (result_io.hpp:18-21).
struct ResultExportView {
    std::string label;
    const SolveResult<std::uint32_t>* result = nullptr;
};

bool hdf5_export_compiled();
// results/<name>_<mode>.h5 with non [A-Za-z0-9_-] characters replaced by '_'
// (result_io.cpp:23-37, 330-335).
std::string default_result_path(const CaseConfig& cfg, const std::string& mode);
// Every run's /summary/<label>/abstraction_ms is `abstraction_ms`, the CPU
// preparation's, as in the reference.
void write_results_hdf5(const std::string& output_path, const CaseConfig& cfg, const Grid4D<std::uint32_t>& state_grid,
                        const InputGrid2D& input_grid, double abstraction_ms,
                        const std::vector<ResultExportView>& runs);
void write_gpu_results_hdf5(const std::string& output_path, const CaseConfig& cfg, const GpuRunReport& gpu_report);

}  // namespace gsc
