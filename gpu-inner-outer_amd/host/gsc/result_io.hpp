// gsc/result_io.hpp — HDF5 result file of a GPU run, with the reference's
// layout (src/result_io.cpp:364-517): /meta, /state, /input, /config,
// /summary/gpu, /result/gpu, /abstraction, /iterations, /timings.
#pragma once

#include <string>

#include "gsc/cuda_api.hpp"

namespace gsc {

bool hdf5_export_compiled();
// results/<name>_<mode>.h5 with non [A-Za-z0-9_-] characters replaced by '_'
// (result_io.cpp:23-37, 330-335).
std::string default_result_path(const CaseConfig& cfg, const std::string& mode);
void write_gpu_results_hdf5(const std::string& output_path, const CaseConfig& cfg, const GpuRunReport& gpu_report);

}  // namespace gsc
