// gsc/hip_api.hpp — this backend's additions to the reference boundary. None
// has a reference counterpart; the reference's own entry points are in
// gsc/cuda_api.hpp.
#pragma once

#include <string>
#include <vector>

#include "gsc/cuda_api.hpp"

namespace gsc {

// The `.cfg` extension keys, kept outside CaseConfig so that CaseConfig keeps
// the reference's layout. Left empty they change nothing.
struct CaseExtensions {
    std::vector<HyperRect4D> extra_targets;  // `target = ...`: a centre in any of them is a candidate
    bool target_band_enabled = false;        // `target_band = c w`: candidate box cells need |x*y - c| <= w
    double target_band_c = 0.0;
    double target_band_w = 0.0;
};

struct RunOptions {
    // Copy the per-pair abstraction triplet to the host (9 B per pair, 3.6 GB
    // at 401M pairs). run_case_cuda_u32 always does, as the reference does.
    bool export_abstraction = true;
};

// run_case_cuda_u32 with the extensions and options above (exported by
// libgsc_hip.so next to the reference's two symbols).
GpuRunReport run_case_hip_u32(const CaseConfig& cfg, const CaseExtensions& ext, const RunOptions& opts);

// load_case_config that also returns the extension keys (CLI side, with the
// other config functions).
CaseConfig load_case_config_ext(const std::string& path, CaseExtensions* ext);

}  // namespace gsc
