// backend.cpp — libgsc_hip.so: the reference's GPU backend symbols
// (cuda_api.hpp:87-88) with their exact signatures, implemented over the C ABI
// of libgsc_amd.so. It takes the place of the reference's CUDA library
// (CMakeLists.txt:24-37, cuda/kernels.cu): a program compiled against the
// reference's headers links it next to the reference's own config/result code.
// Built with hidden visibility; only the GSC_EXPORT functions below leave it.
#include <algorithm>
#include <cstdint>
#include <stdexcept>

#include "gsc/hip_api.hpp"
#include "gsc_amd.h"

#define GSC_EXPORT __attribute__((visibility("default")))

namespace gsc {

namespace {

gsc_rect4 from_rect(const HyperRect4D& r) {
    gsc_rect4 out;
    for (int d = 0; d < kStateDim; ++d) {
        out.lb[d] = r.lb[d];
        out.ub[d] = r.ub[d];
    }
    return out;
}

// The descriptor borrows from `cfg`, `ext` and the two vectors; all must
// outlive its use.
gsc_case_desc to_desc(const CaseConfig& cfg, const CaseExtensions& ext, std::vector<gsc_rect4>* obstacles,
                      std::vector<gsc_rect4>* targets) {
    gsc_case_desc d{};
    d.name = cfg.name.c_str();
    for (int k = 0; k < kStateDim; ++k) {
        d.state_lb[k] = cfg.state_lb[k];
        d.state_ub[k] = cfg.state_ub[k];
        d.state_eta[k] = cfg.state_eta[k];
        d.disturbance_half_width[k] = cfg.disturbance_half_width[k];
    }
    d.wrap_dim = cfg.wrap_dim;
    for (int k = 0; k < kInputDim; ++k) {
        d.input_lb[k] = cfg.input_lb[k];
        d.input_ub[k] = cfg.input_ub[k];
        d.input_eta[k] = cfg.input_eta[k];
    }
    d.candidate = from_rect(cfg.candidate);
    d.map = from_rect(cfg.map);
    for (const auto& r : cfg.obstacles) obstacles->push_back(from_rect(r));
    for (const auto& r : ext.extra_targets) targets->push_back(from_rect(r));
    d.obstacles = obstacles->empty() ? nullptr : obstacles->data();
    d.obstacle_count = static_cast<uint32_t>(obstacles->size());
    d.extra_targets = targets->empty() ? nullptr : targets->data();
    d.extra_target_count = static_cast<uint32_t>(targets->size());
    // kCustom crosses as the state_allow mask (run_case_hip_u32); the reference's
    // GPU kernel evaluates only none/hyperbolic/elliptic (kernels.cu:55-73).
    d.constraint_type = cfg.constraint_type == ConstraintType::kHyperbolic ? GSC_CONSTRAINT_HYPERBOLIC
                        : cfg.constraint_type == ConstraintType::kElliptic ? GSC_CONSTRAINT_ELLIPTIC
                                                                           : GSC_CONSTRAINT_NONE;
    d.hyperbolic_a = cfg.hyperbolic_params.a;
    d.hyperbolic_b = cfg.hyperbolic_params.b;
    d.hyperbolic_c = cfg.hyperbolic_params.c;
    d.elliptic_a = cfg.elliptic_params.a;
    d.elliptic_b = cfg.elliptic_params.b;
    d.sample_time = cfg.sample_time;
    d.integration_substeps = cfg.integration_substeps;
    d.max_iterations = cfg.max_iterations;
    d.verbose = cfg.verbose ? 1 : 0;
    d.target_band_enabled = ext.target_band_enabled ? 1 : 0;
    d.target_band_c = ext.target_band_c;
    d.target_band_w = ext.target_band_w;
    return d;
}

// kCustom: the reference's mask build evaluates the std::function at every cell
// centre (abstraction.hpp:97-101, centre as grid.hpp:38-44); the result
// crosses the C ABI as a per-state byte mask.
std::vector<std::uint8_t> custom_allow_mask(const CaseConfig& cfg, const gsc_case_shape& sh) {
    std::vector<std::uint8_t> allow(sh.state_count, 1);
    if (!cfg.custom_constraint) return allow;
    std::uint64_t s = 0;
    for (std::uint32_t i3 = 0; i3 < sh.state_shape[3]; ++i3)
        for (std::uint32_t i2 = 0; i2 < sh.state_shape[2]; ++i2)
            for (std::uint32_t i1 = 0; i1 < sh.state_shape[1]; ++i1)
                for (std::uint32_t i0 = 0; i0 < sh.state_shape[0]; ++i0, ++s) {
                    const std::uint32_t idx[4] = {i0, i1, i2, i3};
                    double x[kStateDim];
                    for (int d = 0; d < kStateDim; ++d)
                        x[d] = cfg.state_lb[d] + (static_cast<double>(idx[d]) + 0.5) * cfg.state_eta[d];
                    allow[s] = cfg.custom_constraint(x) ? 1 : 0;
                }
    return allow;
}

}  // namespace

GSC_EXPORT bool cuda_backend_compiled() { return gsc_backend_compiled() != 0; }

GSC_EXPORT GpuRunReport run_case_cuda_u32(const CaseConfig& cfg) {
    return run_case_hip_u32(cfg, CaseExtensions{}, RunOptions{});
}

GSC_EXPORT GpuRunReport run_case_hip_u32(const CaseConfig& cfg, const CaseExtensions& ext, const RunOptions& opts) {
    GpuRunReport report;
    report.compiled_with_cuda = true;
    const bool with_abstraction = opts.export_abstraction;
    std::vector<gsc_rect4> obstacles, targets;
    gsc_case_desc desc = to_desc(cfg, ext, &obstacles, &targets);

    gsc_case_shape shape{};
    char err[512] = {0};
    if (gsc_case_shape_of(&desc, &shape, err, sizeof(err)) != GSC_OK) throw std::runtime_error(err);
    const std::uint64_t S = shape.state_count;
    const std::uint64_t P = shape.pair_count;
    std::vector<std::uint8_t> allow;
    if (cfg.constraint_type == ConstraintType::kCustom && cfg.custom_constraint) {
        allow = custom_allow_mask(cfg, shape);
        desc.state_allow = allow.data();
    }

    auto& res = report.result;
    res.valid_mask.assign(S, 0);
    res.candidate_mask.assign(S, 0);
    res.reachable_mask.assign(S, 0);
    res.controller.assign(S, 0);
    res.reach_step.assign(S, 0);
    if (with_abstraction) {
        report.abstraction_min_flat.assign(P, 0);
        report.abstraction_max_flat.assign(P, 0);
        report.abstraction_valid.assign(P, 0);
    }
    std::vector<gsc_iteration_stats> stats(std::max<uint32_t>(1, cfg.max_iterations));

    gsc_result_buffers out{};
    out.valid_mask = res.valid_mask.data();
    out.candidate_mask = res.candidate_mask.data();
    out.reachable_mask = res.reachable_mask.data();
    out.controller = res.controller.data();
    out.reach_step = res.reach_step.data();
    if (with_abstraction) {
        out.abstraction_min_flat = report.abstraction_min_flat.data();
        out.abstraction_max_flat = report.abstraction_max_flat.data();
        out.abstraction_valid = report.abstraction_valid.data();
    }
    out.iteration_stats = stats.data();
    out.iteration_stats_capacity = static_cast<uint32_t>(stats.size());

    gsc_run_report rep{};
    const int rc = gsc_run_case_u32(&desc, &out, &rep);
    if (rc == GSC_ERR_CONFIG || rc == GSC_ERR_INVALID_ARGUMENT) throw std::runtime_error(rep.message);
    report.runtime_available = rep.runtime_available != 0;
    report.executed = rep.executed != 0;
    report.message = rep.message;
    if (rc != GSC_OK) {
        res = SolveResult<std::uint32_t>{};
        report.abstraction_min_flat.clear();
        report.abstraction_max_flat.clear();
        report.abstraction_valid.clear();
        return report;
    }

    res.iterations = rep.iterations;
    res.converged = rep.converged != 0;
    res.message = rep.result_message;
    res.reachable_states = rep.reachable_states;
    res.certified_candidate_states = rep.certified_candidate_states;
    res.solve_ms = rep.solve_ms;
    report.abstraction_ms = rep.abstraction_ms;
    report.solve_ms = rep.solve_ms;
    report.total_ms = rep.total_ms;
    report.pair_count = rep.pair_count;

    auto& g = report.state_grid;
    for (int d = 0; d < kStateDim; ++d) {
        g.lb[d] = cfg.state_lb[d];
        g.ub[d] = cfg.state_ub[d];
        g.eta[d] = cfg.state_eta[d];
        g.shape[d] = shape.state_shape[d];
        g.strides[d] = shape.state_strides[d];
    }
    g.total_size = shape.state_count;
    g.wrap_dim = cfg.wrap_dim;
    auto& ig = report.input_grid;
    for (int d = 0; d < kInputDim; ++d) {
        ig.lb[d] = cfg.input_lb[d];
        ig.ub[d] = cfg.input_ub[d];
        ig.eta[d] = cfg.input_eta[d];
        ig.shape[d] = shape.input_shape[d];
        ig.strides[d] = shape.input_strides[d];
    }
    ig.total_size = shape.input_count;

    const gsc_memory_budget& b = rep.budget;
    report.budget = {b.pair_min_bytes, b.pair_max_bytes, b.pair_valid_bytes, b.pair_satisfied_bytes,
                     b.reachable_prev_bytes, b.reachable_next_bytes, b.valid_mask_bytes, b.candidate_mask_bytes,
                     b.controller_bytes, b.reach_step_bytes, b.prefix_reachable_bytes, b.prefix_valid_bytes,
                     b.total_bytes};
    const gsc_kernel_timings& t = rep.kernel_timings;
    report.kernel_timings = {t.abstraction_kernel_ms, t.abstraction_memcpy_h2d_ms, t.scatter_mask_ms,
                             t.scan_axis0_ms, t.scan_axis1_ms, t.scan_axis2_ms, t.scan_axis3_ms,
                             t.prefix_build_total_ms, t.pair_satisfaction_ms, t.reduce_inputs_ms,
                             t.iteration_memcpy_ms, t.result_memcpy_d2h_ms, t.total_kernel_ms, t.total_memcpy_ms,
                             t.total_compute_ms};
    const gsc_memory_stats& m = rep.memory_stats;
    report.memory_stats = {m.total_allocated_bytes, m.peak_usage_bytes, m.abstraction_bytes,
                           m.prefix_bytes,          m.iteration_bytes,  m.result_bytes};
    for (uint32_t k = 0; k < rep.iteration_stats_written; ++k) {
        const gsc_iteration_stats& s = stats[k];
        report.iteration_stats.push_back({s.iteration, s.newly_reachable, s.newly_certified, s.total_reachable,
                                          s.total_certified, s.iteration_ms, s.prefix_build_ms,
                                          s.satisfaction_check_ms, s.reduction_ms});
    }
    return report;
}

}  // namespace gsc
