// backend.cpp — the reference's C++ entry points (config.hpp:69-80,
// cuda_api.hpp:87-88) implemented over the C ABI of libgsc_amd.so.
#include <cmath>
#include <cstdio>
#include <iomanip>
#include <sstream>
#include <stdexcept>

#include "gsc/cuda_api.hpp"
#include "gsc_amd.h"

namespace gsc {

namespace {

HyperRect4D to_rect(const gsc_rect4& r) {
    HyperRect4D out;
    for (int d = 0; d < kStateDim; ++d) {
        out.lb[d] = r.lb[d];
        out.ub[d] = r.ub[d];
    }
    return out;
}

gsc_rect4 from_rect(const HyperRect4D& r) {
    gsc_rect4 out;
    for (int d = 0; d < kStateDim; ++d) {
        out.lb[d] = r.lb[d];
        out.ub[d] = r.ub[d];
    }
    return out;
}

CaseConfig from_desc(const gsc_case_desc& d) {
    CaseConfig c;
    c.name = d.name ? d.name : "";
    for (int k = 0; k < kStateDim; ++k) {
        c.state_lb[k] = d.state_lb[k];
        c.state_ub[k] = d.state_ub[k];
        c.state_eta[k] = d.state_eta[k];
        c.disturbance_half_width[k] = d.disturbance_half_width[k];
    }
    c.wrap_dim = d.wrap_dim;
    for (int k = 0; k < kInputDim; ++k) {
        c.input_lb[k] = d.input_lb[k];
        c.input_ub[k] = d.input_ub[k];
        c.input_eta[k] = d.input_eta[k];
    }
    c.candidate = to_rect(d.candidate);
    c.map = to_rect(d.map);
    for (uint32_t k = 0; k < d.obstacle_count; ++k) c.obstacles.push_back(to_rect(d.obstacles[k]));
    c.constraint_type = d.constraint_type == GSC_CONSTRAINT_HYPERBOLIC ? ConstraintType::kHyperbolic
                        : d.constraint_type == GSC_CONSTRAINT_ELLIPTIC ? ConstraintType::kElliptic
                                                                       : ConstraintType::kNone;
    c.hyperbolic_params = {d.hyperbolic_a, d.hyperbolic_b, d.hyperbolic_c};
    c.elliptic_params = {d.elliptic_a, d.elliptic_b};
    c.sample_time = d.sample_time;
    c.integration_substeps = d.integration_substeps;
    c.max_iterations = d.max_iterations;
    c.verbose = d.verbose != 0;
    for (uint32_t k = 0; k < d.extra_target_count; ++k) c.extra_targets.push_back(to_rect(d.extra_targets[k]));
    c.target_band_enabled = d.target_band_enabled != 0;
    c.target_band_c = d.target_band_c;
    c.target_band_w = d.target_band_w;
    return c;
}

// The descriptor borrows from `cfg` and `rects`; both must outlive its use.
gsc_case_desc to_desc(const CaseConfig& cfg, std::vector<gsc_rect4>* obstacles, std::vector<gsc_rect4>* targets) {
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
    for (const auto& r : cfg.extra_targets) targets->push_back(from_rect(r));
    d.obstacles = obstacles->empty() ? nullptr : obstacles->data();
    d.obstacle_count = static_cast<uint32_t>(obstacles->size());
    d.extra_targets = targets->empty() ? nullptr : targets->data();
    d.extra_target_count = static_cast<uint32_t>(targets->size());
    // kCustom constraints cannot cross the C ABI; the reference's GPU kernel also
    // evaluates only none/hyperbolic/elliptic (kernels.cu:55-73).
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
    d.target_band_enabled = cfg.target_band_enabled ? 1 : 0;
    d.target_band_c = cfg.target_band_c;
    d.target_band_w = cfg.target_band_w;
    return d;
}

CaseConfig builtin(int which) {
    gsc_config* h = nullptr;
    if (gsc_config_builtin(which, &h) != GSC_OK) throw std::runtime_error("unknown canned case");
    CaseConfig c = from_desc(*gsc_config_desc(h));
    gsc_config_free(h);
    return c;
}

}  // namespace

bool point_in_rect(const HyperRect4D& rect, const double x[kStateDim]) {
    // config.cpp:79-86
    for (int d = 0; d < kStateDim; ++d)
        if (x[d] < rect.lb[d] || x[d] > rect.ub[d]) return false;
    return true;
}

bool satisfies_state_constraint(const CaseConfig& cfg, const double x[kStateDim]) {
    // config.cpp:64-75, 111-127
    switch (cfg.constraint_type) {
        case ConstraintType::kHyperbolic: {
            const double x1_sq = x[0] * x[0];
            const double x2_sq = x[1] * x[1];
            return (x1_sq - x2_sq <= cfg.hyperbolic_params.a) &&
                   (cfg.hyperbolic_params.b * x2_sq - x1_sq <= cfg.hyperbolic_params.c);
        }
        case ConstraintType::kElliptic: {
            const double t1 = (x[0] / cfg.elliptic_params.a) * (x[0] / cfg.elliptic_params.a);
            const double t2 = (x[1] / cfg.elliptic_params.b) * (x[1] / cfg.elliptic_params.b);
            return t1 + t2 <= 1.0;
        }
        case ConstraintType::kCustom:
            return cfg.custom_constraint ? cfg.custom_constraint(x) : true;
        default:
            return true;
    }
}

CaseConfig make_default_case() { return builtin(0); }
CaseConfig make_wraparound_case() { return builtin(1); }
CaseConfig make_smoke_case() { return builtin(2); }
CaseConfig make_hyperbolic_case() { return builtin(3); }

CaseConfig load_case_config(const std::string& path) {
    gsc_config* h = nullptr;
    char err[512] = {0};
    if (gsc_config_load(path.c_str(), &h, err, sizeof(err)) != GSC_OK) throw std::runtime_error(err);
    CaseConfig c = from_desc(*gsc_config_desc(h));
    gsc_config_free(h);
    return c;
}

std::string format_bytes(std::uint64_t bytes) {
    // abstraction.hpp:52-63
    static const char* kUnits[] = {"B", "KiB", "MiB", "GiB", "TiB"};
    double value = static_cast<double>(bytes);
    int unit = 0;
    while (value >= 1024.0 && unit < 4) {
        value /= 1024.0;
        ++unit;
    }
    std::ostringstream oss;
    oss << std::fixed << std::setprecision(2) << value << ' ' << kUnits[unit];
    return oss.str();
}

bool cuda_backend_compiled() { return gsc_backend_compiled() != 0; }

GpuRunReport run_case_cuda_u32(const CaseConfig& cfg, bool with_abstraction) {
    GpuRunReport report;
    report.compiled_with_cuda = true;
    std::vector<gsc_rect4> obstacles, targets;
    const gsc_case_desc desc = to_desc(cfg, &obstacles, &targets);

    gsc_case_shape shape{};
    char err[512] = {0};
    if (gsc_case_shape_of(&desc, &shape, err, sizeof(err)) != GSC_OK) throw std::runtime_error(err);
    const std::uint64_t S = shape.state_count;
    const std::uint64_t P = shape.pair_count;

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
