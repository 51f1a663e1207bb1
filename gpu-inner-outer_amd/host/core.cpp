// core.cpp — the reference's config functions (config.hpp:69-80) for this
// repository's own CLI, served by the C ABI's parser (gsc_config_load /
// gsc_config_builtin in libgsc_amd.so). They are not part of libgsc_hip.so: a
// program built against the reference links the reference's own config.cpp
// next to that library, so the backend library exports nothing it defines.
#include <stdexcept>
#include <string>

#include "gsc/hip_api.hpp"
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

CaseConfig from_desc(const gsc_case_desc& d, CaseExtensions* ext) {
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
    if (ext) {
        *ext = CaseExtensions{};
        for (uint32_t k = 0; k < d.extra_target_count; ++k) ext->extra_targets.push_back(to_rect(d.extra_targets[k]));
        ext->target_band_enabled = d.target_band_enabled != 0;
        ext->target_band_c = d.target_band_c;
        ext->target_band_w = d.target_band_w;
    }
    return c;
}

CaseConfig builtin(int which) {
    gsc_config* h = nullptr;
    if (gsc_config_builtin(which, &h) != GSC_OK) throw std::runtime_error("unknown canned case");
    CaseConfig c = from_desc(*gsc_config_desc(h), nullptr);
    gsc_config_free(h);
    return c;
}

bool hyperbolic_ok(const HyperbolicConstraintParams& p, const double x[kStateDim]) {
    // config.cpp:64-68
    const double x1_sq = x[0] * x[0];
    const double x2_sq = x[1] * x[1];
    return (x1_sq - x2_sq <= p.a) && (p.b * x2_sq - x1_sq <= p.c);
}

bool elliptic_ok(const EllipticConstraintParams& p, const double x[kStateDim]) {
    // config.cpp:71-75
    const double t1 = (x[0] / p.a) * (x[0] / p.a);
    const double t2 = (x[1] / p.b) * (x[1] / p.b);
    return t1 + t2 <= 1.0;
}

}  // namespace

bool point_in_rect(const HyperRect4D& rect, const double x[kStateDim]) {
    // config.cpp:79-86
    for (int d = 0; d < kStateDim; ++d)
        if (x[d] < rect.lb[d] || x[d] > rect.ub[d]) return false;
    return true;
}

StateConstraintFn make_state_constraint(const CaseConfig& cfg) {
    // config.cpp:88-109
    switch (cfg.constraint_type) {
        case ConstraintType::kHyperbolic: {
            const auto p = cfg.hyperbolic_params;
            return [p](const double x[kStateDim]) { return hyperbolic_ok(p, x); };
        }
        case ConstraintType::kElliptic: {
            const auto p = cfg.elliptic_params;
            return [p](const double x[kStateDim]) { return elliptic_ok(p, x); };
        }
        case ConstraintType::kCustom:
            return cfg.custom_constraint;
        default:
            return nullptr;
    }
}

bool satisfies_state_constraint(const CaseConfig& cfg, const double x[kStateDim]) {
    // config.cpp:111-127
    switch (cfg.constraint_type) {
        case ConstraintType::kHyperbolic:
            return hyperbolic_ok(cfg.hyperbolic_params, x);
        case ConstraintType::kElliptic:
            return elliptic_ok(cfg.elliptic_params, x);
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

CaseConfig load_case_config_ext(const std::string& path, CaseExtensions* ext) {
    gsc_config* h = nullptr;
    char err[512] = {0};
    if (gsc_config_load(path.c_str(), &h, err, sizeof(err)) != GSC_OK) throw std::runtime_error(err);
    CaseConfig c = from_desc(*gsc_config_desc(h), ext);
    gsc_config_free(h);
    return c;
}

CaseConfig load_case_config(const std::string& path) { return load_case_config_ext(path, nullptr); }

}  // namespace gsc
