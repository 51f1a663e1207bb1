// gsc/config.hpp — case configuration with the reference's public interface
// (include/gsc/config.hpp:12-82): the same members in the same order with the
// same defaults, so CaseConfig has the reference's layout and a caller
// compiled against either header passes the same object to
// gsc::run_case_cuda_u32. Parsing and the canned cases are served by the C ABI
// (gsc_config_load / gsc_config_builtin in libgsc_amd.so). This backend's
// extension keys live in gsc::CaseExtensions (gsc/hip_api.hpp), outside
// CaseConfig.
#pragma once

#include <cstdint>
#include <functional>
#include <string>
#include <vector>

namespace gsc {

constexpr int kStateDim = 4;
constexpr int kInputDim = 2;

using PairIndex = std::uint64_t;
using PrefixCount = std::uint64_t;
using InputIndex = std::uint32_t;
using ReachStep = std::uint32_t;

constexpr InputIndex kInvalidInput = 0xFFFFFFFFu;
constexpr ReachStep kUnreachableStep = 0xFFFFFFFFu;

struct HyperRect4D {
    double lb[kStateDim]{};
    double ub[kStateDim]{};
};

enum class ConstraintType { kNone, kHyperbolic, kElliptic, kCustom };

struct HyperbolicConstraintParams {
    double a = 4.0;
    double b = 4.0;
    double c = 16.0;
};

struct EllipticConstraintParams {
    double a = 2.0;
    double b = 2.0;
};

using StateConstraintFn = std::function<bool(const double x[kStateDim])>;

// CaseConfig with the reference's members and defaults (config.hpp:41-67).
struct CaseConfig {
    std::string name = "unicycle_default";
    double state_lb[kStateDim]{-2.0, -2.0, -3.141592653589793, 0.0};
    double state_ub[kStateDim]{2.0, 2.0, 3.141592653589793, 2.0};
    double state_eta[kStateDim]{0.5, 0.5, 0.7853981633974483, 0.5};
    int wrap_dim = 2;

    double input_lb[kInputDim]{-0.5, -1.0};
    double input_ub[kInputDim]{0.5, 1.0};
    double input_eta[kInputDim]{0.5, 1.0};

    HyperRect4D candidate{};
    HyperRect4D map{};
    std::vector<HyperRect4D> obstacles;

    ConstraintType constraint_type = ConstraintType::kNone;
    HyperbolicConstraintParams hyperbolic_params;
    EllipticConstraintParams elliptic_params;
    // kCustom: the valid mask takes it at every cell centre; the successor-box
    // scan does not (the reference's GPU kernel treats type 3 as no constraint,
    // kernels.cu:55-73, 79).
    StateConstraintFn custom_constraint;

    double disturbance_half_width[kStateDim]{0.05, 0.05, 0.02, 0.02};
    double sample_time = 0.2;
    std::uint32_t integration_substeps = 32;
    std::uint32_t max_iterations = 32;
    bool verbose = true;
};

// config.cpp:88-109: nullptr for kNone, the parameterised test otherwise.
StateConstraintFn make_state_constraint(const CaseConfig& cfg);
bool point_in_rect(const HyperRect4D& rect, const double x[kStateDim]);
bool satisfies_state_constraint(const CaseConfig& cfg, const double x[kStateDim]);
CaseConfig make_default_case();
CaseConfig make_wraparound_case();
CaseConfig make_smoke_case();
CaseConfig make_hyperbolic_case();
// Throws std::runtime_error with the reference's messages (config.cpp:243-347).
CaseConfig load_case_config(const std::string& path);

}  // namespace gsc
