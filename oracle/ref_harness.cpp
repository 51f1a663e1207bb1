// ref_harness.cpp — drives the reference's own CPU implementation, compiled
// from the reference sources (include/gsc/*.hpp header templates and
// src/config.cpp) by oracle/Makefile into oracle/_ref/. Test infrastructure
// only: it produces golden vectors and the reference CPU timing baseline.
//
//   ref_harness solve <cfg | builtin:NAME> <outdir> [--target 8 numbers]... [--band c w]
//       load_case_config (config.cpp:243) -> prepare_case_cpu (abstraction.hpp:276)
//       -> solve_backward_cpu(verbose) (reachability.hpp:31); writes raw arrays
//       and summary.txt into <outdir>; per-iteration lines go to stdout.
//   ref_harness sample <cfg> <state_begin> <state_end> <iters> [--target ..] [--band c w]
//       times the reference's per-pair work on states [begin, end):
//       successor_box + continuous_box_to_index_box + box_satisfies_constraint
//       (abstraction.hpp:172-205), and `iters` sweeps of the pair-satisfaction
//       test (two query_box_count calls per valid pair, reachability.hpp:78-93)
//       against prefix tables of the valid mask and the candidate (seed) mask.
//
// The extension options rebuild the candidate mask the same way the
// product's extension keys define it (centre in the candidate box and, with
// --band, |x*y - c| <= w; or centre in any --target box) using the reference's
// own Grid4D::center and point_in_rect, then re-check candidate ⊆ valid.

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "gsc/abstraction.hpp"
#include "gsc/config.hpp"
#include "gsc/prefix_sum.hpp"
#include "gsc/reachability.hpp"

namespace {

struct Extension {
    std::vector<gsc::HyperRect4D> targets;
    bool band = false;
    double band_c = 0.0, band_w = 0.0;
};

bool parse_extension(int argc, char** argv, int start, Extension* ext) {
    for (int i = start; i < argc; ++i) {
        const std::string a = argv[i];
        if (a == "--target" && i + 8 < argc) {
            gsc::HyperRect4D r;
            for (int k = 0; k < 4; ++k) r.lb[k] = std::strtod(argv[i + 1 + k], nullptr);
            for (int k = 0; k < 4; ++k) r.ub[k] = std::strtod(argv[i + 5 + k], nullptr);
            ext->targets.push_back(r);
            i += 8;
        } else if (a == "--band" && i + 2 < argc) {
            ext->band = true;
            ext->band_c = std::strtod(argv[i + 1], nullptr);
            ext->band_w = std::strtod(argv[i + 2], nullptr);
            i += 2;
        } else {
            std::cerr << "unknown option " << a << "\n";
            return false;
        }
    }
    return true;
}

bool has_extension(const Extension& ext) { return !ext.targets.empty() || ext.band; }

// With extension options the reference's own candidate test (prepare_case_*,
// abstraction.hpp:257-265) must not run on the unrestricted candidate box, so
// preparation gets a copy whose candidate box holds no centre; the candidate
// mask is then rebuilt here and checked against the valid mask.
gsc::CaseConfig preparation_config(const gsc::CaseConfig& cfg, const Extension& ext) {
    gsc::CaseConfig prep = cfg;
    if (has_extension(ext))
        for (int d = 0; d < 4; ++d) {
            prep.candidate.lb[d] = 1.0;
            prep.candidate.ub[d] = 0.0;
        }
    return prep;
}

void apply_extension(const gsc::CaseConfig& cfg, const Extension& ext, gsc::PreparedCase<std::uint32_t>* p) {
    if (!has_extension(ext)) return;
    const auto& g = p->state_grid;
    for (std::uint64_t s = 0; s < g.total_size; ++s) {
        double x[4];
        g.center(static_cast<std::uint32_t>(s), x);
        bool in = gsc::point_in_rect(cfg.candidate, x);
        if (in && ext.band) in = std::fabs(x[0] * x[1] - ext.band_c) <= ext.band_w;
        if (!in)
            for (const auto& r : ext.targets)
                if (gsc::point_in_rect(r, x)) {
                    in = true;
                    break;
                }
        p->candidate_mask[s] = in ? 1 : 0;
        if (in && !p->valid_mask[s]) throw std::runtime_error("candidate set contains invalid states");
    }
}

template <typename T>
void dump(const std::string& path, const std::vector<T>& v) {
    std::ofstream f(path, std::ios::binary);
    f.write(reinterpret_cast<const char*>(v.data()), static_cast<std::streamsize>(v.size() * sizeof(T)));
}

gsc::CaseConfig builtin_case(const std::string& which) {
    if (which == "default") return gsc::make_default_case();
    if (which == "wraparound") return gsc::make_wraparound_case();
    if (which == "smoke") return gsc::make_smoke_case();
    if (which == "hyperbolic") return gsc::make_hyperbolic_case();
    throw std::runtime_error("unknown builtin case " + which);
}

int run_solve(int argc, char** argv) {
    if (argc < 4) return 1;
    Extension ext;
    if (!parse_extension(argc, argv, 4, &ext)) return 1;
    // `solve builtin:<name> <outdir>` runs one of the reference's canned cases
    // (config.cpp:129-241) instead of reading a file.
    const std::string src = argv[2];
    auto cfg = src.rfind("builtin:", 0) == 0 ? builtin_case(src.substr(8)) : gsc::load_case_config(src);
    // REF_MAX_ITERATIONS overrides max_iterations (to shorten long canned cases).
    if (const char* mi = std::getenv("REF_MAX_ITERATIONS")) cfg.max_iterations = static_cast<std::uint32_t>(std::atoi(mi));
    const std::string out = argv[3];
    auto prepared = gsc::prepare_case_cpu<std::uint32_t>(preparation_config(cfg, ext));
    apply_extension(cfg, ext, &prepared);
    std::cout << std::flush;
    // verbose=true: the per-iteration lines on stdout carry the per-iteration counts.
    const auto result = gsc::solve_backward_cpu(prepared, cfg.max_iterations, true);
    std::fflush(stdout);
    dump(out + "/valid_mask.bin", result.valid_mask);
    dump(out + "/candidate_mask.bin", result.candidate_mask);
    dump(out + "/reachable_mask.bin", result.reachable_mask);
    dump(out + "/controller.bin", result.controller);
    dump(out + "/reach_step.bin", result.reach_step);
    dump(out + "/min_flat.bin", prepared.abstraction.min_flat);
    dump(out + "/max_flat.bin", prepared.abstraction.max_flat);
    dump(out + "/pair_valid.bin", prepared.abstraction.valid);
    std::ofstream s(out + "/summary.txt");
    s << "iterations " << result.iterations << "\n";
    s << "converged " << (result.converged ? 1 : 0) << "\n";
    s << "reachable_states " << result.reachable_states << "\n";
    s << "certified_candidate_states " << result.certified_candidate_states << "\n";
    s << "abstraction_ms " << prepared.abstraction_ms << "\n";
    s << "solve_ms " << result.solve_ms << "\n";
    s << "message " << result.message << "\n";
    return 0;
}

int run_sample(int argc, char** argv) {
    if (argc < 6) return 1;
    Extension ext;
    if (!parse_extension(argc, argv, 6, &ext)) return 1;
    const auto cfg = gsc::load_case_config(argv[2]);
    std::uint64_t begin = std::strtoull(argv[3], nullptr, 10);
    std::uint64_t end = std::strtoull(argv[4], nullptr, 10);
    const int iters = std::atoi(argv[5]);
    auto prepared = gsc::prepare_case_gpu_minimal<std::uint32_t>(preparation_config(cfg, ext));
    apply_extension(cfg, ext, &prepared);
    const auto& g = prepared.state_grid;
    const auto& in = prepared.input_grid;
    if (end > g.total_size) end = g.total_size;
    const std::uint64_t I = in.total_size;
    const std::uint64_t n = (end - begin) * I;
    std::vector<std::uint32_t> mnf(n, 0), mxf(n, 0);
    std::vector<std::uint8_t> val(n, 0);

    // Abstraction of the sample, the loop body of build_abstraction_cpu
    // (abstraction.hpp:172-205) over the sampled states.
    const auto t0 = std::chrono::steady_clock::now();
    for (std::uint64_t s = begin; s < end; ++s) {
        double x[4];
        g.center(static_cast<std::uint32_t>(s), x);
        if (cfg.constraint_type != gsc::ConstraintType::kNone && !gsc::satisfies_state_constraint(cfg, x)) continue;
        for (std::uint64_t i = 0; i < I; ++i) {
            double u[2], mn[4], mx[4];
            in.center(static_cast<gsc::InputIndex>(i), u);
            prepared.model.successor_box(x, u, g.eta, mn, mx);
            gsc::IndexBox4D box{};
            if (!gsc::continuous_box_to_index_box(g, mn, mx, &box)) continue;
            if (!gsc::box_satisfies_constraint(g, box, cfg)) continue;
            const std::uint64_t p = (s - begin) * I + i;
            mnf[p] = g.flatten(box.min);
            mxf[p] = g.flatten(box.max);
            val[p] = 1;
        }
    }
    const auto t1 = std::chrono::steady_clock::now();

    // Pair-satisfaction sweeps (the reference kernel's schedule: every pair).
    const auto pv = gsc::build_prefix_sum_cpu(g, prepared.valid_mask);
    const auto pr = gsc::build_prefix_sum_cpu(g, prepared.candidate_mask);
    std::uint64_t satisfied = 0;
    const auto t2 = std::chrono::steady_clock::now();
    for (int it = 0; it < iters; ++it) {
        for (std::uint64_t p = 0; p < n; ++p) {
            if (!val[p]) continue;
            const auto vc = gsc::query_box_count(g, pv, mnf[p], mxf[p]);
            if (vc == 0) continue;
            const auto rc = gsc::query_box_count(g, pr, mnf[p], mxf[p]);
            satisfied += (rc == vc) ? 1 : 0;
        }
    }
    const auto t3 = std::chrono::steady_clock::now();
    const double abs_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
    const double sat_ms = std::chrono::duration<double, std::milli>(t3 - t2).count();
    std::printf("sample_states %llu\n", static_cast<unsigned long long>(end - begin));
    std::printf("sample_pairs %llu\n", static_cast<unsigned long long>(n));
    std::printf("abstraction_ms %.6f\n", abs_ms);
    std::printf("satisfaction_sweeps %d\n", iters);
    std::printf("satisfaction_ms %.6f\n", sat_ms);
    std::printf("satisfied %llu\n", static_cast<unsigned long long>(satisfied));
    return 0;
}

}  // namespace

int main(int argc, char** argv) {
    try {
        if (argc >= 2 && std::string(argv[1]) == "solve") return run_solve(argc, argv);
        if (argc >= 2 && std::string(argv[1]) == "sample") return run_sample(argc, argv);
        std::cerr << "usage: ref_harness solve <cfg> <outdir> | sample <cfg> <begin> <end> <iters>\n";
        return 1;
    } catch (const std::exception& ex) {
        std::cerr << "error: " << ex.what() << "\n";
        return 1;
    }
}
