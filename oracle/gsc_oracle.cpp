// gsc_oracle.cpp — CPU restatement of the reference's grid symbolic-control
// path, used ONLY as test infrastructure: the parity tests, smoke() and the
// bench's cpu_baseline leg call it as the checker. Nothing in the product
// links or calls it.
//
// It follows the reference CPU path line by line in FP64 with no FMA
// contraction (build with -O2 -ffp-contract=off, never -march=native):
//   grids          include/gsc/grid.hpp:103-168, 253-269
//   cell centres   include/gsc/grid.hpp:38-44, 68-78
//   box rounding   include/gsc/grid.hpp:181-251
//   model          include/gsc/model.hpp:14-112
//   masks          include/gsc/abstraction.hpp:65-106, config.cpp:64-86, 111-127
//   abstraction    include/gsc/abstraction.hpp:110-209
//   prefix / query include/gsc/prefix_sum.hpp:15-105
//   fixpoint       include/gsc/reachability.hpp:30-155
// Each pair and each state update is independent (the fixpoint is a Jacobi
// iteration over a prefix snapshot), so loops over states may be split across
// threads without changing any result.
//
// Pinning: see tests/test_oracle.py — checked against the reference CPU build
// (oracle/_ref), the reference's published run digests (tests/golden/), and
// closed forms (brute-force box counts, ceil-shape table).

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "gsc_amd.h"

namespace {

constexpr int kD = 4;

struct Grid {
    double lb[kD], ub[kD], eta[kD];
    uint32_t shape[kD];
    uint64_t strides[kD];
    uint64_t total;
    int wrap_dim;

    uint32_t flatten(const uint32_t idx[kD]) const {
        // grid.hpp:22-28: u64 arithmetic, cast to the u32 state index.
        return static_cast<uint32_t>(static_cast<uint64_t>(idx[0]) * strides[0] +
                                     static_cast<uint64_t>(idx[1]) * strides[1] +
                                     static_cast<uint64_t>(idx[2]) * strides[2] +
                                     static_cast<uint64_t>(idx[3]) * strides[3]);
    }
    void unflatten(uint32_t flat, uint32_t idx[kD]) const {
        // grid.hpp:30-36
        uint64_t v = flat;
        for (int d = kD - 1; d >= 0; --d) {
            idx[d] = static_cast<uint32_t>(v / strides[d]);
            v %= strides[d];
        }
    }
    void center(uint32_t flat, double out[kD]) const {
        // grid.hpp:38-44
        uint32_t idx[kD];
        unflatten(flat, idx);
        for (int d = 0; d < kD; ++d) out[d] = lb[d] + (static_cast<double>(idx[d]) + 0.5) * eta[d];
    }
    double wrap_value(double x) const {
        // grid.hpp:46-57
        const double low = lb[wrap_dim];
        const double high = ub[wrap_dim];
        const double span = high - low;
        while (x < low) x += span;
        while (x >= high) x -= span;
        return x;
    }
};

struct Inputs {
    double lb[2], ub[2], eta[2];
    uint32_t shape[2];
    uint64_t strides[2];
    uint64_t total;
    void center(uint32_t flat, double out[2]) const {
        // grid.hpp:68-78
        uint64_t v = flat;
        uint32_t idx[2] = {0, 0};
        for (int d = 1; d >= 0; --d) {
            idx[d] = static_cast<uint32_t>(v / strides[d]);
            v %= strides[d];
        }
        for (int d = 0; d < 2; ++d) out[d] = lb[d] + (static_cast<double>(idx[d]) + 0.5) * eta[d];
    }
};

struct Layout {  // PrefixLayout4D, grid.hpp:87-101, 253-269
    uint32_t pshape[kD];
    uint64_t strides[kD];
    uint64_t total;
    uint64_t offset(uint64_t a, uint64_t b, uint64_t c, uint64_t d) const {
        return a * strides[0] + b * strides[1] + c * strides[2] + d * strides[3];
    }
};

struct IBox {
    uint32_t min[kD];
    uint32_t max[kD];
};

Grid make_grid(const gsc_case_desc& c) {
    // grid.hpp:103-139
    Grid g{};
    g.wrap_dim = c.wrap_dim;
    for (int d = 0; d < kD; ++d) {
        g.lb[d] = c.state_lb[d];
        g.ub[d] = c.state_ub[d];
        g.eta[d] = c.state_eta[d];
        if (!(c.state_eta[d] > 0.0)) throw std::runtime_error("state eta must be positive");
        const double width = c.state_ub[d] - c.state_lb[d];
        if (!(width > 0.0)) throw std::runtime_error("state bounds must satisfy ub > lb");
        g.shape[d] = static_cast<uint32_t>(std::ceil(width / c.state_eta[d]));
    }
    g.strides[0] = 1;
    for (int d = 1; d < kD; ++d) g.strides[d] = g.strides[d - 1] * g.shape[d - 1];
    uint64_t total = 1;
    for (int d = 0; d < kD; ++d) total *= g.shape[d];
    g.total = total;
    if (total > 0xFFFFFFFFull) throw std::runtime_error("state grid exceeds the selected state index type");
    return g;
}

Inputs make_inputs(const gsc_case_desc& c) {
    // grid.hpp:141-168
    Inputs in{};
    for (int d = 0; d < 2; ++d) {
        in.lb[d] = c.input_lb[d];
        in.ub[d] = c.input_ub[d];
        in.eta[d] = c.input_eta[d];
        if (!(c.input_eta[d] > 0.0)) throw std::runtime_error("input eta must be positive");
        const double width = c.input_ub[d] - c.input_lb[d];
        if (!(width > 0.0)) throw std::runtime_error("input bounds must satisfy ub > lb");
        in.shape[d] = static_cast<uint32_t>(std::ceil(width / c.input_eta[d]));
    }
    in.strides[0] = 1;
    in.strides[1] = in.shape[0];
    in.total = static_cast<uint64_t>(in.shape[0]) * in.shape[1];
    if (in.total > 0xFFFFFFFFull) throw std::runtime_error("input grid exceeds uint32_t");
    return in;
}

Layout make_layout(const Grid& g) {
    Layout l{};
    for (int d = 0; d < kD; ++d) l.pshape[d] = g.shape[d] + 1;
    l.strides[0] = 1;
    for (int d = 1; d < kD; ++d) l.strides[d] = l.strides[d - 1] * l.pshape[d - 1];
    l.total = static_cast<uint64_t>(l.pshape[0]) * l.pshape[1] * l.pshape[2] * l.pshape[3];
    return l;
}

bool point_in_rect(const gsc_rect4& r, const double x[kD]) {
    // config.cpp:79-86
    for (int d = 0; d < kD; ++d)
        if (x[d] < r.lb[d] || x[d] > r.ub[d]) return false;
    return true;
}

bool satisfies_constraint(const gsc_case_desc& c, const double x[kD]) {
    // config.cpp:64-75, 111-127
    if (c.constraint_type == GSC_CONSTRAINT_HYPERBOLIC) {
        const double x1_sq = x[0] * x[0];
        const double x2_sq = x[1] * x[1];
        return (x1_sq - x2_sq <= c.hyperbolic_a) && (c.hyperbolic_b * x2_sq - x1_sq <= c.hyperbolic_c);
    }
    if (c.constraint_type == GSC_CONSTRAINT_ELLIPTIC) {
        const double t1 = (x[0] / c.elliptic_a) * (x[0] / c.elliptic_a);
        const double t2 = (x[1] / c.elliptic_b) * (x[1] / c.elliptic_b);
        return t1 + t2 <= 1.0;
    }
    return true;
}

// Extension (no reference counterpart): candidate = (centre in `candidate` and,
// if enabled, |x*y - c| <= w) or centre in any extra target box.
bool in_candidate(const gsc_case_desc& c, const double x[kD]) {
    bool in = point_in_rect(c.candidate, x);
    if (in && c.target_band_enabled) in = std::fabs(x[0] * x[1] - c.target_band_c) <= c.target_band_w;
    if (!in)
        for (uint32_t k = 0; k < c.extra_target_count; ++k)
            if (point_in_rect(c.extra_targets[k], x)) return true;
    return in;
}

struct Model {  // Unicycle4DModel, model.hpp:9-113
    double sample_time;
    uint32_t substeps;
    double dw[kD];

    static double wrap_angle(double x) {
        constexpr double kPi = 3.14159265358979323846;
        constexpr double kTwoPi = 2.0 * kPi;
        while (x < -kPi) x += kTwoPi;
        while (x >= kPi) x -= kTwoPi;
        return x;
    }
    void nominal_step(const double x[kD], const double u[2], double out[kD]) const {
        // model.hpp:26-46
        for (int d = 0; d < kD; ++d) out[d] = x[d];
        const double dt = sample_time / static_cast<double>(substeps);
        for (uint32_t s = 0; s < substeps; ++s) {
            const double dx = out[3] * std::cos(out[2]);
            const double dy = out[3] * std::sin(out[2]);
            const double dtheta = u[1];
            const double dv = u[0];
            out[0] += dt * dx;
            out[1] += dt * dy;
            out[2] = wrap_angle(out[2] + dt * dtheta);
            out[3] += dt * dv;
        }
    }
    void successor_box(const double x[kD], const double u[2], const double eta[kD], double mn[kD],
                       double mx[kD]) const {
        // model.hpp:55-112
        double nominal[kD];
        nominal_step(x, u, nominal);
        const double h[kD] = {0.5 * eta[0], 0.5 * eta[1], 0.5 * eta[2], 0.5 * eta[3]};
        const double abs_v = std::fabs(x[3]);
        double jfx[kD];
        jfx[0] = h[0] + sample_time * abs_v * h[2] + sample_time * h[3];
        jfx[1] = h[1] + sample_time * abs_v * h[2] + sample_time * h[3];
        jfx[2] = h[2];
        jfx[3] = h[3];
        for (int d = 0; d < kD; ++d) {
            const double radius = jfx[d] + sample_time * dw[d];
            mn[d] = nominal[d] - radius;
            mx[d] = nominal[d] + radius;
        }
    }
};

Model make_model(const gsc_case_desc& c) {
    Model m{};
    m.sample_time = c.sample_time;
    m.substeps = c.integration_substeps;
    for (int d = 0; d < kD; ++d) m.dw[d] = c.disturbance_half_width[d];
    return m;
}

int64_t clamp_cell(int64_t v, uint32_t n) {
    if (v < 0) v = 0;
    if (v >= static_cast<int64_t>(n)) v = static_cast<int64_t>(n) - 1;
    return v;
}

bool to_index_box(const Grid& g, const double mn[kD], const double mx[kD], IBox* out) {
    // grid.hpp:181-251. Note the two separate clamps of lower and upper.
    constexpr double kEps = 1e-9;
    for (int d = 0; d < kD; ++d) {
        if (d == g.wrap_dim) continue;
        if (mn[d] < g.lb[d] - kEps || mx[d] > g.ub[d] + kEps) return false;
    }
    const double wrap_span = g.ub[g.wrap_dim] - g.lb[g.wrap_dim];
    const double raw_wrap_width = mx[g.wrap_dim] - mn[g.wrap_dim];
    for (int d = 0; d < kD; ++d) {
        double lo_c = mn[d];
        double hi_c = mx[d];
        if (d == g.wrap_dim) {
            if (raw_wrap_width >= wrap_span - kEps) {
                out->min[d] = 0;
                out->max[d] = g.shape[d] - 1;
                continue;
            }
            lo_c = g.wrap_value(mn[d]);
            hi_c = g.wrap_value(mx[d]);
        }
        const int64_t lower = static_cast<int64_t>(std::floor((lo_c - g.lb[d]) / g.eta[d]));
        const int64_t upper = static_cast<int64_t>(std::ceil((hi_c - g.lb[d]) / g.eta[d])) - 1;
        out->min[d] = static_cast<uint32_t>(clamp_cell(lower, g.shape[d]));
        out->max[d] = static_cast<uint32_t>(clamp_cell(upper, g.shape[d]));
    }
    return true;
}

bool box_satisfies_constraint(const Grid& g, const IBox& box, const gsc_case_desc& c) {
    // abstraction.hpp:110-159, restated with its exact loop structure (including
    // the θ loop's wrap walk, which is keyed on wrap_dim == 2).
    if (c.constraint_type == GSC_CONSTRAINT_NONE) return true;
    const int wd = g.wrap_dim;
    const bool has_wrap = box.min[wd] > box.max[wd];
    uint32_t idx[kD];
    for (idx[3] = box.min[3]; idx[3] <= box.max[3]; ++idx[3]) {
        for (idx[2] = box.min[2];; ++idx[2]) {
            bool in2;
            if (wd == 2 && has_wrap)
                in2 = (idx[2] >= box.min[2] || idx[2] <= box.max[2]);
            else
                in2 = (idx[2] <= box.max[2]);
            if (!in2 && idx[2] > box.max[2] && (!has_wrap || idx[2] >= g.shape[2])) break;
            for (idx[1] = box.min[1]; idx[1] <= box.max[1]; ++idx[1]) {
                for (idx[0] = box.min[0]; idx[0] <= box.max[0]; ++idx[0]) {
                    double ctr[kD];
                    g.center(g.flatten(idx), ctr);
                    if (!satisfies_constraint(c, ctr)) return false;
                }
            }
            if (wd == 2 && has_wrap && idx[2] == g.shape[2] - 1)
                idx[2] = static_cast<uint32_t>(-1);
            else if (idx[2] == box.max[2])
                break;
        }
    }
    return true;
}

template <typename F>
void parallel_for(uint64_t begin, uint64_t end, int threads, F&& fn) {
    if (threads <= 1 || end - begin < 4096) {
        fn(begin, end);
        return;
    }
    std::vector<std::thread> pool;
    const uint64_t chunk = 1024;
    std::atomic<uint64_t> next{begin};
    for (int t = 0; t < threads; ++t)
        pool.emplace_back([&]() {
            for (;;) {
                const uint64_t b = next.fetch_add(chunk);
                if (b >= end) return;
                fn(b, std::min(end, b + chunk));
            }
        });
    for (auto& th : pool) th.join();
}

void copy_err(const char* msg, char* err, size_t len) {
    if (err && len) {
        std::snprintf(err, len, "%s", msg);
    }
}

std::vector<uint64_t> prefix_u64(const Grid& g, const Layout& l, const uint8_t* mask) {
    // prefix_sum.hpp:15-63 (16-term inclusion–exclusion recurrence).
    std::vector<uint64_t> P(l.total, 0);
    for (uint64_t w = 1; w < l.pshape[3]; ++w)
        for (uint64_t z = 1; z < l.pshape[2]; ++z)
            for (uint64_t y = 1; y < l.pshape[1]; ++y)
                for (uint64_t x = 1; x < l.pshape[0]; ++x) {
                    const uint32_t idx[kD] = {uint32_t(x - 1), uint32_t(y - 1), uint32_t(z - 1), uint32_t(w - 1)};
                    const uint64_t value = mask[g.flatten(idx)] ? 1 : 0;
                    P[l.offset(x, y, z, w)] =
                        value + P[l.offset(x - 1, y, z, w)] + P[l.offset(x, y - 1, z, w)] +
                        P[l.offset(x, y, z - 1, w)] + P[l.offset(x, y, z, w - 1)] -
                        P[l.offset(x - 1, y - 1, z, w)] - P[l.offset(x - 1, y, z - 1, w)] -
                        P[l.offset(x - 1, y, z, w - 1)] - P[l.offset(x, y - 1, z - 1, w)] -
                        P[l.offset(x, y - 1, z, w - 1)] - P[l.offset(x, y, z - 1, w - 1)] +
                        P[l.offset(x - 1, y - 1, z - 1, w)] + P[l.offset(x - 1, y - 1, z, w - 1)] +
                        P[l.offset(x - 1, y, z - 1, w - 1)] + P[l.offset(x, y - 1, z - 1, w - 1)] -
                        P[l.offset(x - 1, y - 1, z - 1, w - 1)];
                }
    return P;
}

uint64_t query_single(const Layout& l, const uint64_t* P, const IBox& b) {
    // prefix_sum.hpp:65-85 (u64 modular arithmetic, as the reference).
    const uint32_t lo[kD] = {b.min[0], b.min[1], b.min[2], b.min[3]};
    const uint32_t hi[kD] = {b.max[0] + 1, b.max[1] + 1, b.max[2] + 1, b.max[3] + 1};
    uint64_t total = 0;
    for (int m = 0; m < 16; ++m) {
        const uint64_t v = P[l.offset((m & 1) ? lo[0] : hi[0], (m & 2) ? lo[1] : hi[1], (m & 4) ? lo[2] : hi[2],
                                      (m & 8) ? lo[3] : hi[3])];
        if ((__builtin_popcount(static_cast<unsigned>(m)) & 1) == 0)
            total += v;
        else
            total -= v;
    }
    return total;
}

uint64_t query_count(const Grid& g, const Layout& l, const uint64_t* P, uint32_t min_flat, uint32_t max_flat) {
    // prefix_sum.hpp:87-105: split a box whose wrap-dimension range wraps.
    IBox b{};
    g.unflatten(min_flat, b.min);
    g.unflatten(max_flat, b.max);
    const int wd = g.wrap_dim;
    if (b.min[wd] <= b.max[wd]) return query_single(l, P, b);
    IBox left = b, right = b;
    left.max[wd] = g.shape[wd] - 1;
    right.min[wd] = 0;
    return query_single(l, P, left) + query_single(l, P, right);
}

}  // namespace

extern "C" {

int oracle_shape(const gsc_case_desc* c, gsc_case_shape* out, char* err, size_t err_len) {
    try {
        const Grid g = make_grid(*c);
        const Inputs in = make_inputs(*c);
        const Layout l = make_layout(g);
        std::memset(out, 0, sizeof(*out));
        for (int d = 0; d < kD; ++d) {
            out->state_shape[d] = g.shape[d];
            out->state_strides[d] = g.strides[d];
            out->prefix_shape[d] = l.pshape[d];
        }
        out->state_count = g.total;
        for (int d = 0; d < 2; ++d) {
            out->input_shape[d] = in.shape[d];
            out->input_strides[d] = in.strides[d];
        }
        out->input_count = in.total;
        out->pair_count = g.total * in.total;
        out->prefix_count = l.total;
        return 0;
    } catch (const std::exception& ex) {
        copy_err(ex.what(), err, err_len);
        return GSC_ERR_CONFIG;
    }
}

// build_valid_mask / build_candidate_mask and the candidate ⊆ valid check
// (abstraction.hpp:65-106, 257-265).
int oracle_masks(const gsc_case_desc* c, uint8_t* valid, uint8_t* candidate, char* err, size_t err_len) {
    try {
        const Grid g = make_grid(*c);
        for (uint64_t s = 0; s < g.total; ++s) {
            double x[kD];
            g.center(static_cast<uint32_t>(s), x);
            bool ok = point_in_rect(c->map, x);
            if (ok)
                for (uint32_t k = 0; k < c->obstacle_count; ++k)
                    if (point_in_rect(c->obstacles[k], x)) {
                        ok = false;
                        break;
                    }
            if (ok && c->constraint_type != GSC_CONSTRAINT_NONE) ok = satisfies_constraint(*c, x);
            valid[s] = ok ? 1 : 0;
            candidate[s] = in_candidate(*c, x) ? 1 : 0;
        }
        for (uint64_t s = 0; s < g.total; ++s)
            if (candidate[s] && !valid[s]) throw std::runtime_error("candidate set contains invalid states");
        return 0;
    } catch (const std::exception& ex) {
        copy_err(ex.what(), err, err_len);
        return GSC_ERR_CONFIG;
    }
}

// Successor box of one (state, input) pair in continuous coordinates (model.hpp:91-112).
int oracle_successor_box(const gsc_case_desc* c, uint32_t state, uint32_t input, double* mn, double* mx) {
    try {
        const Grid g = make_grid(*c);
        const Inputs in = make_inputs(*c);
        const Model m = make_model(*c);
        double x[kD], u[2];
        g.center(state, x);
        in.center(input, u);
        m.successor_box(x, u, g.eta, mn, mx);
        return 0;
    } catch (...) {
        return GSC_ERR_CONFIG;
    }
}

// build_abstraction_cpu (abstraction.hpp:161-209) for states [state_begin, state_end);
// outputs are indexed from pair state_begin*I.
int oracle_abstraction(const gsc_case_desc* c, uint64_t state_begin, uint64_t state_end, uint32_t* min_flat,
                       uint32_t* max_flat, uint8_t* pvalid, int threads) {
    try {
        const Grid g = make_grid(*c);
        const Inputs in = make_inputs(*c);
        const Model m = make_model(*c);
        if (state_end > g.total) state_end = g.total;
        const uint64_t I = in.total;
        parallel_for(state_begin, state_end, threads, [&](uint64_t b, uint64_t e) {
            for (uint64_t s = b; s < e; ++s) {
                const uint64_t base = (s - state_begin) * I;
                for (uint64_t i = 0; i < I; ++i) {
                    min_flat[base + i] = 0;
                    max_flat[base + i] = 0;
                    pvalid[base + i] = 0;
                }
                double x[kD];
                g.center(static_cast<uint32_t>(s), x);
                if (c->constraint_type != GSC_CONSTRAINT_NONE && !satisfies_constraint(*c, x)) continue;
                for (uint64_t i = 0; i < I; ++i) {
                    double u[2], mn[kD], mx[kD];
                    in.center(static_cast<uint32_t>(i), u);
                    m.successor_box(x, u, g.eta, mn, mx);
                    IBox box{};
                    if (!to_index_box(g, mn, mx, &box)) continue;
                    if (!box_satisfies_constraint(g, box, *c)) continue;
                    min_flat[base + i] = g.flatten(box.min);
                    max_flat[base + i] = g.flatten(box.max);
                    pvalid[base + i] = 1;
                }
            }
        });
        return 0;
    } catch (...) {
        return GSC_ERR_CONFIG;
    }
}

// Padded u64 prefix table of a mask (prefix_sum.hpp:15-63).
int oracle_prefix_sum(const gsc_case_desc* c, const uint8_t* mask, uint64_t* table) {
    try {
        const Grid g = make_grid(*c);
        const Layout l = make_layout(g);
        const auto P = prefix_u64(g, l, mask);
        std::memcpy(table, P.data(), P.size() * sizeof(uint64_t));
        return 0;
    } catch (...) {
        return GSC_ERR_CONFIG;
    }
}

// query_box_count (prefix_sum.hpp:87-105) for n (min_flat, max_flat) pairs.
int oracle_query_box_count(const gsc_case_desc* c, const uint64_t* table, const uint32_t* min_flat,
                           const uint32_t* max_flat, uint64_t n, uint64_t* out) {
    try {
        const Grid g = make_grid(*c);
        const Layout l = make_layout(g);
        for (uint64_t k = 0; k < n; ++k) out[k] = query_count(g, l, table, min_flat[k], max_flat[k]);
        return 0;
    } catch (...) {
        return GSC_ERR_CONFIG;
    }
}

// solve_backward_cpu (reachability.hpp:30-155): Jacobi fixpoint over the
// prefix snapshot of the previous reachable set; first satisfied input wins.
// newly_reachable / newly_certified receive per-iteration counts (capacity
// stats_capacity). Returns the stop reason in `message`.
int oracle_solve(const gsc_case_desc* c, const uint8_t* valid_mask, const uint8_t* candidate_mask,
                 const uint32_t* min_flat, const uint32_t* max_flat, const uint8_t* pvalid, uint8_t* reachable,
                 uint32_t* controller, uint32_t* reach_step, uint64_t* newly_reachable, uint64_t* newly_certified,
                 uint32_t stats_capacity, uint32_t* iterations, int32_t* converged, char* message, size_t msg_len,
                 int threads) {
    try {
        const Grid g = make_grid(*c);
        const Inputs in = make_inputs(*c);
        const Layout l = make_layout(g);
        const uint64_t S = g.total, I = in.total;
        uint64_t candidate_count = 0;
        for (uint64_t s = 0; s < S; ++s) {
            reachable[s] = candidate_mask[s];
            controller[s] = GSC_INVALID_INPUT;
            reach_step[s] = GSC_UNREACHABLE_STEP;
            if (candidate_mask[s]) {
                reach_step[s] = 0;
                ++candidate_count;
            }
        }
        const auto PV = prefix_u64(g, l, valid_mask);
        uint64_t certified = 0;
        *iterations = 0;
        *converged = 0;
        std::string msg;
        std::vector<uint8_t> next(S);
        for (uint32_t iter = 1; iter <= c->max_iterations; ++iter) {
            const auto PR = prefix_u64(g, l, reachable);
            std::memcpy(next.data(), reachable, S);
            std::atomic<uint64_t> new_reach{0}, new_cert{0};
            parallel_for(0, S, threads, [&](uint64_t b, uint64_t e) {
                uint64_t nr = 0, nc = 0;
                for (uint64_t s = b; s < e; ++s) {
                    if (!valid_mask[s]) continue;
                    const bool is_cand = candidate_mask[s] != 0;
                    const bool reached = reachable[s] != 0;
                    const bool certified_s = controller[s] != GSC_INVALID_INPUT;
                    if ((!is_cand && reached) || (is_cand && certified_s)) continue;
                    for (uint64_t i = 0; i < I; ++i) {
                        const uint64_t p = s * I + i;
                        if (!pvalid[p]) continue;
                        const uint64_t vc = query_count(g, l, PV.data(), min_flat[p], max_flat[p]);
                        if (vc == 0) continue;
                        const uint64_t rc = query_count(g, l, PR.data(), min_flat[p], max_flat[p]);
                        if (rc == vc) {
                            if (!reached) {
                                next[s] = 1;
                                ++nr;
                            }
                            if (!certified_s) {
                                controller[s] = static_cast<uint32_t>(i);
                                if (is_cand)
                                    ++nc;
                                else
                                    reach_step[s] = iter;
                            }
                            break;
                        }
                    }
                }
                new_reach += nr;
                new_cert += nc;
            });
            std::memcpy(reachable, next.data(), S);
            certified += new_cert.load();
            *iterations = iter;
            if (iter - 1 < stats_capacity) {
                newly_reachable[iter - 1] = new_reach.load();
                newly_certified[iter - 1] = new_cert.load();
            }
            if (certified == candidate_count) {
                *converged = 1;
                msg = "all candidate states obtained a controller";
                break;
            }
            if (new_reach.load() == 0 && new_cert.load() == 0) {
                msg = "iteration stalled before candidate closure";
                break;
            }
        }
        if (!*converged && msg.empty()) msg = "maximum iterations reached";
        copy_err(msg.c_str(), message, msg_len);
        return 0;
    } catch (const std::exception& ex) {
        copy_err(ex.what(), message, msg_len);
        return GSC_ERR_CONFIG;
    }
}

}  // extern "C"
