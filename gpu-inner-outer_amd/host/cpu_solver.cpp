// cpu_solver.cpp — the host solver behind the CLI's --cpu and --compare modes.
//
// Same results as the reference's CPU path (abstraction.hpp:161-209, 287-306;
// reachability.hpp:31-155), computed differently:
//
//  - Separable successor boxes. The θ and v trajectories (model.hpp:34-45)
//    depend only on (θ cell, ω) and (v cell, a), and the nominal x (y) is the
//    cell centre plus K increments that depend only on the (θ, v) plane and the
//    input. The radius depends on the v cell alone. So the x cell range of a
//    pair is a function of (plane, input, x cell), its y range of (plane,
//    input, y cell), its θ range of (θ cell, ω) and its v range of (v cell, a):
//    the abstraction is four interval tables instead of 9 B per pair.
//  - Every interval is computed with the reference's arithmetic in its order:
//    centres (grid.hpp:42), increments added one substep at a time
//    (model.hpp:38-44), radius (model.hpp:71-73, 84-89, 107-111), cells
//    (grid.hpp:181-251). The file is compiled with -ffp-contract=off.
//  - The box constraint scan (abstraction.hpp:110-159) reads x and y only, so
//    "every cell of the box satisfies the constraint" is a 2D prefix query over
//    the failing (x, y) cells, with the scan's empty-loop cases kept.
//  - The fixpoint is the reference's Jacobi iteration (reachability.hpp:54-132):
//    a pair is satisfied iff its box holds at least one valid cell and no valid
//    unreached one (reachable_count == valid_count with valid_count > 0, as
//    reached ⊆ valid), counted on u32 4D prefix tables (prefix_sum.hpp:15-63,
//    query_box_count :94-112). States are processed in parallel; each thread
//    writes only its own states, and every count reads the iteration's
//    snapshot, so the update order does not matter. Before a pair's count,
//    three rows of its box are probed in a bitmask of the valid unreached
//    cells (0.5 MB at 4M states, cache-resident): a set bit proves the pair
//    unsatisfied, which settles almost every pair without the count.
//  - Frontier skipping: a state that was tested and not satisfied keeps that
//    answer until a cell of one of its boxes turns from unreached to reached.
//    Each state keeps the union of its valid boxes (its hull); when the
//    unreached count over the hull is the same in this iteration's table as
//    in the previous one, nothing in it changed and the state is skipped (the
//    unreached set only shrinks, so equal counts mean equal sets). In the
//    first iteration the previous table is the valid one: a hull with no
//    reached cell satisfies no input. Used when the wrap axis is θ or v.
#include "gsc/cpu_solver.hpp"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <functional>
#include <iostream>
#include <stdexcept>
#include <thread>
#include <vector>

namespace gsc {

namespace {

constexpr double kEps = 1e-9;                   // grid.hpp:186
constexpr std::uint32_t kReject = 0xFFFFFFFFu;  // interval or hull: no box
constexpr std::uint32_t kMaxXYCells = 65535;    // x / y cells in a packed interval

using Clock = std::chrono::steady_clock;

double ms_since(Clock::time_point t0) {
    return std::chrono::duration<double, std::milli>(Clock::now() - t0).count();
}

// An x or y cell interval packed in one word: lo in bits 0-15, hi in bits
// 16-31; lo > hi is a range that wraps (wrap axis x or y).
inline std::uint32_t pack(std::uint32_t lo, std::uint32_t hi) { return lo | (hi << 16); }
inline std::uint32_t lo_of(std::uint32_t r) { return r & 0xFFFFu; }
inline std::uint32_t hi_of(std::uint32_t r) { return r >> 16; }

// Runs body(begin, end) over [0, n) in chunks of `grain`, handed out to
// `threads` threads in order of request. body must not throw.
void parallel_for(unsigned threads, std::uint64_t n, std::uint64_t grain,
                  const std::function<void(std::uint64_t, std::uint64_t)>& body) {
    if (n == 0) return;
    grain = std::max<std::uint64_t>(1, grain);
    const std::uint64_t chunks = (n + grain - 1) / grain;
    const unsigned t_count = static_cast<unsigned>(std::min<std::uint64_t>(std::max(1u, threads), chunks));
    std::atomic<std::uint64_t> next{0};
    auto worker = [&]() {
        for (;;) {
            const std::uint64_t b = next.fetch_add(grain, std::memory_order_relaxed);
            if (b >= n) break;
            body(b, std::min(n, b + grain));
        }
    };
    if (t_count == 1) {
        worker();
        return;
    }
    std::vector<std::thread> pool;
    pool.reserve(t_count - 1);
    for (unsigned t = 1; t < t_count; ++t) pool.emplace_back(worker);
    worker();
    for (auto& th : pool) th.join();
}

double wrap_value(double low, double high, double x) {
    // grid.hpp:46-57
    const double span = high - low;
    while (x < low) x += span;
    while (x >= high) x -= span;
    return x;
}

double wrap_angle(double x) {
    // model.hpp:14-24 (literal π, not the grid bounds)
    constexpr double kPi = 3.14159265358979323846;
    constexpr double kTwoPi = 2.0 * kPi;
    while (x < -kPi) x += kTwoPi;
    while (x >= kPi) x -= kTwoPi;
    return x;
}

struct Case {
    Grid4D<std::uint32_t> g;
    InputGrid2D in;
    std::uint32_t n[4]{};
    std::uint64_t S = 0;
    std::uint32_t I = 0, n_a = 0, n_w = 0, K = 0;
    std::uint32_t PQ = 0, planes = 0;
    int wd = 2;
    double T = 0.0, dt = 0.0;
    std::uint64_t ps[4]{};  // padded prefix strides (grid.hpp:253-269)
    std::uint64_t PS = 0;
    bool constrained = false;
    bool frontier = false;

    double center(int d, std::uint32_t idx) const {
        return g.lb[d] + (static_cast<double>(idx) + 0.5) * g.eta[d];  // grid.hpp:42
    }
    double input_center(int d, std::uint32_t idx) const {
        return in.lb[d] + (static_cast<double>(idx) + 0.5) * in.eta[d];  // grid.hpp:76
    }

    // One dimension of continuous_box_to_index_box (grid.hpp:181-251): false
    // when a non-wrap dimension leaves [lb - eps, ub + eps].
    bool cells(int d, double mn, double mx, std::uint32_t* lo, std::uint32_t* hi) const {
        double a = mn, b = mx;
        if (d == wd) {
            const double span = g.ub[d] - g.lb[d];
            if (mx - mn >= span - kEps) {
                *lo = 0;
                *hi = n[d] - 1;
                return true;
            }
            a = wrap_value(g.lb[d], g.ub[d], mn);
            b = wrap_value(g.lb[d], g.ub[d], mx);
        } else if (mn < g.lb[d] - kEps || mx > g.ub[d] + kEps) {
            return false;
        }
        auto lower = static_cast<std::int64_t>(std::floor((a - g.lb[d]) / g.eta[d]));
        auto upper = static_cast<std::int64_t>(std::ceil((b - g.lb[d]) / g.eta[d])) - 1;
        const auto top = static_cast<std::int64_t>(n[d]);
        if (lower < 0) lower = 0;
        if (upper < 0) upper = 0;
        if (lower >= top) lower = top - 1;
        if (upper >= top) upper = top - 1;
        *lo = static_cast<std::uint32_t>(lower);
        *hi = static_cast<std::uint32_t>(upper);
        return true;
    }
};

// The reference's grids (grid.hpp:103-168), with its error messages.
void build_grids(const CaseConfig& cfg, Case& c) {
    if (cfg.wrap_dim < 0 || cfg.wrap_dim > 3) throw std::runtime_error("wrap_dim must be 0, 1, 2 or 3");
    auto& g = c.g;
    g.wrap_dim = cfg.wrap_dim;
    for (int d = 0; d < kStateDim; ++d) {
        g.lb[d] = cfg.state_lb[d];
        g.ub[d] = cfg.state_ub[d];
        g.eta[d] = cfg.state_eta[d];
        if (!(g.eta[d] > 0.0)) throw std::runtime_error("state eta must be positive");
        const double width = g.ub[d] - g.lb[d];
        if (!(width > 0.0)) throw std::runtime_error("state bounds must satisfy ub > lb");
        g.shape[d] = static_cast<std::uint32_t>(std::ceil(width / g.eta[d]));
    }
    g.strides[0] = 1;
    for (int d = 1; d < kStateDim; ++d) g.strides[d] = g.strides[d - 1] * g.shape[d - 1];
    g.total_size = g.strides[3] * g.shape[3];
    if (g.total_size > 0xFFFFFFFFull) throw std::runtime_error("state grid exceeds the selected state index type");

    auto& in = c.in;
    for (int d = 0; d < kInputDim; ++d) {
        in.lb[d] = cfg.input_lb[d];
        in.ub[d] = cfg.input_ub[d];
        in.eta[d] = cfg.input_eta[d];
        if (!(in.eta[d] > 0.0)) throw std::runtime_error("input eta must be positive");
        const double width = in.ub[d] - in.lb[d];
        if (!(width > 0.0)) throw std::runtime_error("input bounds must satisfy ub > lb");
        in.shape[d] = static_cast<std::uint32_t>(std::ceil(width / in.eta[d]));
    }
    in.strides[0] = 1;
    in.strides[1] = in.shape[0];
    in.total_size = static_cast<std::uint64_t>(in.shape[0]) * in.shape[1];
    if (in.total_size > 0xFFFFFFFFull) throw std::runtime_error("input grid exceeds uint32_t");

    for (int d = 0; d < kStateDim; ++d) c.n[d] = g.shape[d];
    c.S = g.total_size;
    c.I = static_cast<std::uint32_t>(in.total_size);
    c.n_a = in.shape[0];
    c.n_w = in.shape[1];
    c.K = cfg.integration_substeps;
    c.wd = cfg.wrap_dim;
    c.T = cfg.sample_time;
    c.dt = cfg.sample_time / static_cast<double>(cfg.integration_substeps);  // model.hpp:34
    c.PQ = c.n[0] * c.n[1];
    c.planes = c.n[2] * c.n[3];
    c.ps[0] = 1;
    for (int d = 1; d < kStateDim; ++d) c.ps[d] = c.ps[d - 1] * (c.n[d - 1] + 1ull);
    c.PS = c.ps[3] * (c.n[3] + 1ull);
    c.constrained = cfg.constraint_type != ConstraintType::kNone;
    c.frontier = c.wd == 2 || c.wd == 3;
}

MemoryBudget reference_budget(const Case& c) {
    // abstraction.hpp:211-245 with 32-bit state indices and 64-bit prefix counts
    MemoryBudget b;
    const std::uint64_t P = c.S * c.I;
    b.pair_min_bytes = P * 4;
    b.pair_max_bytes = P * 4;
    b.pair_valid_bytes = P;
    b.pair_satisfied_bytes = P;
    b.reachable_prev_bytes = c.S;
    b.reachable_next_bytes = c.S;
    b.valid_mask_bytes = c.S;
    b.candidate_mask_bytes = c.S;
    b.controller_bytes = c.S * 4;
    b.reach_step_bytes = c.S * 4;
    b.prefix_reachable_bytes = c.PS * 8;
    b.prefix_valid_bytes = c.PS * 8;
    b.total_bytes = b.pair_min_bytes + b.pair_max_bytes + b.pair_valid_bytes + b.pair_satisfied_bytes +
                    b.reachable_prev_bytes + b.reachable_next_bytes + b.valid_mask_bytes + b.candidate_mask_bytes +
                    b.controller_bytes + b.reach_step_bytes + b.prefix_reachable_bytes + b.prefix_valid_bytes;
    return b;
}

// valid = centre in the map, in no obstacle, satisfying the constraint;
// candidate = centre in the candidate box (abstraction.hpp:65-106), with this
// backend's extension keys applied as the GPU path applies them.
void build_masks(const CaseConfig& cfg, const CaseExtensions& ext, const Case& c, unsigned threads,
                 std::vector<std::uint8_t>& valid, std::vector<std::uint8_t>& cand) {
    valid.assign(c.S, 0);
    cand.assign(c.S, 0);
    parallel_for(threads, c.S, 1u << 14, [&](std::uint64_t b, std::uint64_t e) {
        for (std::uint64_t s = b; s < e; ++s) {
            std::uint64_t v = s;
            double x[kStateDim];
            for (int d = 0; d < kStateDim; ++d) {
                x[d] = c.center(d, static_cast<std::uint32_t>(v % c.n[d]));
                v /= c.n[d];
            }
            bool ok = point_in_rect(cfg.map, x);
            if (ok)
                for (const auto& o : cfg.obstacles)
                    if (point_in_rect(o, x)) {
                        ok = false;
                        break;
                    }
            if (ok && c.constrained) ok = satisfies_state_constraint(cfg, x);
            bool is_cand = point_in_rect(cfg.candidate, x);
            if (is_cand && ext.target_band_enabled)
                is_cand = std::fabs(x[0] * x[1] - ext.target_band_c) <= ext.target_band_w;
            if (!is_cand)
                for (const auto& t : ext.extra_targets)
                    if (point_in_rect(t, x)) {
                        is_cand = true;
                        break;
                    }
            valid[s] = ok ? 1 : 0;
            cand[s] = is_cand ? 1 : 0;
        }
    });
}

// The separable abstraction.
struct Abstraction {
    // θ range of the box per (θ cell, ω input), v range per (v cell, a input);
    // lo = kReject when the range leaves the domain.
    std::vector<std::uint32_t> th_lo, th_hi, v_lo, v_hi;
    // x interval per (plane, x cell, input), y interval per (plane, y cell,
    // input): packed, or kReject.
    std::vector<std::uint32_t> xr, yr;
    // 2D prefix of constraint failures over (x, y), (n_x + 1) x (n_y + 1).
    std::vector<std::uint32_t> fail;
    // Frontier hulls: per state, the union of its valid boxes' x and y
    // intervals (packed; kReject for a state with no valid pair); per θ cell
    // and per v cell, the union of the θ / v ranges over the inputs.
    std::vector<std::uint32_t> hull_x, hull_y;
    std::vector<std::uint32_t> hull_th_lo, hull_th_hi, hull_v_lo, hull_v_hi;
    std::uint64_t valid_pairs = 0;
};

// Cells of the pair (state (ix, iy) of plane p, input i) into lo/hi; false for
// an invalid pair (abstraction.hpp:187-205). `th` / `vv` are the plane's θ / v
// ranges per input.
struct PlaneRanges {
    std::vector<std::uint32_t> th_lo, th_hi, v_lo, v_hi;
    void load(const Case& c, const Abstraction& a, std::uint32_t p) {
        const std::uint32_t it = p % c.n[2], iv = p / c.n[2];
        th_lo.resize(c.I);
        th_hi.resize(c.I);
        v_lo.resize(c.I);
        v_hi.resize(c.I);
        for (std::uint32_t i = 0; i < c.I; ++i) {
            const std::uint32_t ia = i % c.n_a, iw = i / c.n_a;
            th_lo[i] = a.th_lo[it * c.n_w + iw];
            th_hi[i] = a.th_hi[it * c.n_w + iw];
            v_lo[i] = a.v_lo[iv * c.n_a + ia];
            v_hi[i] = a.v_hi[iv * c.n_a + ia];
        }
    }
};

inline bool pair_cells(const Case& c, const Abstraction& a, const PlaneRanges& pr, std::uint64_t xrow,
                       std::uint64_t yrow, std::uint32_t i, std::uint32_t lo[4], std::uint32_t hi[4]) {
    const std::uint32_t x = a.xr[xrow + i], y = a.yr[yrow + i];
    if (x == kReject || y == kReject || pr.th_lo[i] == kReject || pr.v_lo[i] == kReject) return false;
    lo[0] = lo_of(x);
    hi[0] = hi_of(x);
    lo[1] = lo_of(y);
    hi[1] = hi_of(y);
    lo[2] = pr.th_lo[i];
    hi[2] = pr.th_hi[i];
    lo[3] = pr.v_lo[i];
    hi[3] = pr.v_hi[i];
    if (c.constrained) {
        // abstraction.hpp:110-159: a cell is tested only when every loop of the
        // scan is non-empty; the θ walk is empty only for an inverted range
        // without a wrap (which the grid never produces), the others whenever
        // their range is inverted.
        const int wd = c.wd;
        const bool has_wrap = lo[wd] > hi[wd];
        const bool empty = (lo[3] > hi[3]) || (lo[2] > hi[2] && !has_wrap) || lo[1] > hi[1] || lo[0] > hi[0];
        if (!empty) {
            const std::uint64_t w = c.n[0] + 1ull;
            const std::uint32_t* F = a.fail.data();
            const std::uint32_t f = F[(hi[1] + 1) * w + hi[0] + 1] - F[(hi[1] + 1) * w + lo[0]] -
                                    F[lo[1] * w + hi[0] + 1] + F[lo[1] * w + lo[0]];
            if (f != 0) return false;
        }
    }
    return true;
}

inline bool state_passes_constraint(const Case& c, const Abstraction& a, std::uint32_t ix, std::uint32_t iy) {
    if (!c.constrained) return true;
    const std::uint64_t w = c.n[0] + 1ull;
    const std::uint32_t* F = a.fail.data();
    return F[(iy + 1) * w + ix + 1] - F[(iy + 1) * w + ix] - F[iy * w + ix + 1] + F[iy * w + ix] == 0;
}

void build_abstraction(const CaseConfig& cfg, const Case& c, unsigned threads, Abstraction& a) {
    const std::uint32_t K = c.K;
    const std::uint32_t n0 = c.n[0], n1 = c.n[1], n2 = c.n[2], n3 = c.n[3];
    const double T = c.T;
    const double half[kStateDim] = {0.5 * c.g.eta[0], 0.5 * c.g.eta[1], 0.5 * c.g.eta[2], 0.5 * c.g.eta[3]};
    double jfw[kStateDim];
    for (int d = 0; d < kStateDim; ++d) jfw[d] = T * cfg.disturbance_half_width[d];  // model.hpp:84-89

    // θ trajectory per (θ cell, ω): cos / sin of θ at each substep, then the
    // box's θ range (model.hpp:35-44, 107-111).
    std::vector<double> tcos(static_cast<std::size_t>(n2) * c.n_w * K), tsin(tcos.size());
    a.th_lo.assign(static_cast<std::size_t>(n2) * c.n_w, kReject);
    a.th_hi.assign(a.th_lo.size(), kReject);
    const double rad2 = half[2] + jfw[2];
    for (std::uint32_t it = 0; it < n2; ++it)
        for (std::uint32_t iw = 0; iw < c.n_w; ++iw) {
            const std::size_t row = static_cast<std::size_t>(it) * c.n_w + iw;
            const double omega = c.input_center(1, iw);
            double th = c.center(2, it);
            for (std::uint32_t k = 0; k < K; ++k) {
                tcos[row * K + k] = std::cos(th);
                tsin[row * K + k] = std::sin(th);
                th = wrap_angle(th + c.dt * omega);
            }
            std::uint32_t lo, hi;
            if (c.cells(2, th - rad2, th + rad2, &lo, &hi)) {
                a.th_lo[row] = lo;
                a.th_hi[row] = hi;
            }
        }
    // v trajectory per (v cell, a), the box's v range, and the x / y radius
    // of the v cell (model.hpp:71-73: |Jf_x| reads the state's own v).
    std::vector<double> vel(static_cast<std::size_t>(n3) * c.n_a * K);
    a.v_lo.assign(static_cast<std::size_t>(n3) * c.n_a, kReject);
    a.v_hi.assign(a.v_lo.size(), kReject);
    std::vector<double> rad0(n3), rad1(n3);
    const double rad3 = half[3] + jfw[3];
    for (std::uint32_t iv = 0; iv < n3; ++iv) {
        const double abs_v = std::fabs(c.center(3, iv));
        rad0[iv] = (half[0] + T * abs_v * half[2] + T * half[3]) + jfw[0];
        rad1[iv] = (half[1] + T * abs_v * half[2] + T * half[3]) + jfw[1];
        for (std::uint32_t ia = 0; ia < c.n_a; ++ia) {
            const std::size_t row = static_cast<std::size_t>(iv) * c.n_a + ia;
            const double acc = c.input_center(0, ia);
            double v = c.center(3, iv);
            for (std::uint32_t k = 0; k < K; ++k) {
                vel[row * K + k] = v;
                v += c.dt * acc;
            }
            std::uint32_t lo, hi;
            if (c.cells(3, v - rad3, v + rad3, &lo, &hi)) {
                a.v_lo[row] = lo;
                a.v_hi[row] = hi;
            }
        }
    }

    // x / y intervals per (plane, input): the increments dt*(v_k cos θ_k) and
    // dt*(v_k sin θ_k) added to every cell centre one substep at a time.
    a.xr.assign(static_cast<std::size_t>(c.planes) * n0 * c.I, kReject);
    a.yr.assign(static_cast<std::size_t>(c.planes) * n1 * c.I, kReject);
    parallel_for(threads, c.planes, 1, [&](std::uint64_t pb, std::uint64_t pe) {
        std::vector<double> incx(K), incy(K), px(n0), py(n1);
        for (std::uint64_t p = pb; p < pe; ++p) {
            const std::uint32_t it = static_cast<std::uint32_t>(p % n2), iv = static_cast<std::uint32_t>(p / n2);
            for (std::uint32_t i = 0; i < c.I; ++i) {
                const std::uint32_t ia = i % c.n_a, iw = i / c.n_a;
                const double* cs = tcos.data() + (static_cast<std::size_t>(it) * c.n_w + iw) * K;
                const double* sn = tsin.data() + (static_cast<std::size_t>(it) * c.n_w + iw) * K;
                const double* vk = vel.data() + (static_cast<std::size_t>(iv) * c.n_a + ia) * K;
                for (std::uint32_t k = 0; k < K; ++k) {
                    incx[k] = c.dt * (vk[k] * cs[k]);
                    incy[k] = c.dt * (vk[k] * sn[k]);
                }
                for (std::uint32_t ix = 0; ix < n0; ++ix) px[ix] = c.center(0, ix);
                for (std::uint32_t iy = 0; iy < n1; ++iy) py[iy] = c.center(1, iy);
                for (std::uint32_t k = 0; k < K; ++k) {
                    const double dx = incx[k], dy = incy[k];
                    for (std::uint32_t ix = 0; ix < n0; ++ix) px[ix] = px[ix] + dx;
                    for (std::uint32_t iy = 0; iy < n1; ++iy) py[iy] = py[iy] + dy;
                }
                std::uint32_t lo, hi;
                for (std::uint32_t ix = 0; ix < n0; ++ix)
                    if (c.cells(0, px[ix] - rad0[iv], px[ix] + rad0[iv], &lo, &hi))
                        a.xr[(p * n0 + ix) * c.I + i] = pack(lo, hi);
                for (std::uint32_t iy = 0; iy < n1; ++iy)
                    if (c.cells(1, py[iy] - rad1[iv], py[iy] + rad1[iv], &lo, &hi))
                        a.yr[(p * n1 + iy) * c.I + i] = pack(lo, hi);
            }
        }
    });

    // Constraint failures over (x, y) (config.cpp:111-127 reads x[0], x[1]).
    if (c.constrained) {
        const std::uint64_t w = n0 + 1ull;
        a.fail.assign(w * (n1 + 1ull), 0);
        for (std::uint32_t iy = 0; iy < n1; ++iy)
            for (std::uint32_t ix = 0; ix < n0; ++ix) {
                double x[kStateDim] = {c.center(0, ix), c.center(1, iy), c.center(2, 0), c.center(3, 0)};
                const std::uint32_t bad = satisfies_state_constraint(cfg, x) ? 0u : 1u;
                a.fail[(iy + 1) * w + ix + 1] =
                    bad + a.fail[(iy + 1) * w + ix] + a.fail[iy * w + ix + 1] - a.fail[iy * w + ix];
            }
    }

    // θ / v hulls per cell: union over the inputs; a wrapped range takes the
    // whole axis.
    auto cell_hulls = [](const std::vector<std::uint32_t>& lo, const std::vector<std::uint32_t>& hi,
                         std::uint32_t cells, std::uint32_t per_cell, std::vector<std::uint32_t>& hlo,
                         std::vector<std::uint32_t>& hhi) {
        hlo.assign(cells, kReject);
        hhi.assign(cells, 0);
        for (std::uint32_t k = 0; k < cells; ++k)
            for (std::uint32_t j = 0; j < per_cell; ++j) {
                const std::uint32_t l = lo[static_cast<std::size_t>(k) * per_cell + j];
                const std::uint32_t h = hi[static_cast<std::size_t>(k) * per_cell + j];
                if (l == kReject) continue;
                if (l > h) {
                    hlo[k] = 0;
                    hhi[k] = cells - 1;
                    break;
                }
                hlo[k] = hlo[k] == kReject ? l : std::min(hlo[k], l);
                hhi[k] = std::max(hhi[k], h);
            }
    };
    cell_hulls(a.th_lo, a.th_hi, n2, c.n_w, a.hull_th_lo, a.hull_th_hi);
    cell_hulls(a.v_lo, a.v_hi, n3, c.n_a, a.hull_v_lo, a.hull_v_hi);

    // Per-state x / y hulls over the valid pairs, and the valid-pair count.
    if (c.frontier) {
        a.hull_x.assign(c.S, kReject);
        a.hull_y.assign(c.S, kReject);
    }
    std::atomic<std::uint64_t> valid_pairs{0};
    const std::uint64_t rows = static_cast<std::uint64_t>(c.planes) * n1;
    parallel_for(threads, rows, 16, [&](std::uint64_t rb, std::uint64_t re) {
        PlaneRanges pr;
        std::uint32_t loaded = kReject;
        std::uint64_t count = 0;
        for (std::uint64_t r = rb; r < re; ++r) {
            const std::uint32_t p = static_cast<std::uint32_t>(r / n1), iy = static_cast<std::uint32_t>(r % n1);
            if (p != loaded) {
                pr.load(c, a, p);
                loaded = p;
            }
            const std::uint64_t yrow = (static_cast<std::uint64_t>(p) * n1 + iy) * c.I;
            for (std::uint32_t ix = 0; ix < n0; ++ix) {
                if (!state_passes_constraint(c, a, ix, iy)) continue;  // abstraction.hpp:177-181
                const std::uint64_t xrow = (static_cast<std::uint64_t>(p) * n0 + ix) * c.I;
                std::uint32_t xl = kReject, xh = 0, yl = kReject, yh = 0;
                std::uint32_t lo[4], hi[4];
                for (std::uint32_t i = 0; i < c.I; ++i) {
                    if (!pair_cells(c, a, pr, xrow, yrow, i, lo, hi)) continue;
                    ++count;
                    xl = std::min(xl, lo[0]);
                    xh = std::max(xh, hi[0]);
                    yl = std::min(yl, lo[1]);
                    yh = std::max(yh, hi[1]);
                }
                if (c.frontier && xl != kReject) {
                    const std::uint64_t s = static_cast<std::uint64_t>(r) * n0 + ix;
                    a.hull_x[s] = pack(xl, xh);
                    a.hull_y[s] = pack(yl, yh);
                }
            }
        }
        valid_pairs.fetch_add(count, std::memory_order_relaxed);
    });
    a.valid_pairs = valid_pairs.load();
}

// 4D inclusive prefix sums of valid ∧ ¬reached (reached == nullptr: of valid)
// into P, whose zero borders are never written (prefix_sum.hpp:15-63): a row
// pass along x, then one running-sum pass per further axis.
void build_prefix(const Case& c, const std::vector<std::uint8_t>& valid, const std::uint8_t* reached,
                  std::vector<std::uint32_t>& P, unsigned threads) {
    if (P.size() != c.PS) P.assign(c.PS, 0);
    const std::uint32_t n0 = c.n[0], n1 = c.n[1], n2 = c.n[2], n3 = c.n[3];
    const std::uint64_t* ps = c.ps;
    std::uint32_t* D = P.data();
    parallel_for(threads, static_cast<std::uint64_t>(n1) * n2 * n3, 64, [&](std::uint64_t b, std::uint64_t e) {
        for (std::uint64_t line = b; line < e; ++line) {
            const std::uint64_t y = line % n1, z = (line / n1) % n2, w = line / (static_cast<std::uint64_t>(n1) * n2);
            const std::uint64_t src = line * n0;
            std::uint32_t* row = D + (y + 1) * ps[1] + (z + 1) * ps[2] + (w + 1) * ps[3] + 1;
            std::uint32_t run = 0;
            for (std::uint32_t x = 0; x < n0; ++x) {
                run += (valid[src + x] && !(reached && reached[src + x])) ? 1u : 0u;
                row[x] = run;
            }
        }
    });
    // y: per (z, w) plane, each row adds the one before.
    parallel_for(threads, static_cast<std::uint64_t>(n2) * n3, 1, [&](std::uint64_t b, std::uint64_t e) {
        for (std::uint64_t zw = b; zw < e; ++zw) {
            const std::uint64_t z = zw % n2, w = zw / n2;
            std::uint32_t* base = D + (z + 1) * ps[2] + (w + 1) * ps[3] + 1;
            for (std::uint32_t y = 2; y <= n1; ++y) {
                std::uint32_t* row = base + y * ps[1];
                const std::uint32_t* prev = row - ps[1];
                for (std::uint32_t x = 0; x < n0; ++x) row[x] += prev[x];
            }
        }
    });
    // z: per (y, w), along z.
    parallel_for(threads, static_cast<std::uint64_t>(n1) * n3, 16, [&](std::uint64_t b, std::uint64_t e) {
        for (std::uint64_t yw = b; yw < e; ++yw) {
            const std::uint64_t y = yw % n1, w = yw / n1;
            std::uint32_t* base = D + (y + 1) * ps[1] + (w + 1) * ps[3] + 1;
            for (std::uint32_t z = 2; z <= n2; ++z) {
                std::uint32_t* row = base + z * ps[2];
                const std::uint32_t* prev = row - ps[2];
                for (std::uint32_t x = 0; x < n0; ++x) row[x] += prev[x];
            }
        }
    });
    // w: per (y, z), along w.
    parallel_for(threads, static_cast<std::uint64_t>(n1) * n2, 16, [&](std::uint64_t b, std::uint64_t e) {
        for (std::uint64_t yz = b; yz < e; ++yz) {
            const std::uint64_t y = yz % n1, z = yz / n1;
            std::uint32_t* base = D + (y + 1) * ps[1] + (z + 1) * ps[2] + 1;
            for (std::uint32_t w = 2; w <= n3; ++w) {
                std::uint32_t* row = base + w * ps[3];
                const std::uint32_t* prev = row - ps[3];
                for (std::uint32_t x = 0; x < n0; ++x) row[x] += prev[x];
            }
        }
    });
}

// Cells counted in the box [lo, hi] (inclusive, lo <= hi on every axis):
// query_box_single (prefix_sum.hpp:66-89), modulo 2^32.
inline std::uint32_t box_sum(const std::uint32_t* P, const std::uint64_t ps[4], const std::uint32_t lo[4],
                             const std::uint32_t hi[4]) {
    const std::uint64_t a[4] = {lo[0] * ps[0], lo[1] * ps[1], lo[2] * ps[2], lo[3] * ps[3]};
    const std::uint64_t b[4] = {(hi[0] + 1ull) * ps[0], (hi[1] + 1ull) * ps[1], (hi[2] + 1ull) * ps[2],
                                (hi[3] + 1ull) * ps[3]};
    std::uint32_t pos = 0, neg = 0;
    for (int m = 0; m < 16; ++m) {
        const std::uint64_t off = ((m & 1) ? a[0] : b[0]) + ((m & 2) ? a[1] : b[1]) + ((m & 4) ? a[2] : b[2]) +
                                  ((m & 8) ? a[3] : b[3]);
        if (__builtin_popcount(static_cast<unsigned>(m)) & 1)
            neg += P[off];
        else
            pos += P[off];
    }
    return pos - neg;
}

// query_box_count (prefix_sum.hpp:94-112): a range inverted along the wrap
// axis is split in two.
inline std::uint32_t box_count(const Case& c, const std::uint32_t* P, const std::uint32_t lo[4],
                               const std::uint32_t hi[4]) {
    const int wd = c.wd;
    if (lo[wd] <= hi[wd]) return box_sum(P, c.ps, lo, hi);
    std::uint32_t l2[4] = {lo[0], lo[1], lo[2], lo[3]}, h2[4] = {hi[0], hi[1], hi[2], hi[3]};
    h2[wd] = c.n[wd] - 1;
    const std::uint32_t left = box_sum(P, c.ps, l2, h2);
    l2[wd] = 0;
    h2[wd] = hi[wd];
    return left + box_sum(P, c.ps, l2, h2);
}

// Row probes: one bit per state (flat order), 1 for a valid, unreached state,
// and the same ORed with each state's θ + 1, v + 1 and θ + 1 / v + 1 neighbours
// (the next θ (v) cell wrapping to the first when θ (v) is the wrap axis, none
// past the last otherwise). A set bit in a box's x-range on one of its rows
// proves the box holds a valid unreached cell; the OR bitmask is read for boxes
// that span two θ and two v cells, where its four slices lie inside the box.
struct ProbeBits {
    std::vector<std::uint64_t> plain, ored;  // (S + 63) / 64 + 1 words each
};

void build_probe_bits(const Case& c, const std::vector<std::uint8_t>& valid, const std::vector<std::uint8_t>& reach,
                      ProbeBits& pb, unsigned threads) {
    const std::uint64_t words = (c.S + 63) / 64 + 1;
    pb.plain.assign(words, 0);
    pb.ored.assign(words, 0);
    parallel_for(threads, c.S / 64 + 1, 256, [&](std::uint64_t b, std::uint64_t e) {
        for (std::uint64_t w = b; w < e; ++w) {
            std::uint64_t m = 0;
            for (std::uint64_t s = w * 64, k = 0; k < 64 && s < c.S; ++k, ++s)
                m |= static_cast<std::uint64_t>(valid[s] && !reach[s]) << k;
            pb.plain[w] = m;
        }
    });
    const std::uint64_t st2 = c.PQ, st3 = static_cast<std::uint64_t>(c.PQ) * c.n[2];
    auto bit = [&](std::uint64_t s) { return (pb.plain[s >> 6] >> (s & 63)) & 1u; };
    parallel_for(threads, c.S / 64 + 1, 256, [&](std::uint64_t b, std::uint64_t e) {
        for (std::uint64_t w = b; w < e; ++w) {
            std::uint64_t m = 0;
            for (std::uint64_t s = w * 64, k = 0; k < 64 && s < c.S; ++k, ++s) {
                const std::uint64_t plane = s / st2;
                const std::uint32_t z = static_cast<std::uint32_t>(plane % c.n[2]);
                const std::uint32_t v = static_cast<std::uint32_t>(plane / c.n[2]);
                const bool zl = z + 1 < c.n[2], vl = v + 1 < c.n[3];
                const bool has_z = zl || c.wd == 2, has_v = vl || c.wd == 3;
                const std::uint64_t sz = zl ? s + st2 : s - z * st2;
                const std::uint64_t sv = vl ? s + st3 : s - v * st3;
                const std::uint64_t szv = vl ? sz + st3 : sz - v * st3;
                std::uint64_t u = bit(s);
                if (has_z) u |= bit(sz);
                if (has_v) u |= bit(sv);
                if (has_z && has_v) u |= bit(szv);
                m |= u << k;
            }
            pb.ored[w] = m;
        }
    });
}

// Any set bit in [pos, pos + ext] (ext < 64).
inline bool bits_any(const std::vector<std::uint64_t>& b, std::uint64_t pos, std::uint32_t ext) {
    const std::uint64_t w = pos >> 6, sh = pos & 63;
    std::uint64_t v = b[w] >> sh;
    if (sh) v |= b[w + 1] << (64 - sh);
    const std::uint64_t mask = ext >= 63 ? ~0ull : ((1ull << (ext + 1)) - 1);
    return (v & mask) != 0;
}

// A valid unreached cell seen on the box's middle, first or last y row at its
// lower θ and v (the four θ / v corner slices when the box spans them).
inline bool probe_unsatisfied(const Case& c, const ProbeBits& pb, const std::uint32_t lo[4],
                              const std::uint32_t hi[4]) {
    if (lo[0] > hi[0] || lo[1] > hi[1] || hi[0] - lo[0] >= 64) return false;
    const bool span = lo[2] != hi[2] && lo[3] != hi[3];
    const std::vector<std::uint64_t>& b = span ? pb.ored : pb.plain;
    const std::uint64_t base = lo[0] + static_cast<std::uint64_t>(c.PQ) * (lo[2] + static_cast<std::uint64_t>(c.n[2]) * lo[3]);
    const std::uint32_t ext = hi[0] - lo[0];
    const std::uint32_t ym = lo[1] + ((hi[1] - lo[1]) >> 1);
    return bits_any(b, base + static_cast<std::uint64_t>(ym) * c.n[0], ext) ||
           bits_any(b, base + static_cast<std::uint64_t>(lo[1]) * c.n[0], ext) ||
           bits_any(b, base + static_cast<std::uint64_t>(hi[1]) * c.n[0], ext);
}

}  // namespace

unsigned default_cpu_threads() {
    for (const char* var : {"GSC_CPU_THREADS", "OMP_NUM_THREADS"}) {
        const char* v = std::getenv(var);
        if (v && *v) {
            const long n = std::strtol(v, nullptr, 10);
            if (n > 0) return static_cast<unsigned>(std::min(n, 1024L));
        }
    }
    return std::max(1u, std::thread::hardware_concurrency());
}

CpuRunReport run_case_cpu_u32(const CaseConfig& cfg, const CaseExtensions& ext, unsigned threads) {
    if (threads == 0) threads = default_cpu_threads();
    if (cfg.constraint_type == ConstraintType::kCustom)
        throw std::runtime_error("the CPU solver takes hyperbolic and elliptic constraints only");
    Case c;
    build_grids(cfg, c);
    if (c.n[0] > kMaxXYCells || c.n[1] > kMaxXYCells)
        throw std::runtime_error("the CPU solver supports at most 65535 cells along x and y");

    CpuRunReport rep;
    rep.threads = threads;
    rep.state_grid = c.g;
    rep.input_grid = c.in;
    rep.pair_count = c.S * c.I;
    rep.budget = reference_budget(c);
    auto& res = rep.result;

    // abstraction.hpp:287-306, with its progress lines.
    std::cout << "CPU: Building valid mask for " << c.S << " states..." << std::endl;
    std::cout << "CPU: Building candidate mask for " << c.S << " states..." << std::endl;
    build_masks(cfg, ext, c, threads, res.valid_mask, res.candidate_mask);
    for (std::uint64_t s = 0; s < c.S; ++s)
        if (res.candidate_mask[s] && !res.valid_mask[s])
            throw std::runtime_error("candidate set contains invalid states");

    std::cout << "CPU: Starting abstraction phase for " << c.S << " states and " << c.I << " inputs..." << std::endl;
    const auto ta = Clock::now();
    Abstraction a;
    build_abstraction(cfg, c, threads, a);
    rep.abstraction_ms = ms_since(ta);
    rep.valid_pairs = a.valid_pairs;
    std::cout << "CPU: Abstraction phase completed in " << rep.abstraction_ms << " ms" << std::endl;
    std::cout << "CPU: Valid pairs: " << rep.pair_count << std::endl;

    // reachability.hpp:35-155
    const std::vector<std::uint8_t>& valid = res.valid_mask;
    const std::vector<std::uint8_t>& cand = res.candidate_mask;
    std::vector<std::uint8_t>& reach = res.reachable_mask;
    reach = cand;
    res.controller.assign(c.S, kInvalidInput);
    res.reach_step.assign(c.S, kUnreachableStep);
    std::uint64_t candidate_count = 0;
    for (std::uint64_t s = 0; s < c.S; ++s)
        if (cand[s]) {
            res.reach_step[s] = 0;
            ++candidate_count;
        }

    std::vector<std::uint32_t> pv, pu[2];
    ProbeBits pb;
    build_prefix(c, valid, nullptr, pv, threads);
    std::uint64_t certified = 0;
    const auto t0 = Clock::now();
    const std::uint32_t n0 = c.n[0], n1 = c.n[1];
    const std::uint64_t rows = static_cast<std::uint64_t>(c.planes) * n1;

    for (std::uint32_t iter = 1; iter <= cfg.max_iterations; ++iter) {
        std::vector<std::uint32_t>& cur = pu[iter & 1];
        build_prefix(c, valid, reach.data(), cur, threads);
        build_probe_bits(c, valid, reach, pb, threads);
        const std::uint32_t* U = cur.data();
        const std::uint32_t* Uprev = iter == 1 ? pv.data() : pu[(iter + 1) & 1].data();
        const std::uint32_t* V = pv.data();
        std::atomic<std::uint64_t> new_reach_total{0}, new_cert_total{0};

        parallel_for(threads, rows, 8, [&](std::uint64_t rb, std::uint64_t re) {
            PlaneRanges pr;
            std::uint32_t loaded = kReject;
            std::uint64_t new_reach = 0, new_cert = 0;
            std::uint32_t lo[4], hi[4];
            for (std::uint64_t r = rb; r < re; ++r) {
                const std::uint32_t p = static_cast<std::uint32_t>(r / n1), iy = static_cast<std::uint32_t>(r % n1);
                const std::uint64_t yrow = (static_cast<std::uint64_t>(p) * n1 + iy) * c.I;
                for (std::uint32_t ix = 0; ix < n0; ++ix) {
                    const std::uint64_t s = r * n0 + ix;
                    if (!valid[s]) continue;
                    const bool is_cand = cand[s] != 0;
                    const bool already_reached = reach[s] != 0;
                    const bool already_certified = res.controller[s] != kInvalidInput;
                    if ((!is_cand && already_reached) || (is_cand && already_certified)) continue;
                    if (c.frontier) {
                        const std::uint32_t hx = a.hull_x[s];
                        if (hx == kReject) continue;  // no valid pair: never satisfied
                        const std::uint32_t hy = a.hull_y[s];
                        const std::uint32_t it = p % c.n[2], iv = p / c.n[2];
                        const std::uint32_t hl[4] = {lo_of(hx), lo_of(hy), a.hull_th_lo[it], a.hull_v_lo[iv]};
                        const std::uint32_t hh[4] = {hi_of(hx), hi_of(hy), a.hull_th_hi[it], a.hull_v_hi[iv]};
                        if (box_sum(U, c.ps, hl, hh) == box_sum(Uprev, c.ps, hl, hh)) continue;
                    }
                    if (p != loaded) {
                        pr.load(c, a, p);
                        loaded = p;
                    }
                    if (!state_passes_constraint(c, a, ix, iy)) continue;
                    const std::uint64_t xrow = (static_cast<std::uint64_t>(p) * n0 + ix) * c.I;
                    for (std::uint32_t i = 0; i < c.I; ++i) {
                        if (!pair_cells(c, a, pr, xrow, yrow, i, lo, hi)) continue;
                        if (probe_unsatisfied(c, pb, lo, hi)) continue;
                        if (box_count(c, U, lo, hi) != 0) continue;
                        if (box_count(c, V, lo, hi) == 0) continue;
                        // reachability.hpp:89-104
                        if (!already_reached) {
                            reach[s] = 1;
                            ++new_reach;
                        }
                        if (!already_certified) {
                            res.controller[s] = i;
                            if (is_cand)
                                ++new_cert;
                            else
                                res.reach_step[s] = iter;
                        }
                        break;
                    }
                }
            }
            new_reach_total.fetch_add(new_reach, std::memory_order_relaxed);
            new_cert_total.fetch_add(new_cert, std::memory_order_relaxed);
        });

        const std::uint64_t new_reachable = new_reach_total.load();
        const std::uint64_t new_candidate_certified = new_cert_total.load();
        certified += new_candidate_certified;
        res.iterations = iter;
        if (cfg.verbose) {
            // reachability.hpp:113-128: a candidate has a controller exactly
            // when it has been certified.
            const double coverage = candidate_count == 0 ? 100.0
                                                         : 100.0 * static_cast<double>(certified) /
                                                               static_cast<double>(candidate_count);
            std::printf("iter %u: new_reachable=%llu candidate_controller_coverage=%.1f%% (%llu/%llu)\n", iter,
                        static_cast<unsigned long long>(new_reachable), coverage,
                        static_cast<unsigned long long>(certified), static_cast<unsigned long long>(candidate_count));
        }
        if (certified == candidate_count) {
            res.converged = true;
            res.message = "all candidate states obtained a controller";
            break;
        }
        if (new_reachable == 0 && new_candidate_certified == 0) {
            res.message = "iteration stalled before candidate closure";
            break;
        }
    }

    res.solve_ms = ms_since(t0);
    res.reachable_states = static_cast<std::uint64_t>(std::count(reach.begin(), reach.end(), std::uint8_t{1}));
    res.certified_candidate_states = certified;
    if (!res.converged && res.message.empty()) res.message = "maximum iterations reached";
    std::fflush(stdout);
    return rep;
}

}  // namespace gsc
