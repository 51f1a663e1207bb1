// gsc_device.hpp — data layout and FP64 cell-box arithmetic shared by the
// host (table building) and the gfx950 kernels.
//
// Every function here that turns FP64 coordinates into integer cells must give
// the same bits on the host and on the device. Both sides are compiled with
// -ffp-contract=off (no FMA contraction), keep the reference's expression
// order (grid.hpp:42, 76, 181-251; model.hpp:34-45, 59-75, 107-111) and use
// only IEEE-exact operations (+, -, *, correctly rounded /, floor, ceil).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "gsc_amd.h"

namespace gsc_amd {

#define GSC_HD __host__ __device__

constexpr double kEps = 1e-9;  // grid.hpp:186

// ---- pair record (4 bytes per (state, input) pair) ------------------------
// The reference stores min_flat u32 + max_flat u32 + valid u8 per pair
// (abstraction.hpp:32-38). Here the successor box is split: its θ and v ranges
// depend only on (θ cell, ω input) and (v cell, a input) — the θ and v
// trajectories of model.hpp:26-46 do not involve x or y — so they live in two
// small tables (box2, box3). The per-pair word keeps the x and y ranges as
// the cell where the box's middle x-row starts, relative to the state's own
// cell q = ix + iy * n_x, and the two extents:
//   bits  0-19  rel = (lo0 - ix) + (lo1 + ext1 / 2 - iy) * n_x, two's
//               complement (-2^19 .. 2^19 - 1). q + rel is the middle row's
//               first cell in the plane; lo0 = (q + rel) mod n_x and
//               lo1 = (q + rel) / n_x - ext1 / 2, unique because lo0 lies in
//               [0, n_x). The satisfaction probe reads that row as is.
//   bits 20-24  31 - x extent (box.max[0] - box.min[0], 0..31): the probe's
//               shift amount
//   bits 25-29  y extent
//   bit  30     dead: the box holds no valid cell, so the pair can never be
//               satisfied (the reference's `valid_count == 0` test,
//               reachability.hpp:78-85 / kernels.cu:327-332). The valid set
//               never changes, so the answer is memoised: the abstraction
//               writes 0 ("not known dead") and the satisfaction pass sets the
//               bit the first time it counts zero valid cells in the box.
//   bit  31     escape: bits 0-30 index an overflow entry holding the full
//               (min_flat, max_flat) pair; used for boxes that do not fit the
//               fields above (wrapped x/y ranges, large boxes, inverted ranges)
// 0xFFFFFFFF marks an invalid pair (valid = 0 in the reference).
constexpr uint32_t kRecInvalid = 0xFFFFFFFFu;
constexpr uint32_t kRecEscape = 0x80000000u;
constexpr uint32_t kRecDead = 0x40000000u;
constexpr uint32_t kRecIndexMask = 0x7FFFFFFFu;
constexpr int kRelMin = -(1 << 19);
constexpr int kRelMax = (1 << 19) - 1;
constexpr uint32_t kExtMax = 31;
// box2/box3 entry whose range is out of the state domain (pair invalid).
constexpr uint32_t kBoxReject = 0xFFFFFFFFu;

// The middle-row offset of a box with x / y offsets off0, off1 (lo - cell) and
// y extent ext1, on a grid n_x cells wide.
GSC_HD inline int64_t rec_rel_of(int off0, int off1, uint32_t ext1, uint32_t n0) {
    return static_cast<int64_t>(off0) + (static_cast<int64_t>(off1) + static_cast<int64_t>(ext1 >> 1)) * n0;
}
GSC_HD inline uint32_t rec_pack(int rel, uint32_t ext0, uint32_t ext1, bool dead) {
    return (static_cast<uint32_t>(rel) & 0xFFFFFu) | ((31u - ext0) << 20) | (ext1 << 25) | (dead ? kRecDead : 0u);
}
GSC_HD inline int rec_rel(uint32_t r) { return static_cast<int32_t>(r << 12) >> 12; }
GSC_HD inline uint32_t rec_shift0(uint32_t r) { return (r >> 20) & 31u; }  // 31 - x extent
GSC_HD inline uint32_t rec_ext0(uint32_t r) { return 31u - rec_shift0(r); }
GSC_HD inline uint32_t rec_ext1(uint32_t r) { return (r >> 25) & 31u; }
// The box's lower x and y cells, for the state at cell q of its plane.
GSC_HD inline void rec_lo(uint32_t r, uint32_t q, uint32_t n0, uint32_t* lo0, uint32_t* lo1) {
    const uint32_t t = q + static_cast<uint32_t>(rec_rel(r));  // the middle row's first cell, in [0, PQ)
    const uint32_t ym = t / n0;
    *lo0 = t - ym * n0;
    *lo1 = ym - (rec_ext1(r) >> 1);
}

// Record placement. Within a (θ, v) plane (plane base = plane * I_pad * PQ,
// PQ = cells per plane) the records come in blocks of kRecBlock consecutive
// inputs: input i of cell q sits at (i / kRecBlock * PQ + q) * kRecBlock +
// i % kRecBlock. One state's block is one 32-byte sector, loaded with two
// 16-byte vector loads, and neighbouring cells' blocks are adjacent, so the
// satisfaction pass fetches whole sectors of the states it tests however
// sparse they are.
constexpr uint32_t kRecBlock = 8;
GSC_HD inline uint32_t rec_offset(uint32_t PQ, uint32_t q, uint32_t i) {
    return ((i / kRecBlock) * PQ + q) * kRecBlock + i % kRecBlock;
}

// ---- state hull (4 bytes per state) ----------------------------------------
// The union of the x / y ranges of a state's valid successor boxes, relative to
// the state's own cell: int8 dx_lo, dx_hi, dy_lo, dy_hi in bytes 0..3. Together
// with the per-θ-cell and per-v-cell unions of the θ / v ranges it bounds every
// box the state can be tested against, so a state whose hull saw no newly
// reached cell since its last test gets the same (unsatisfied) answer and can
// be skipped. Two sentinels:
constexpr uint32_t kHullNone = 0x807F807Fu;  // no valid pair: never satisfiable
constexpr uint32_t kHullAll = 0x7F807F80u;   // not representable (escaped box): always test
constexpr int kHullMaxOffset = 100;

// Chunk hull (uint4 x0, x1, y0, y1 in absolute cells) of 64 consecutive cells
// of a plane: the union of their state hulls, or one of two sentinels.
constexpr uint32_t kChunkAll = 0xFFFFFFFFu;   // x0 == kChunkAll: some cell always tests
constexpr uint32_t kChunkNone = 0xFFFFFFFEu;  // x0 == kChunkNone: no cell can be satisfied

GSC_HD inline uint32_t hull_pack(int dxlo, int dxhi, int dylo, int dyhi) {
    return (static_cast<uint32_t>(dxlo) & 0xFFu) | ((static_cast<uint32_t>(dxhi) & 0xFFu) << 8) |
           ((static_cast<uint32_t>(dylo) & 0xFFu) << 16) | ((static_cast<uint32_t>(dyhi) & 0xFFu) << 24);
}
GSC_HD inline int hull_byte(uint32_t h, int k) { return static_cast<int8_t>((h >> (8 * k)) & 0xFFu); }

// Overflow entry for escaped pairs.
struct OverflowEntry {
    uint32_t min_flat;
    uint32_t max_flat;
    uint32_t dead;
    uint32_t reserved;
};

// ---- case parameters passed by value to every kernel -----------------------
struct DevCase {
    double lb[4], ub[4], eta[4];
    double lb_minus_eps[4];  // lb[d] - kEps   (grid.hpp:193)
    // Filtered rounding of the nominal x / y (dims 0, 1; set_cell_filter,
    // dim_to_cells_fast), in quotient units (coordinate - lb) / eta: the domain
    // limits lb - kEps, ub + kEps moved out and in by the error bound, the
    // bound itself and 1 - bound, and 1 / eta.
    double flt_rej_lo[2], flt_in_lo[2], flt_in_hi[2], flt_rej_hi[2];
    double flt_q[2], flt_q1[2];
    double inv_eta[2];
    double ub_plus_eps[4];   // ub[d] + kEps
    double wrap_span_minus_eps;  // (ub[w] - lb[w]) - kEps (grid.hpp:198-207)
    uint32_t shape[4];
    uint32_t stride[4];   // state strides, x fastest (grid.hpp:126-129)
    uint32_t pstride[4];  // padded prefix strides (grid.hpp:253-269)
    uint32_t S;           // states
    uint32_t I;           // inputs
    uint32_t I_pad;       // inputs rounded up to a whole number of satisfaction blocks;
                          // records of the padding inputs are invalid
    uint32_t PQ;          // cells per (θ, v) plane: n_x * n_y
    uint32_t chunks_per_plane;  // ceil(PQ / 64): 64-cell wave tasks per plane
    uint32_t n_tasks;     // planes * chunks_per_plane
    uint32_t groups_per_plane;  // ceil(PQ / 256): satisfaction workgroups per plane
    uint32_t n_groups;    // planes * groups_per_plane
    uint32_t n_a;         // input shape[0]: a (dv = u[0], model.hpp:39)
    uint32_t n_w;         // input shape[1]: ω (dθ = u[1], model.hpp:38)
    uint32_t K;           // integration substeps
    int32_t wrap_dim;
    double dt;            // sample_time / substeps (model.hpp:34)
    int32_t constraint_type;
    double hyp_a, hyp_b, hyp_c, ell_a, ell_b;
    gsc_rect4 candidate, map;
    int32_t band_enabled;
    double band_c, band_w;
    uint32_t n_obstacles, n_targets;
    // Row pitch of the 2D constraint-failure prefix table: 1 << fail_shift
    // entries, the least power of two >= n_x + 1, so a row offset is a shift.
    uint32_t fail_shift;
    int32_t hull_enabled;  // frontier skipping available (wrap along θ or v only)
    // Coarse prefix tables: the full tables sampled at x, y multiples of
    // kCoarse (and at n_x, n_y), θ and v at full resolution.
    uint32_t cshape[2];    // coarse x / y positions: ceil(n / kCoarse) + 1
    uint32_t cstride[4];
};

constexpr uint32_t kCoarse = 4;

GSC_HD inline double cell_center(double lb, uint32_t idx, double eta) {
    // grid.hpp:42 / :76
    return lb + (static_cast<double>(idx) + 0.5) * eta;
}

// config.cpp:64-75, 111-127 (the constraint reads x[0] and x[1] only).
GSC_HD inline bool constraint_ok(int type, double ha, double hb, double hc, double ea, double eb, double x0,
                                 double x1) {
    if (type == GSC_CONSTRAINT_HYPERBOLIC) {
        const double x1_sq = x0 * x0;
        const double x2_sq = x1 * x1;
        return (x1_sq - x2_sq <= ha) && (hb * x2_sq - x1_sq <= hc);
    }
    if (type == GSC_CONSTRAINT_ELLIPTIC) {
        const double t1 = (x0 / ea) * (x0 / ea);
        const double t2 = (x1 / eb) * (x1 / eb);
        return t1 + t2 <= 1.0;
    }
    return true;
}

GSC_HD inline bool point_in_rect(const gsc_rect4& r, const double x[4]) {
    // config.cpp:79-86
    for (int d = 0; d < 4; ++d)
        if (x[d] < r.lb[d] || x[d] > r.ub[d]) return false;
    return true;
}

// Unicycle4DModel::wrap_angle, model.hpp:14-24 (literal π, not the grid bounds).
GSC_HD inline double wrap_angle(double x) {
    constexpr double kPi = 3.14159265358979323846;
    constexpr double kTwoPi = 2.0 * kPi;
    while (x < -kPi) x += kTwoPi;
    while (x >= kPi) x -= kTwoPi;
    return x;
}

// Grid4D::wrap_value, grid.hpp:46-57.
GSC_HD inline double wrap_value(double low, double high, double x) {
    const double span = high - low;
    while (x < low) x += span;
    while (x >= high) x -= span;
    return x;
}

// One dimension of continuous_box_to_index_box (grid.hpp:181-251).
// Returns false when a non-wrap dimension leaves [lb - eps, ub + eps].
GSC_HD inline bool dim_to_cells(const DevCase& c, int d, double mn, double mx, uint32_t* lo, uint32_t* hi) {
    double a = mn, b = mx;
    if (d == c.wrap_dim) {
        const double raw_width = mx - mn;
        if (raw_width >= c.wrap_span_minus_eps) {
            *lo = 0;
            *hi = c.shape[d] - 1;
            return true;
        }
        a = wrap_value(c.lb[d], c.ub[d], mn);
        b = wrap_value(c.lb[d], c.ub[d], mx);
    } else {
        if (mn < c.lb_minus_eps[d] || mx > c.ub_plus_eps[d]) return false;
    }
    int64_t lower = static_cast<int64_t>(floor((a - c.lb[d]) / c.eta[d]));
    int64_t upper = static_cast<int64_t>(ceil((b - c.lb[d]) / c.eta[d])) - 1;
    const int64_t n = static_cast<int64_t>(c.shape[d]);
    if (lower < 0) lower = 0;
    if (upper < 0) upper = 0;
    if (lower >= n) lower = n - 1;
    if (upper >= n) upper = n - 1;
    *lo = static_cast<uint32_t>(lower);
    *hi = static_cast<uint32_t>(upper);
    return true;
}

// ---- filtered rounding of the nominal x / y coordinate ----------------------
// The reference adds the K trajectory increments to the cell centre one at a
// time, rounding after every step (x <- x + inc_k, model.hpp:40-41), takes the
// box edges p -/+ rad and rounds (edge - lb) / eta to cells (grid.hpp:181-251).
// The sum S of the increments and the radius are the same for every cell of a
// (θ, v) plane and input, so the kernels form the edge quotients as
//   v0 + (S -/+ rad) / eta,   v0 = (x0 - lb) / eta per cell,
// with the second term from a per-(plane, input) table. With u = 2^-53 and X
// bounding |x0| + sum|inc_k| + rad, the reference's sequential sum and S differ
// from the exact sum by at most uK(|x0| + A) and uKA, and every other rounding
// on either side (the reference's edge, subtraction and division; here the
// reciprocal, two products and the sum) is a few u of |lb| + X over eta or of
// the quotient, so
//   |q_ref - q| <= ((2K + 10) X + 6|lb|) u / eta + 4u (qmax + 1)
// where qmax bounds the quotients inside the domain. set_cell_filter widens
// that by kFilterWiden; dim_to_cells_fast returns the reference's answer only
// when every threshold — the domain limits lb - kEps / ub + kEps in quotient
// units, the floor of the lower quotient, the ceil of the upper one — is clear
// of the bound, and kCellsUndecided otherwise, for the caller to run the exact
// step-by-step sum. Non-wrap dimensions only.
constexpr double kUnitRoundoff = 1.1102230246251565e-16;  // 2^-53
constexpr double kFilterWiden = 1024.0;
enum : int { kCellsUndecided = 0, kCellsReject = 1, kCellsOk = 2 };

// Constants of dimension d (0 or 1) for cell centres |x0| <= x_abs, K-step
// increment sums with sum |inc_k| <= abs_sum, and radii |rad| <= rad_max.
// Needs lb, eta, shape, lb_minus_eps and ub_plus_eps of dimension d.
inline void set_cell_filter(DevCase& c, int d, uint32_t K, double x_abs, double abs_sum, double rad_max) {
    const double u = kUnitRoundoff;
    const double X = x_abs + abs_sum + rad_max;
    const double ie = 1.0 / c.eta[d];
    const double aie = fabs(ie);
    const double qmax = static_cast<double>(c.shape[d]) + 3.0 + 2.0 * kEps * aie;
    const double eq = kFilterWiden * (((2.0 * K + 10.0) * X + 6.0 * fabs(c.lb[d])) * u * aie + 4.0 * u * (qmax + 1.0));
    const double t_lo = (c.lb_minus_eps[d] - c.lb[d]) * ie;
    const double t_hi = (c.ub_plus_eps[d] - c.lb[d]) * ie;
    c.inv_eta[d] = ie;
    c.flt_q[d] = eq;
    c.flt_q1[d] = 1.0 - eq;
    c.flt_rej_lo[d] = t_lo - eq;
    c.flt_in_lo[d] = t_lo + eq;
    c.flt_in_hi[d] = t_hi - eq;
    c.flt_rej_hi[d] = t_hi + eq;
}

// v0 = (x0 - lb) * inv_eta of the cell; qa, qb = (S - rad) * inv_eta and
// (S + rad) * inv_eta of the (plane, input).
GSC_HD inline int dim_to_cells_fast(const DevCase& c, int d, double v0, double qa, double qb, uint32_t* lo,
                                    uint32_t* hi) {
    const double ql = v0 + qa, qh = v0 + qb;
    const bool reject = (ql < c.flt_rej_lo[d]) | (qh > c.flt_rej_hi[d]);
    const bool inside = (ql > c.flt_in_lo[d]) & (qh < c.flt_in_hi[d]);
    const double fl = floor(ql), ce = ceil(qh);
    const double f = ql - fl, g = ce - qh;  // distances to the thresholds, in [0, 1)
    const bool clear = (f > c.flt_q[d]) & (f < c.flt_q1[d]) & (g > c.flt_q[d]) & (g < c.flt_q1[d]);
    // dim_to_cells' clamps, on the (integral) doubles.
    const double top = static_cast<double>(c.shape[d]) - 1.0;
    *lo = static_cast<uint32_t>(fmin(fmax(fl, 0.0), top));
    *hi = static_cast<uint32_t>(fmin(fmax(ce - 1.0, 0.0), top));
    return reject ? kCellsReject : (inside & clear) ? kCellsOk : kCellsUndecided;
}

// ---- per-(plane, input) cell offsets ---------------------------------------
// For a cell centre, v0 = (x0 - lb) / eta is i + 0.5 up to a few roundings,
// so the lower quotient is i + (0.5 + qa) and the upper i + (0.5 + qb) for
// every cell i of the plane. When the fractional parts of ua = 0.5 + qa, of
// ub = 0.5 + qb and of the domain limits relative to them are clear of the
// filter bound, the reference's cells are i + floor(ua) and i + ceil(ub) - 1
// (clamped to the grid as dim_to_cells clamps them) and the box leaves the
// domain exactly for i outside [rej_lo, rej_hi]: the cell's own rounding
// (v0 against i + 0.5, and the sums ua, ub, t - ua) adds at most 4u qmax +
// X u / eta + 2u (qmax + |q|) to the bound above, less than twice it, and the
// bound used is kFilterWiden times it. cell_offsets returns false when any
// fractional part is too close to an integer; the kernels then take
// dim_to_cells_fast for that (plane, input).
struct CellOffsets {
    int32_t off_lo, off_hi;  // cell offsets of the lower and upper edge
    int32_t rej_lo, rej_hi;  // cells i < rej_lo or i > rej_hi leave the domain
};
constexpr int32_t kOffsetsUndecided = INT32_MIN;  // off_lo of a (plane, input) left to dim_to_cells_fast

GSC_HD inline bool cell_offsets(const DevCase& c, int d, double qa, double qb, CellOffsets* o) {
    const double eq = c.flt_q[d];
    const double ua = 0.5 + qa, ub = 0.5 + qb;
    // Quotients of the domain limits lb - kEps / ub + kEps, as set_cell_filter forms them.
    const double t_lo = (c.lb_minus_eps[d] - c.lb[d]) * c.inv_eta[d];
    const double t_hi = (c.ub_plus_eps[d] - c.lb[d]) * c.inv_eta[d];
    const double A = t_lo - ua, B = t_hi - ub;
    const double lim = 1073741824.0;  // 2^30: all four stay far inside int32
    if (!(fabs(ua) < lim && fabs(ub) < lim && fabs(A) < lim && fabs(B) < lim)) return false;
    const double fa = floor(ua), cb = ceil(ub), fA = floor(A), fB = floor(B);
    const double r1 = ua - fa, r2 = cb - ub, r3 = A - fA, r4 = B - fB;  // in [0, 1)
    const bool clear = (r1 > eq) & (r1 < 1.0 - eq) & (r2 > eq) & (r2 < 1.0 - eq) & (r3 > eq) & (r3 < 1.0 - eq) &
                       (r4 > eq) & (r4 < 1.0 - eq);
    if (!clear) return false;
    o->off_lo = static_cast<int32_t>(fa);
    o->off_hi = static_cast<int32_t>(cb) - 1;
    o->rej_lo = static_cast<int32_t>(fA) + 1;  // i + ua < t_lo  <=>  i < A  <=>  i <= floor(A)
    o->rej_hi = static_cast<int32_t>(fB);      // i + ub > t_hi  <=>  i > B  <=>  i >= floor(B) + 1
    return true;
}

// The cells of cell i from its (plane, input) offsets: false when the box
// leaves the domain (dim_to_cells' rejection). The clamped range is written
// either way (cells in [0, n - 1]), so callers select rather than branch.
GSC_HD inline bool cells_from_offsets(const CellOffsets& o, int32_t i, uint32_t n, uint32_t* lo, uint32_t* hi) {
    const int32_t top = static_cast<int32_t>(n) - 1;
    const int32_t a = i + o.off_lo, b = i + o.off_hi;
    *lo = static_cast<uint32_t>(a < 0 ? 0 : a > top ? top : a);
    *hi = static_cast<uint32_t>(b < 0 ? 0 : b > top ? top : b);
    return (i >= o.rej_lo) & (i <= o.rej_hi);
}

// 16-corner inclusion–exclusion on a padded u32 prefix table (prefix_sum.hpp:65-85)
// for a box with lo <= hi in every dimension. Exact: the true count lies in [0, S].
GSC_HD inline uint32_t box_count_u32(const uint32_t* __restrict__ P, const uint32_t ps[4], const uint32_t lo[4],
                                     const uint32_t hi[4]) {
    const uint32_t x0 = lo[0], x1 = hi[0] + 1;
    const uint32_t y0 = lo[1] * ps[1], y1 = (hi[1] + 1) * ps[1];
    const uint32_t z0 = lo[2] * ps[2], z1 = (hi[2] + 1) * ps[2];
    const uint32_t w0 = lo[3] * ps[3], w1 = (hi[3] + 1) * ps[3];
    const uint32_t a11 = w1 + z1, a10 = w1 + z0, a01 = w0 + z1, a00 = w0 + z0;
    uint32_t s = 0;
    s += P[x1 + y1 + a11];
    s -= P[x0 + y1 + a11];
    s -= P[x1 + y0 + a11];
    s += P[x0 + y0 + a11];
    s -= P[x1 + y1 + a10];
    s += P[x0 + y1 + a10];
    s += P[x1 + y0 + a10];
    s -= P[x0 + y0 + a10];
    s -= P[x1 + y1 + a01];
    s += P[x0 + y1 + a01];
    s += P[x1 + y0 + a01];
    s -= P[x0 + y0 + a01];
    s += P[x1 + y1 + a00];
    s -= P[x0 + y1 + a00];
    s -= P[x1 + y0 + a00];
    s += P[x0 + y0 + a00];
    return s;
}

// query_box_count (prefix_sum.hpp:87-105): a box whose wrap-dimension range
// wraps (min > max) is split in two.
GSC_HD inline uint32_t box_count_wrap_u32(const uint32_t* __restrict__ P, const uint32_t ps[4], int wd,
                                          const uint32_t shape_wd, const uint32_t lo[4], const uint32_t hi[4]) {
    if (lo[wd] <= hi[wd]) return box_count_u32(P, ps, lo, hi);
    uint32_t hl[4] = {hi[0], hi[1], hi[2], hi[3]};
    uint32_t lr[4] = {lo[0], lo[1], lo[2], lo[3]};
    hl[wd] = shape_wd - 1;
    lr[wd] = 0;
    return box_count_u32(P, ps, lo, hl) + box_count_u32(P, ps, lr, hi);
}

// Same query with signed 64-bit accumulation, exact for any (even inverted)
// box: used for escaped pairs, where the reference's u64 modular sums of
// inverted ranges must be matched as integers.
GSC_HD inline int64_t box_count_i64_single(const uint32_t* __restrict__ P, const uint32_t ps[4], const uint32_t lo[4],
                                           const uint32_t hi[4]) {
    int64_t s = 0;
    for (int m = 0; m < 16; ++m) {
        uint64_t off = 0;
        for (int d = 0; d < 4; ++d) {
            const uint64_t v = (m >> d) & 1 ? lo[d] : static_cast<uint64_t>(hi[d]) + 1;
            off += v * ps[d];
        }
        const int64_t val = static_cast<int64_t>(P[off]);
        s += (__builtin_popcount(static_cast<unsigned>(m)) & 1) ? -val : val;
    }
    return s;
}

GSC_HD inline int64_t box_count_i64(const uint32_t* __restrict__ P, const uint32_t ps[4], int wd, uint32_t shape_wd,
                                    const uint32_t lo[4], const uint32_t hi[4]) {
    if (lo[wd] <= hi[wd]) return box_count_i64_single(P, ps, lo, hi);
    uint32_t hl[4] = {hi[0], hi[1], hi[2], hi[3]};
    uint32_t lr[4] = {lo[0], lo[1], lo[2], lo[3]};
    hl[wd] = shape_wd - 1;
    lr[wd] = 0;
    return box_count_i64_single(P, ps, lo, hl) + box_count_i64_single(P, ps, lr, hi);
}

}  // namespace gsc_amd
