// Property check of the filtered cell rounding (dim_to_cells_fast) against the
// exact step-by-step path (the reference's x <- x + inc_k for k < K, then
// dim_to_cells), on random grids and on coordinates placed next to every
// threshold the filter guards: cell edges (floor / ceil of (edge - lb) / eta)
// and the domain limits lb - kEps, ub + kEps. Whenever the filter decides, its
// answer must equal the exact one; undecided cases are counted.
//
// A second part checks the per-(plane, input) cell offsets the same way, with
// the cell centre kept and the increments moved instead.
//
// Usage: cell_filter_check <trials> <seed>; prints "checked <n> decided <d>
// rejected <r> mismatches <m>" for each part and exits 1 on any mismatch.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "gsc_device.hpp"

using namespace gsc_amd;

int main(int argc, char** argv) {
    const long trials = argc > 1 ? std::atol(argv[1]) : 1000000;
    const unsigned seed = argc > 2 ? static_cast<unsigned>(std::atoi(argv[2])) : 1u;
    std::mt19937_64 rng(seed);
    std::uniform_real_distribution<double> U(0.0, 1.0);
    auto pick = [&](double a, double b) { return a + (b - a) * U(rng); };
    long decided = 0, mismatches = 0, rejected = 0;
    std::vector<double> inc;
    for (long t = 0; t < trials; ++t) {
        DevCase c{};
        const int d = 0;
        const double lb = pick(-50.0, 50.0);
        const double eta = std::exp(pick(std::log(1e-3), std::log(2.0)));
        const uint32_t n = 1 + static_cast<uint32_t>(U(rng) * 999.0);
        c.lb[d] = lb;
        c.eta[d] = eta;
        c.ub[d] = lb + n * eta;
        c.shape[d] = n;
        c.lb_minus_eps[d] = c.lb[d] - kEps;
        c.ub_plus_eps[d] = c.ub[d] + kEps;
        c.wrap_dim = 2;
        const uint32_t K = (const uint32_t[]){1, 8, 32, 64}[rng() % 4];
        const double h = std::exp(pick(std::log(1e-6), std::log(1.0)));
        inc.resize(K);
        for (uint32_t k = 0; k < K; ++k) inc[k] = pick(-h, h);
        const double rad = pick(0.0, 5.0) * eta;
        double x0 = cell_center(lb, static_cast<uint32_t>(U(rng) * n), eta);
        auto exact_coord = [&](double start) {
            double p = start;
            for (uint32_t k = 0; k < K; ++k) p = p + inc[k];
            return p;
        };
        const int mode = static_cast<int>(rng() % 4);
        if (mode != 0) {
            // Move the start so that one box edge lands next to a threshold.
            const double p = exact_coord(x0);
            const double edge = (rng() & 1) ? p - rad : p + rad;
            double target;
            if (mode == 1)
                target = lb + std::floor(pick(0.0, n)) * eta;  // a cell edge
            else if (mode == 2)
                target = (rng() & 1) ? c.lb_minus_eps[d] : c.ub_plus_eps[d];  // a domain limit
            else
                target = lb + std::floor(pick(0.0, n)) * eta + 0.5 * eta;  // a cell centre (control)
            const double delta = (rng() & 1 ? 1.0 : -1.0) * std::ldexp(1.0, -static_cast<int>(rng() % 60)) *
                                 std::fabs(target + 1.0) * 1e-2;
            x0 += (target + ((rng() % 3) == 0 ? 0.0 : delta)) - edge;
        }
        // Exact path.
        const double p = exact_coord(x0);
        uint32_t elo = 0, ehi = 0;
        const bool eok = dim_to_cells(c, d, p - rad, p + rad, &elo, &ehi);
        // Filtered path, evaluated as the kernels do (sum in step order, then the
        // quotient terms), with the case-wide bounds the solver would derive:
        // |x0|, sum |inc_k| <= K h, |rad|.
        double S = 0.0;
        for (uint32_t k = 0; k < K; ++k) S = S + inc[k];
        set_cell_filter(c, d, K, std::max(std::max(std::fabs(c.lb[d]), std::fabs(c.ub[d])), std::fabs(x0)), K * h, rad);
        const double ie = c.inv_eta[d];
        uint32_t flo = 0, fhi = 0;
        const int r = dim_to_cells_fast(c, d, (x0 - lb) * ie, (S - rad) * ie, (S + rad) * ie, &flo, &fhi);
        if (r == kCellsUndecided) continue;
        ++decided;
        if (r == kCellsReject) ++rejected;
        const bool same = (r == kCellsReject) ? !eok : (eok && flo == elo && fhi == ehi);
        if (!same) {
            if (mismatches < 10)
                std::printf("mismatch: trial %ld lb %.17g eta %.17g n %u K %u x0 %.17g rad %.17g: exact %d [%u,%u], "
                            "filtered %d [%u,%u]\n",
                            t, lb, eta, n, K, x0, rad, eok ? 1 : 0, elo, ehi, r, flo, fhi);
            ++mismatches;
        }
    }
    std::printf("checked %ld decided %ld rejected %ld mismatches %ld\n", trials, decided, rejected, mismatches);

    // Per-(plane, input) cell offsets (cell_offsets / cells_from_offsets): x0
    // stays a cell centre, and the last increment is moved so that one box
    // edge lands next to a threshold (cell edge, domain limit, or a cell centre
    // as control). Whenever the offsets decide, the cells (or the rejection)
    // must equal the exact path's.
    long u_decided = 0, u_rejected = 0, u_mismatches = 0;
    for (long t = 0; t < trials; ++t) {
        DevCase c{};
        const int d = 1;
        const double lb = pick(-50.0, 50.0);
        const double eta = std::exp(pick(std::log(1e-3), std::log(2.0)));
        const uint32_t n = 1 + static_cast<uint32_t>(U(rng) * 999.0);
        c.lb[d] = lb;
        c.eta[d] = eta;
        c.ub[d] = lb + n * eta;
        c.shape[d] = n;
        c.lb_minus_eps[d] = c.lb[d] - kEps;
        c.ub_plus_eps[d] = c.ub[d] + kEps;
        c.wrap_dim = 2;
        const uint32_t K = (const uint32_t[]){1, 8, 32, 64}[rng() % 4];
        const double h = std::exp(pick(std::log(1e-6), std::log(1.0)));
        inc.resize(K);
        for (uint32_t k = 0; k < K; ++k) inc[k] = pick(-h, h);
        const double rad = pick(0.0, 5.0) * eta;
        const uint32_t idx = static_cast<uint32_t>(U(rng) * n);
        const double x0 = cell_center(lb, idx, eta);
        auto exact_coord = [&]() {
            double p = x0;
            for (uint32_t k = 0; k < K; ++k) p = p + inc[k];
            return p;
        };
        const int mode = static_cast<int>(rng() % 4);
        if (mode != 0) {
            const double p = exact_coord();
            const double edge = (rng() & 1) ? p - rad : p + rad;
            double target;
            if (mode == 1)
                target = lb + std::floor(pick(0.0, n)) * eta;
            else if (mode == 2)
                target = (rng() & 1) ? c.lb_minus_eps[d] : c.ub_plus_eps[d];
            else
                target = lb + std::floor(pick(0.0, n)) * eta + 0.5 * eta;
            const double delta = (rng() & 1 ? 1.0 : -1.0) * std::ldexp(1.0, -static_cast<int>(rng() % 60)) *
                                 std::fabs(target + 1.0) * 1e-2;
            inc[K - 1] += (target + ((rng() % 3) == 0 ? 0.0 : delta)) - edge;
        }
        const double p = exact_coord();
        uint32_t elo = 0, ehi = 0;
        const bool eok = dim_to_cells(c, d, p - rad, p + rad, &elo, &ehi);
        double S = 0.0, abs_sum = 0.0;
        for (uint32_t k = 0; k < K; ++k) {
            S = S + inc[k];
            abs_sum += std::fabs(inc[k]);
        }
        set_cell_filter(c, d, K, std::max(std::max(std::fabs(c.lb[d]), std::fabs(c.ub[d])), std::fabs(x0)), abs_sum,
                        rad);
        const double ie = c.inv_eta[d];
        CellOffsets o;
        if (!cell_offsets(c, d, (S - rad) * ie, (S + rad) * ie, &o)) continue;
        ++u_decided;
        uint32_t flo = 0, fhi = 0;
        const bool ok = cells_from_offsets(o, static_cast<int32_t>(idx), n, &flo, &fhi);
        if (!ok) ++u_rejected;
        const bool same = ok ? (eok && flo == elo && fhi == ehi) : !eok;
        if (!same) {
            if (u_mismatches < 10)
                std::printf("offsets mismatch: trial %ld lb %.17g eta %.17g n %u K %u idx %u rad %.17g S %.17g: exact %d "
                            "[%u,%u], offsets %d [%u,%u]\n",
                            t, lb, eta, n, K, idx, rad, S, eok ? 1 : 0, elo, ehi, ok ? 1 : 0, flo, fhi);
            ++u_mismatches;
        }
    }
    std::printf("offsets checked %ld decided %ld rejected %ld mismatches %ld\n", trials, u_decided, u_rejected,
                u_mismatches);
    return mismatches == 0 && u_mismatches == 0 ? 0 : 1;
}
