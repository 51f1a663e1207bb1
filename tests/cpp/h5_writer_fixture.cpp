// Writes a result file from a hand-made GpuRunReport with the backend's HDF5
// writer (host/result_io.cpp), so the layout can be checked without a GPU.
//   h5_writer_fixture <out.h5>
// Sizes: 4x3x2x2 states (48), 2x3 inputs (6), 3 iterations, 2 obstacles.
// Array values are simple functions of the index that the test recomputes.
#include <cstdio>

#include "gsc/result_io.hpp"

int main(int argc, char** argv) {
    if (argc != 2) {
        std::fprintf(stderr, "usage: h5_writer_fixture <out.h5>\n");
        return 1;
    }
    gsc::CaseConfig cfg;
    cfg.name = "fixture case";
    gsc::HyperRect4D o{};
    for (int d = 0; d < gsc::kStateDim; ++d) {
        o.lb[d] = d;
        o.ub[d] = d + 0.5;
    }
    cfg.obstacles = {o, o};
    gsc::GpuRunReport r;
    r.compiled_with_cuda = true;
    r.runtime_available = true;
    r.executed = true;
    r.message = "ok";
    const unsigned shape[4] = {4, 3, 2, 2};
    std::uint64_t stride = 1;
    for (int d = 0; d < 4; ++d) {
        r.state_grid.shape[d] = shape[d];
        r.state_grid.strides[d] = stride;
        r.state_grid.lb[d] = cfg.state_lb[d];
        r.state_grid.ub[d] = cfg.state_ub[d];
        r.state_grid.eta[d] = cfg.state_eta[d];
        stride *= shape[d];
    }
    r.state_grid.total_size = stride;
    r.input_grid.shape[0] = 2;
    r.input_grid.shape[1] = 3;
    r.input_grid.strides[0] = 1;
    r.input_grid.strides[1] = 2;
    r.input_grid.total_size = 6;
    for (int d = 0; d < 2; ++d) {
        r.input_grid.lb[d] = cfg.input_lb[d];
        r.input_grid.ub[d] = cfg.input_ub[d];
        r.input_grid.eta[d] = cfg.input_eta[d];
    }
    const std::size_t S = 48, P = S * 6;
    r.pair_count = P;
    auto& res = r.result;
    res.valid_mask.resize(S);
    res.candidate_mask.resize(S);
    res.reachable_mask.resize(S);
    res.controller.resize(S);
    res.reach_step.resize(S);
    for (std::size_t s = 0; s < S; ++s) {
        res.valid_mask[s] = s % 5 != 0;
        res.candidate_mask[s] = s % 7 == 1;
        res.reachable_mask[s] = s % 3 == 0;
        res.controller[s] = static_cast<std::uint32_t>(s * 2654435761u);
        res.reach_step[s] = static_cast<std::uint32_t>(s % 4);
    }
    res.iterations = 3;
    res.converged = true;
    res.message = "all candidate states obtained a controller";
    res.reachable_states = 16;
    res.certified_candidate_states = 7;
    res.solve_ms = 1.5;
    r.abstraction_ms = 2.5;
    r.solve_ms = 1.5;
    r.total_ms = 4.0;
    r.abstraction_min_flat.resize(P);
    r.abstraction_max_flat.resize(P);
    r.abstraction_valid.resize(P);
    for (std::size_t p = 0; p < P; ++p) {
        r.abstraction_min_flat[p] = static_cast<std::uint32_t>(p % S);
        r.abstraction_max_flat[p] = static_cast<std::uint32_t>((p * 7) % S);
        r.abstraction_valid[p] = p % 4 != 0;
    }
    for (std::uint32_t it = 1; it <= 3; ++it) {
        gsc::IterationStats st;
        st.iteration = it;
        st.newly_reachable = 10 * it;
        st.newly_certified = it;
        st.total_reachable = 100 + it;
        st.total_certified = 200 + it;
        st.iteration_ms = 0.25 * it;
        r.iteration_stats.push_back(st);
    }
    gsc::write_gpu_results_hdf5(argv[1], cfg, r);
    return 0;
}
