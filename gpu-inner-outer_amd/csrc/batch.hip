// batch.hip — many independent cases advanced together on one device.
//
// Each case keeps its own gsc_solver (tables, masks, pair records, prefix
// tables, iteration state). What the batch adds is the launch structure: every
// stage of an iteration is ONE launch over the concatenated work of all cases
// (block ranges per case in `starts`), and `iterations_per_launch` iterations
// are captured once in a HIP graph that is replayed until every case has hit
// its own stop rule. Cases that stopped return at once from every stage, so a
// replay costs only the work of the cases still iterating.
//
// Tile and schedule choices — satisfaction workgroup size, inputs per block of
// record loads, iterations per replay, order of the cases in the grid — are
// plain options; gsc_batch_tune() times the candidates on the live batch and
// keeps the fastest. None of them changes a result.
#include <memory>
#include <numeric>

#include "solver.hpp"

namespace gsc_amd {
namespace {

enum Stage : uint32_t {
    kSeed = 0, kRows, kAxis1, kAxis2, kAxis3, kCoarseStage, kBits, kChunk, kSat, kPlane, kLast,
    kFrontierBits, kFrontierV, kFrontierPlane, kStages
};

bool options_valid(const gsc_batch_options& o) {
    return satisfy_variant_supported(static_cast<uint32_t>(o.tile_wg), static_cast<uint32_t>(o.tile_ib)) &&
           o.iterations_per_launch >= 1 && o.iterations_per_launch <= static_cast<int32_t>(kRing) &&
           (o.order == 0 || o.order == 1);
}

}  // namespace
}  // namespace gsc_amd

struct gsc_batch {
    gsc_batch_options opts{};
    std::vector<std::unique_ptr<gsc_solver>> cases;  // caller order
    std::vector<uint32_t> order;                      // grid position -> caller index
    std::vector<uint32_t> position;                   // caller index -> grid position
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;

    BatchInstance* d_inst = nullptr;
    uint32_t* d_starts = nullptr;  // kStages x (n + 1)
    BatchStatus* d_status = nullptr;
    IterRecord* d_records = nullptr;  // n x kRing, grid order
    BatchStatus* h_status = nullptr;  // pinned
    IterRecord* h_records = nullptr;  // pinned
    uint32_t grid[kStages] = {};
    // Fused prefix build (plane kernel, θ pass, v pass with the coarse sample)
    // when every case's plane fits in LDS; its dynamic LDS bytes.
    bool fused_prefix = false;
    uint32_t plane_lds = 0;
    // Bitmask frontier (frontier_*_batch_kernel) when every case qualifies
    // (gsc_solver::bits_frontier); its dynamic LDS bytes.
    bool bits_frontier = false;
    uint32_t frontier_lds = 0;
    std::vector<void*> device_allocs;

    // One captured, instantiated graph per (tile_wg, tile_ib, iterations per
    // launch, frontier mode), kept until the batch is destroyed. Its nodes read the instance
    // table and block ranges from d_inst / d_starts at run time, and its grid
    // sizes depend only on the tile width (the per-stage totals do not depend
    // on the case order), so neither a re-plan nor a new abstraction outdates
    // it. Tuning therefore never destroys and re-instantiates graphs while the
    // batch lives.
    struct GraphEntry {
        int32_t wg, ib, g;
        bool bits;
        hipGraph_t graph;
        hipGraphExec_t exec;
    };
    std::vector<GraphEntry> graphs;
    hipGraphExec_t exec = nullptr;  // the entry of the current options, once captured
    bool planned = false;
    bool abstraction_ready = false;
    bool solved = false;

    std::vector<std::vector<gsc_iteration_stats>> stats;  // caller order
    std::vector<gsc_solve_summary> case_summary;          // caller order
    gsc_batch_summary summary{};
    std::string last_error;

    ~gsc_batch() {
        if (stream) (void)hipStreamSynchronize(stream);
        drop_graphs();
        for (void* p : device_allocs) (void)hipFree(p);
        if (h_status) (void)hipHostFree(h_status);
        if (h_records) (void)hipHostFree(h_records);
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        cases.clear();
        if (stream) (void)hipStreamDestroy(stream);
    }

    uint32_t n() const { return static_cast<uint32_t>(cases.size()); }

    template <typename T>
    T* dalloc(uint64_t count) {
        void* p = nullptr;
        GSC_HIP(hipMalloc(&p, std::max<uint64_t>(count, 1) * sizeof(T)));
        device_allocs.push_back(p);
        return static_cast<T*>(p);
    }

    // Destructor only, after the stream has drained.
    void drop_graphs() {
        for (GraphEntry& e : graphs) {
            if (e.exec) (void)hipGraphExecDestroy(e.exec);
            if (e.graph) (void)hipGraphDestroy(e.graph);
        }
        graphs.clear();
        exec = nullptr;
    }

    // The cached graph of the current options; nullptr when not captured yet.
    hipGraphExec_t find_exec() const {
        for (const GraphEntry& e : graphs)
            if (e.wg == opts.tile_wg && e.ib == opts.tile_ib && e.g == opts.iterations_per_launch &&
                e.bits == bits_frontier)
                return e.exec;
        return nullptr;
    }

    void create(const gsc_case_desc* const* descs, uint32_t count);
    void build_abstraction();
    void plan();
    void enqueue_iterations(uint32_t iterations);
    void solve();
};

void gsc_batch::create(const gsc_case_desc* const* descs, uint32_t count) {
    GSC_HIP(hipSetDevice(opts.device));
    GSC_HIP(hipStreamCreate(&stream));
    GSC_HIP(hipEventCreate(&ev0));
    GSC_HIP(hipEventCreate(&ev1));
    for (uint32_t k = 0; k < count; ++k) {
        auto s = std::make_unique<gsc_solver>();
        gsc_solver_default_options(&s->opts);
        s->opts.device = opts.device;
        s->opts.time_kernels = 0;
        copy_case(s.get(), *descs[k]);
        try {
            s->setup();
        } catch (const ConfigError& e) {
            throw ConfigError("case " + std::to_string(k) + ": " + e.what());
        }
        cases.push_back(std::move(s));
    }
    const uint32_t N = n();
    d_inst = dalloc<BatchInstance>(N);
    d_starts = dalloc<uint32_t>(static_cast<uint64_t>(kStages) * (N + 1));
    d_status = dalloc<BatchStatus>(N);
    d_records = dalloc<IterRecord>(static_cast<uint64_t>(N) * kRing);
    GSC_HIP(hipHostMalloc(reinterpret_cast<void**>(&h_status), sizeof(BatchStatus) * N, hipHostMallocDefault));
    GSC_HIP(hipHostMalloc(reinterpret_cast<void**>(&h_records), sizeof(IterRecord) * N * kRing, hipHostMallocDefault));
    std::memset(h_status, 0, sizeof(BatchStatus) * N);
    std::memset(h_records, 0, sizeof(IterRecord) * N * kRing);
    stats.assign(N, {});
    case_summary.assign(N, gsc_solve_summary{});
    summary.cases = N;
    for (const auto& c : cases) summary.pairs += c->shape.pair_count;
}

void gsc_batch::build_abstraction() {
    const auto t0 = std::chrono::steady_clock::now();
    // Every case on its own stream, all in flight together; then the waits.
    for (auto& c : cases) c->launch_abstraction();
    for (auto& c : cases) c->finish_abstraction();
    summary.abstraction_ms = ms_since(t0);
    abstraction_ready = true;
    planned = false;  // overflow tables may have moved
}

// Grid order, per-stage block ranges and the per-case argument blocks.
void gsc_batch::plan() {
    const uint32_t N = n();
    order.resize(N);
    std::iota(order.begin(), order.end(), 0u);
    if (opts.order == 1)
        std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
            return cases[a]->shape.pair_count > cases[b]->shape.pair_count;
        });
    position.assign(N, 0);
    for (uint32_t p = 0; p < N; ++p) position[order[p]] = p;

    std::vector<BatchInstance> h_inst(N);
    std::vector<uint32_t> starts(static_cast<size_t>(kStages) * (N + 1), 0);
    uint64_t total[kStages] = {};
    bool fused = true;
    uint64_t lds = 0, flds = 0;
    bool bits = N > 0;
    for (uint32_t p = 0; p < N; ++p) bits = bits && cases[p]->bits_frontier();
    for (uint32_t p = 0; p < N; ++p) {
        const gsc_solver& s = *cases[order[p]];
        const DevCase& c = s.dc;
        uint64_t blocks[kStages];
        blocks[kSeed] = (c.S + 255) / 256;
        blocks[kRows] = (static_cast<uint64_t>(c.shape[1]) * c.shape[2] * c.shape[3] + 3) / 4;
        for (int axis = 1; axis <= 3; ++axis) {
            const int o1 = axis == 1 ? 2 : 1;
            const int o2 = axis == 3 ? 2 : 3;
            blocks[kAxis1 + axis - 1] = (static_cast<uint64_t>(c.shape[0]) * c.shape[o1] * c.shape[o2] + 255) / 256;
        }
        blocks[kCoarseStage] = (s.coarse_count + 255) / 256;
        blocks[kBits] = (c.S + 255) / 256;
        blocks[kChunk] = (c.n_tasks + 255) / 256;
        blocks[kPlane] = static_cast<uint64_t>(c.shape[2]) * c.shape[3];
        blocks[kLast] = (static_cast<uint64_t>(c.shape[0]) * c.shape[1] * c.shape[2] + 255) / 256;
        blocks[kSat] = satisfy_blocks(c, static_cast<uint32_t>(opts.tile_wg));
        const uint64_t planes = static_cast<uint64_t>(c.shape[2]) * c.shape[3];
        blocks[kFrontierBits] = ((c.S + 7ull) / 8ull + 255) / 256;
        blocks[kFrontierV] = (planes * c.shape[1] * ((c.shape[0] + 31ull) / 32ull) + 255) / 256;
        blocks[kFrontierPlane] = planes;
        flds = std::max<uint64_t>(flds, frontier_lds_bytes(c));
        for (uint32_t st = 0; st < kStages; ++st) {
            starts[st * (N + 1) + p] = static_cast<uint32_t>(total[st]);
            total[st] += blocks[st];
            if (total[st] > 0x7FFFFFFFull) throw ConfigError("batch too large for one launch grid");
        }
        BatchInstance& b = h_inst[p];
        b.sat.c = c;
        b.sat.prefix_unreached[0] = s.prefix_unreached[0];
        b.sat.prefix_unreached[1] = s.prefix_unreached[1];
        b.sat.coarse_unreached[0] = s.coarse_unreached[0];
        b.sat.coarse_unreached[1] = s.coarse_unreached[1];
        b.sat.coarse_valid = s.coarse_valid;
        b.sat.hull = s.hull;
        b.sat.chunk_hull = s.chunk_hull;
        b.sat.chunk_flags = s.chunk_flags;
        b.plane_pitch = s.plane_pitch;
        fused = fused && s.plane_pitch != 0;
        lds = std::max<uint64_t>(lds, plane_lds_bytes(s.plane_pitch, c.shape[1], c.PQ));
        b.sat.hull2 = s.hull2;
        b.sat.hull3 = s.hull3;
        b.sat.prefix_valid = s.prefix_valid;
        b.sat.unreached_bits = s.ubits;
        b.sat.probe_base = s.probe_base;
        // Written by chunk_flags_batch_kernel, which runs where frontier skipping does.
        b.sat.probe_or_bits = s.probe_or_used() ? s.probe_or_bits : nullptr;
        b.sat.valid_bits = s.valid_bits;
        b.sat.frontier_bits = bits ? s.frontier_words : nullptr;
        b.fr = s.frontier_args();
        b.sat.rec = s.rec;
        b.sat.overflow = s.overflow;
        b.sat.box2 = s.box2;
        b.sat.box3 = s.box3;
        b.sat.valid = s.valid;
        b.sat.cand = s.cand;
        b.sat.reach = s.reach;
        b.sat.ctrl = s.ctrl;
        b.sat.rstep = s.rstep;
        b.sat.ctl = s.ctl;
        b.sat.sweep_all = 0;
        b.prefix_unreached[0] = s.prefix_unreached[0];
        b.prefix_unreached[1] = s.prefix_unreached[1];
        b.coarse_unreached[0] = s.coarse_unreached[0];
        b.coarse_unreached[1] = s.coarse_unreached[1];
        b.unreached_bits = s.ubits;
        b.records = d_records + static_cast<size_t>(p) * kRing;
        b.candidate_count = s.candidate_count;
        b.max_iterations = s.cfg.max_iterations;
    }
    fused_prefix = fused && N > 0;
    plane_lds = static_cast<uint32_t>(lds);
    bits_frontier = bits;
    frontier_lds = static_cast<uint32_t>(flds);
    for (uint32_t st = 0; st < kStages; ++st) {
        starts[st * (N + 1) + N] = static_cast<uint32_t>(total[st]);
        grid[st] = static_cast<uint32_t>(total[st]);
    }
    GSC_HIP(hipMemcpyAsync(d_inst, h_inst.data(), sizeof(BatchInstance) * N, hipMemcpyHostToDevice, stream));
    GSC_HIP(hipMemcpyAsync(d_starts, starts.data(), sizeof(uint32_t) * starts.size(), hipMemcpyHostToDevice, stream));
    GSC_HIP(hipStreamSynchronize(stream));
    planned = true;
}

// `iterations` iterations of every unfinished case, then the status and record
// copies to pinned host memory. Captured as is into the graph.
void gsc_batch::enqueue_iterations(uint32_t iterations) {
    const uint32_t N = n();
    const uint32_t* st = d_starts;
    for (uint32_t it = 0; it < iterations; ++it) {
        if (bits_frontier) {
            // The unreached bitmask and the states reached since the last
            // iteration, the v pass, then a workgroup per plane of every case.
            hipLaunchKernelGGL(frontier_bits_batch_kernel, dim3(grid[kFrontierBits]), dim3(256), 0, stream, d_inst,
                               st + kFrontierBits * (N + 1), N);
            hipLaunchKernelGGL(frontier_v_batch_kernel, dim3(grid[kFrontierV]), dim3(256), 0, stream, d_inst,
                               st + kFrontierV * (N + 1), N);
            hipLaunchKernelGGL(frontier_plane_batch_kernel, dim3(grid[kFrontierPlane]), dim3(kFrontierThreads),
                               frontier_lds, stream, d_inst, st + kFrontierPlane * (N + 1), N);
        } else if (fused_prefix) {
            hipLaunchKernelGGL(prefix_plane_batch_kernel, dim3(grid[kPlane]), dim3(kPlaneThreads), plane_lds, stream,
                               d_inst, st + kPlane * (N + 1), N);
            hipLaunchKernelGGL(prefix_axis_batch_kernel, dim3(grid[kAxis2]), dim3(256), 0, stream, d_inst,
                               st + kAxis2 * (N + 1), N, 2);
            hipLaunchKernelGGL(prefix_last_axis_batch_kernel, dim3(grid[kLast]), dim3(256), 0, stream, d_inst,
                               st + kLast * (N + 1), N);
        } else {
            hipLaunchKernelGGL(prefix_rows_batch_kernel, dim3(grid[kRows]), dim3(256), 0, stream, d_inst,
                               st + kRows * (N + 1), N);
            for (int axis = 1; axis <= 3; ++axis) {
                const uint32_t s = kAxis1 + axis - 1;
                hipLaunchKernelGGL(prefix_axis_batch_kernel, dim3(grid[s]), dim3(256), 0, stream, d_inst, st + s * (N + 1),
                                   N, axis);
            }
            hipLaunchKernelGGL(coarse_batch_kernel, dim3(grid[kCoarseStage]), dim3(256), 0, stream, d_inst,
                               st + kCoarseStage * (N + 1), N);
            hipLaunchKernelGGL(unreached_bits_batch_kernel, dim3(grid[kBits]), dim3(256), 0, stream, d_inst,
                               st + kBits * (N + 1), N);
        }
        if (!bits_frontier)
            hipLaunchKernelGGL(chunk_flags_batch_kernel, dim3(grid[kChunk]), dim3(256), 0, stream, d_inst,
                               st + kChunk * (N + 1), N);
        GSC_HIP(launch_satisfy_batch(d_inst, st + kSat * (N + 1), N, grid[kSat], static_cast<uint32_t>(opts.tile_wg),
                                     static_cast<uint32_t>(opts.tile_ib), bits_frontier, stream));
        hipLaunchKernelGGL(finish_batch_kernel, dim3(N), dim3(64), 0, stream, d_inst, N, d_status);
        GSC_HIP(hipGetLastError());
    }
    GSC_HIP(hipMemcpyAsync(h_status, d_status, sizeof(BatchStatus) * N, hipMemcpyDeviceToHost, stream));
    GSC_HIP(hipMemcpyAsync(h_records, d_records, sizeof(IterRecord) * N * kRing, hipMemcpyDeviceToHost, stream));
}

void gsc_batch::solve() {
    if (!abstraction_ready) build_abstraction();
    if (!planned) plan();
    const uint32_t N = n();
    const uint32_t G = static_cast<uint32_t>(opts.iterations_per_launch);
    exec = opts.use_graph ? find_exec() : nullptr;
    if (opts.use_graph && !exec) {
        GSC_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
        try {
            enqueue_iterations(G);
        } catch (...) {
            hipGraph_t g = nullptr;
            (void)hipStreamEndCapture(stream, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
        }
        hipGraph_t graph = nullptr;
        GSC_HIP(hipStreamEndCapture(stream, &graph));
        const hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        if (e != hipSuccess) {
            (void)hipGraphDestroy(graph);
            exec = nullptr;
            throw DeviceError(e, hipGetErrorString(e));
        }
        graphs.push_back(GraphEntry{opts.tile_wg, opts.tile_ib, opts.iterations_per_launch, bits_frontier, graph, exec});
    }

    const auto t0 = std::chrono::steady_clock::now();
    hipLaunchKernelGGL(seed_batch_kernel, dim3(grid[kSeed]), dim3(256), 0, stream, d_inst, d_starts + kSeed * (N + 1),
                       N);
    GSC_HIP(hipGetLastError());
    for (auto& s : stats) s.clear();
    std::vector<uint32_t> seen(N, 0);
    uint32_t replays = 0, stopped = 0;
    float gpu_ms = 0.f;
    for (;;) {
        GSC_HIP(hipEventRecord(ev0, stream));
        if (opts.use_graph)
            GSC_HIP(hipGraphLaunch(exec, stream));
        else
            enqueue_iterations(G);
        GSC_HIP(hipEventRecord(ev1, stream));
        GSC_HIP(hipStreamSynchronize(stream));
        float ms = 0.f;
        GSC_HIP(hipEventElapsedTime(&ms, ev0, ev1));
        gpu_ms += ms;
        ++replays;
        stopped = 0;
        for (uint32_t p = 0; p < N; ++p) {
            const BatchStatus& st = h_status[p];
            const uint32_t k = order[p];
            for (uint32_t it = seen[p] + 1; it <= st.iterations; ++it) {
                const IterRecord& r = h_records[static_cast<size_t>(p) * kRing + it % kRing];
                gsc_iteration_stats s;
                std::memset(&s, 0, sizeof(s));
                s.iteration = r.iteration;
                s.newly_reachable = r.newly_reachable;
                s.newly_certified = r.newly_certified;
                s.total_reachable = r.total_reachable;
                s.total_certified = r.total_certified;
                stats[k].push_back(s);
                case_summary[k].pairs_tested += r.pairs_tested;
                case_summary[k].boxes_counted += r.boxes_counted;
                case_summary[k].test_slots += r.test_slots;
            }
            seen[p] = st.iterations;
            if (st.done) ++stopped;
        }
        if (stopped == N) break;
    }
    const double wall = ms_since(t0);
    uint32_t max_iter = 0;
    for (uint32_t p = 0; p < N; ++p) {
        const BatchStatus& st = h_status[p];
        const uint32_t k = order[p];
        gsc_solver& s = *cases[k];
        gsc_solve_summary& cs = case_summary[k];
        const uint64_t pt = cs.pairs_tested, bc = cs.boxes_counted, ts = cs.test_slots;
        std::memset(&cs, 0, sizeof(cs));
        cs.iterations = st.iterations;
        cs.converged = st.done == kDoneConverged ? 1 : 0;
        std::snprintf(cs.message, sizeof(cs.message), "%s", done_message(st.done));
        cs.reachable_states = st.total_reachable;
        cs.certified_candidate_states = st.total_certified;
        cs.candidate_states = s.candidate_count;
        cs.abstraction_ms = s.abstraction_kernel_ms;
        cs.solve_ms = wall;
        cs.pairs_tested = pt;
        cs.boxes_counted = bc;
        cs.test_slots = ts;
        cs.pair_bytes_read = pt * sizeof(uint32_t);
        // The batch runs the bitmask frontier only when every case can.
        cs.frontier = s.frontier_mode() == GSC_FRONTIER_BITS && !bits_frontier ? GSC_FRONTIER_PREFIX : s.frontier_mode();
        s.summary = cs;
        s.iter_stats = stats[k];
        s.timings = gsc_kernel_timings{};
        s.timings.abstraction_kernel_ms = s.abstraction_kernel_ms;
        s.solved = true;
        max_iter = std::max(max_iter, st.iterations);
    }
    summary.replays = replays;
    summary.max_iterations = max_iter;
    summary.stopped = stopped;
    summary.solve_ms = wall;
    summary.gpu_ms = gpu_ms;
    solved = true;
}

// ===========================================================================
// C ABI
// ===========================================================================
namespace {

template <typename F>
int guarded(gsc_batch* b, F&& f) {
    try {
        f();
        return GSC_OK;
    } catch (const ConfigError& e) {
        if (b) b->last_error = e.what();
        return GSC_ERR_CONFIG;
    } catch (const DeviceError& e) {
        if (b) b->last_error = e.what();
        return status_of(e);
    } catch (const std::exception& e) {
        if (b) b->last_error = e.what();
        return GSC_ERR_DEVICE;
    }
}

}  // namespace

extern "C" {

void gsc_batch_default_options(gsc_batch_options* o) {
    if (!o) return;
    std::memset(o, 0, sizeof(*o));
    o->device = 0;
    o->tile_wg = 512;
    o->tile_ib = 8;
    o->iterations_per_launch = 8;
    o->order = 1;
    o->use_graph = 1;
}

int gsc_batch_create(const gsc_case_desc* const* cases, uint32_t count, const gsc_batch_options* opts,
                     gsc_batch** out, char* err, size_t err_len) {
    if (!cases || !out || count == 0) return GSC_ERR_INVALID_ARGUMENT;
    for (uint32_t k = 0; k < count; ++k)
        if (!cases[k]) return GSC_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    auto* b = new gsc_batch();
    if (opts)
        b->opts = *opts;
    else
        gsc_batch_default_options(&b->opts);
    if (!options_valid(b->opts)) {
        put(err, err_len, "invalid batch options");
        delete b;
        return GSC_ERR_INVALID_ARGUMENT;
    }
    const int rc = guarded(b, [&] { b->create(cases, count); });
    if (rc != GSC_OK) {
        put(err, err_len, b->last_error);
        delete b;
        return rc;
    }
    *out = b;
    return GSC_OK;
}

int gsc_batch_build_abstraction(gsc_batch* b, double* wall_ms) {
    if (!b) return GSC_ERR_INVALID_ARGUMENT;
    const int rc = guarded(b, [&] { b->build_abstraction(); });
    if (wall_ms) *wall_ms = b->summary.abstraction_ms;
    return rc;
}

int gsc_batch_solve(gsc_batch* b, gsc_batch_summary* out) {
    if (!b) return GSC_ERR_INVALID_ARGUMENT;
    const int rc = guarded(b, [&] { b->solve(); });
    if (rc == GSC_OK && out) *out = b->summary;
    return rc;
}

int gsc_batch_get_options(const gsc_batch* b, gsc_batch_options* out) {
    if (!b || !out) return GSC_ERR_INVALID_ARGUMENT;
    *out = b->opts;
    return GSC_OK;
}

int gsc_batch_set_options(gsc_batch* b, const gsc_batch_options* opts) {
    if (!b || !opts) return GSC_ERR_INVALID_ARGUMENT;
    if (!options_valid(*opts) || opts->device != b->opts.device) return GSC_ERR_INVALID_ARGUMENT;
    return guarded(b, [&] {
        // A new tile width or order needs new block ranges; the graph of the new
        // options is looked up (or captured) by the next solve.
        const bool replan = opts->tile_wg != b->opts.tile_wg || opts->order != b->opts.order;
        b->opts = *opts;
        if (replan) b->planned = false;
    });
}

int gsc_batch_tune(gsc_batch* b, gsc_batch_tune_entry* entries, uint32_t capacity, uint32_t* written) {
    if (!b) return GSC_ERR_INVALID_ARGUMENT;
    uint32_t count = 0;
    const int rc = guarded(b, [&] {
        if (!b->abstraction_ready) b->build_abstraction();
        auto time_config = [&](const gsc_batch_options& o) {
            if (gsc_batch_set_options(b, &o) != GSC_OK) throw std::runtime_error("tuner produced invalid options");
            double best = 1e300;
            for (int rep = 0; rep < 3; ++rep) {  // first solve also captures the graph
                b->solve();
                if (rep > 0) best = std::min(best, b->summary.solve_ms);
            }
            if (entries && count < capacity)
                entries[count] = gsc_batch_tune_entry{o.tile_wg, o.tile_ib, o.iterations_per_launch, o.order, best};
            ++count;
            return best;
        };
        // Coordinate search: tile shape, then iterations per replay, then case order.
        gsc_batch_options best = b->opts;
        best.use_graph = 1;
        double best_ms = 1e300;
        for (int wg : {128, 256, 512, 1024})
            for (int ib : {4, 8}) {
                gsc_batch_options o = best;
                o.tile_wg = wg;
                o.tile_ib = ib;
                const double t = time_config(o);
                if (t < best_ms) {
                    best_ms = t;
                    best.tile_wg = wg;
                    best.tile_ib = ib;
                }
            }
        gsc_batch_options stage = best;
        for (int g : {4, 16, 32}) {
            gsc_batch_options o = stage;
            o.iterations_per_launch = g;
            const double t = time_config(o);
            if (t < best_ms) {
                best_ms = t;
                best.iterations_per_launch = g;
            }
        }
        {
            gsc_batch_options o = best;
            o.order = 1 - best.order;
            const double t = time_config(o);
            if (t < best_ms) {
                best_ms = t;
                best.order = o.order;
            }
        }
        if (gsc_batch_set_options(b, &best) != GSC_OK) throw std::runtime_error("tuner produced invalid options");
        b->solve();  // leave results of the chosen configuration in place
    });
    if (written) *written = std::min(count, capacity);
    return rc;
}

int gsc_batch_case_summary(gsc_batch* b, uint32_t index, gsc_solve_summary* out) {
    if (!b || !out || index >= b->n()) return GSC_ERR_INVALID_ARGUMENT;
    if (!b->solved) {
        b->last_error = "batch not solved";
        return GSC_ERR_INVALID_ARGUMENT;
    }
    *out = b->case_summary[index];
    return GSC_OK;
}

int gsc_batch_download(gsc_batch* b, uint32_t index, const gsc_result_buffers* out, gsc_run_report* report) {
    if (!b || index >= b->n()) return GSC_ERR_INVALID_ARGUMENT;
    if (report) std::memset(report, 0, sizeof(*report));
    return guarded(b, [&] { b->cases[index]->download(out, report); });
}

const char* gsc_batch_last_error(const gsc_batch* b) { return b ? b->last_error.c_str() : ""; }

void gsc_batch_destroy(gsc_batch* b) { delete b; }

}  // extern "C"
