// serve.hip — a request loop over one device with continuous batching.
//
// Submitted cases queue in arrival order. A worker thread admits them while
// the HBM budget (one DeviceArena carved out at creation) and the slot count
// allow, and prepares each admitted case — masks, tables, abstraction — on a
// preparation stream WHILE the cases already admitted iterate on the iteration
// stream. Every step advances all admitted cases by `iterations_per_step`
// fused iterations (the batched stage kernels of batch.hip over the current
// set); between steps, cases that stopped leave (results copied to host
// memory, HBM returned to the budget) and prepared cases join. Nothing is
// allocated with hipMalloc while cases are on the device.
#include <condition_variable>
#include <deque>
#include <memory>
#include <thread>
#include <unordered_map>

#include "solver.hpp"

namespace gsc_amd {
namespace {

// The prefix-table stages (kRows .. kLast, kChunk, kSat) cover the cases that
// run frontier skipping on prefix tables, the kFrontier* stages and kSatBits
// the cases on the bitmask frontier (gsc_solver::bits_frontier): each case
// keeps its mode for its whole solve, and a stage's block ranges give the
// other mode's cases no blocks.
enum Stage : uint32_t {
    kSeed = 0, kRows, kAxis1, kAxis2, kAxis3, kCoarseStage, kBits, kChunk, kSat, kPlane, kLast,
    kFrontierBits, kFrontierV, kFrontierPlane, kSatBits, kStages
};

using Clock = std::chrono::steady_clock;

struct Request {
    uint64_t id = 0;
    std::unique_ptr<gsc_solver> spec;    // the submitted case (copied), never set up
    std::unique_ptr<gsc_solver> solver;  // device state while admitted
    gsc_request_state state = GSC_REQUEST_QUEUED;
    int status = GSC_OK;
    std::string error;
    gsc_case_shape shape{};
    uint64_t est_bytes = 0;
    Clock::time_point t_submit{}, t_admit{}, t_stop{}, t_done{};
    int slot = -1;
    bool in_retire = false;  // handed to the retire thread: only it touches the solver now
    uint32_t seen = 0;  // iterations already copied into `stats`
    // Results, on the host once the case stopped: one block of 11 B per state
    // (controller, reach step, then the valid, candidate and reachable masks),
    // in the server's pinned result pool or, when the pool is full, in
    // `fallback`. The pool block goes back when the request is destroyed.
    uint8_t* results = nullptr;
    DeviceArena* pool = nullptr;  // owner of `results` when it is a pool block
    std::vector<uint8_t> fallback;
    std::vector<gsc_iteration_stats> stats;

    ~Request() {
        if (results && pool) pool->release(results);
    }
    const uint32_t* ctrl() const { return reinterpret_cast<const uint32_t*>(results); }
    const uint32_t* rstep() const { return reinterpret_cast<const uint32_t*>(results) + shape.state_count; }
    const uint8_t* valid() const { return results + 8 * shape.state_count; }
    const uint8_t* cand() const { return results + 9 * shape.state_count; }
    const uint8_t* reach() const { return results + 10 * shape.state_count; }
    gsc_solve_summary summary{};
    gsc_run_report report{};
};

}  // namespace
}  // namespace gsc_amd

struct gsc_server {
    gsc_server_options opts{};
    std::unique_ptr<DeviceArena> arena;
    // Finished results until fetched: one pinned host block carved like the
    // device arena (its bookkeeping is plain offset arithmetic).
    void* pinned_results = nullptr;
    std::unique_ptr<DeviceArena> result_pool;
    hipStream_t iter_stream = nullptr, prep_stream = nullptr;
    BatchInstance* d_inst = nullptr;
    uint32_t* d_starts = nullptr;
    BatchStatus* d_status = nullptr;
    IterRecord* d_records = nullptr;  // max_active x kRing, by slot
    uint32_t* h_words = nullptr;      // pinned, kHostWords words per slot, lent to the slot's solver
    BatchStatus* h_status = nullptr;  // pinned
    IterRecord* h_records = nullptr;  // pinned
    uint32_t grid[kStages] = {};
    // Fused prefix build (plane kernel, θ pass, v pass with the coarse sample)
    // when every case's plane fits in LDS; its dynamic LDS bytes.
    bool fused_prefix = false;
    uint32_t plane_lds = 0;
    // Cases of the current step in each frontier mode, and the bitmask
    // frontier's dynamic LDS bytes.
    uint32_t n_prefix_mode = 0, n_bits_mode = 0;
    uint32_t frontier_lds = 0;

    std::mutex m;
    std::condition_variable cv_work, cv_done;
    std::deque<Request*> queue;
    std::unordered_map<uint64_t, std::unique_ptr<Request>> requests;
    std::vector<Request*> active;     // iterating, in grid order (worker only)
    std::vector<int> free_slots;      // guarded by `m`
    bool plan_dirty = true;           // worker only
    bool trace = false;               // GSC_SERVE_TRACE set: per-step host timings on stderr
    Clock::time_point t_created = Clock::now();
    uint64_t next_id = 1;
    bool stopping = false;
    std::string fatal;
    gsc_server_stats stats{};
    std::thread worker;
    // Stopped cases are copied back by a second thread on its own stream, so
    // the worker's next step runs while their results travel to the host.
    hipStream_t retire_stream = nullptr;
    std::deque<Request*> retire_q;  // guarded by `m`
    uint32_t retiring = 0;          // handed over, not yet done; guarded by `m`
    bool retire_stop = false;       // guarded by `m`
    std::condition_variable cv_retire;
    std::thread retirer;

    ~gsc_server();
    void run();
    void step_loop();
    void plan();
    void enqueue_step();
    void admit(std::vector<Request*>& admitted, std::vector<Request*>& prepared);
    void retire(Request* r, const BatchStatus& st);
    void retire_loop();
    void finish_retire(Request* r);
    void fail(Request* r, int status, const std::string& why);
    void release_slot(Request* r) {  // caller holds `m`
        if (r->slot >= 0) free_slots.push_back(r->slot);
        r->slot = -1;
    }
    void fail_all(const std::string& why);
};

namespace {

double ms_between(Clock::time_point a, Clock::time_point b) {
    return std::chrono::duration<double, std::milli>(b - a).count();
}

}  // namespace

gsc_server::~gsc_server() {
    {
        std::lock_guard<std::mutex> lock(m);
        stopping = true;
    }
    cv_work.notify_all();
    if (worker.joinable()) worker.join();
    {
        // The worker has handed over every stopped case; the retire thread
        // finishes them and leaves.
        std::lock_guard<std::mutex> lock(m);
        retire_stop = true;
    }
    cv_retire.notify_all();
    if (retirer.joinable()) retirer.join();
    requests.clear();  // solvers and results release their pieces before the pools go
    result_pool.reset();
    if (pinned_results) (void)hipHostFree(pinned_results);
    if (retire_stream) (void)hipStreamSynchronize(retire_stream);
    if (iter_stream) (void)hipStreamSynchronize(iter_stream);
    if (prep_stream) (void)hipStreamSynchronize(prep_stream);
    for (void* p : {static_cast<void*>(d_inst), static_cast<void*>(d_starts), static_cast<void*>(d_status),
                    static_cast<void*>(d_records)})
        if (p) (void)hipFree(p);
    if (h_status) (void)hipHostFree(h_status);
    if (h_records) (void)hipHostFree(h_records);
    if (h_words) (void)hipHostFree(h_words);
    arena.reset();
    if (iter_stream) (void)hipStreamDestroy(iter_stream);
    if (prep_stream) (void)hipStreamDestroy(prep_stream);
    if (retire_stream) (void)hipStreamDestroy(retire_stream);
}

// Caller holds `m`.
void gsc_server::fail(Request* r, int status, const std::string& why) {
    r->state = GSC_REQUEST_FAILED;
    r->status = status;
    r->error = why;
    r->t_done = Clock::now();
    r->solver.reset();
    ++stats.failed;
    cv_done.notify_all();
}

void gsc_server::fail_all(const std::string& why) {
    std::lock_guard<std::mutex> lock(m);
    fatal = why;
    for (auto& kv : requests)
        if ((kv.second->state == GSC_REQUEST_QUEUED || kv.second->state == GSC_REQUEST_RUNNING) &&
            !kv.second->in_retire)  // the retire thread finishes those
            fail(kv.second.get(), GSC_ERR_DEVICE, why);
    queue.clear();
    active.clear();
}

void gsc_server::plan() {
    const uint32_t N = static_cast<uint32_t>(active.size());
    std::vector<BatchInstance> h_inst(N);
    std::vector<uint32_t> starts(static_cast<size_t>(kStages) * (N + 1), 0);
    uint64_t total[kStages] = {};
    bool fused = true;
    uint64_t lds = 0, flds = 0;
    uint32_t n_bits = 0;
    for (uint32_t p = 0; p < N; ++p) {
        const gsc_solver& s = *active[p]->solver;
        const DevCase& c = s.dc;
        const bool bits = s.bits_frontier();
        n_bits += bits ? 1u : 0u;
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
        blocks[kSatBits] = blocks[kSat];
        // Only the stages of the case's own mode get blocks.
        for (uint32_t st = kRows; st < kStages; ++st) {
            const bool bits_stage = st == kFrontierBits || st == kFrontierV || st == kFrontierPlane || st == kSatBits;
            if (bits_stage != bits) blocks[st] = 0;
        }
        if (bits) flds = std::max<uint64_t>(flds, frontier_lds_bytes(c));
        for (uint32_t st = 0; st < kStages; ++st) {
            starts[st * (N + 1) + p] = static_cast<uint32_t>(total[st]);
            total[st] += blocks[st];
            if (total[st] > 0x7FFFFFFFull) throw ConfigError("admitted cases exceed one launch grid");
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
        if (!bits) {
            fused = fused && s.plane_pitch != 0;
            lds = std::max<uint64_t>(lds, plane_lds_bytes(s.plane_pitch, c.shape[1], c.PQ));
        }
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
        b.records = d_records + static_cast<size_t>(active[p]->slot) * kRing;
        b.candidate_count = s.candidate_count;
        b.max_iterations = s.cfg.max_iterations;
    }
    n_bits_mode = n_bits;
    n_prefix_mode = N - n_bits;
    fused_prefix = fused && n_prefix_mode > 0;
    plane_lds = static_cast<uint32_t>(lds);
    frontier_lds = static_cast<uint32_t>(flds);
    for (uint32_t st = 0; st < kStages; ++st) {
        starts[st * (N + 1) + N] = static_cast<uint32_t>(total[st]);
        grid[st] = static_cast<uint32_t>(total[st]);
    }
    if (N) {
        GSC_HIP(hipMemcpyAsync(d_inst, h_inst.data(), sizeof(BatchInstance) * N, hipMemcpyHostToDevice, iter_stream));
        GSC_HIP(hipMemcpyAsync(d_starts, starts.data(), sizeof(uint32_t) * starts.size(), hipMemcpyHostToDevice,
                               iter_stream));
        GSC_HIP(hipStreamSynchronize(iter_stream));  // the host vectors go out of scope
    }
    plan_dirty = false;
}

void gsc_server::enqueue_step() {
    const uint32_t N = static_cast<uint32_t>(active.size());
    const uint32_t* st = d_starts;
    for (int32_t it = 0; it < opts.iterations_per_step; ++it) {
        if (n_prefix_mode) {
            if (fused_prefix) {
                hipLaunchKernelGGL(prefix_plane_batch_kernel, dim3(grid[kPlane]), dim3(kPlaneThreads), plane_lds, iter_stream,
                                   d_inst, st + kPlane * (N + 1), N);
                hipLaunchKernelGGL(prefix_axis_batch_kernel, dim3(grid[kAxis2]), dim3(256), 0, iter_stream, d_inst,
                                   st + kAxis2 * (N + 1), N, 2);
                hipLaunchKernelGGL(prefix_last_axis_batch_kernel, dim3(grid[kLast]), dim3(256), 0, iter_stream, d_inst,
                                   st + kLast * (N + 1), N);
            } else {
                hipLaunchKernelGGL(prefix_rows_batch_kernel, dim3(grid[kRows]), dim3(256), 0, iter_stream, d_inst,
                                   st + kRows * (N + 1), N);
                for (int axis = 1; axis <= 3; ++axis) {
                    const uint32_t s = kAxis1 + axis - 1;
                    hipLaunchKernelGGL(prefix_axis_batch_kernel, dim3(grid[s]), dim3(256), 0, iter_stream, d_inst,
                                       st + s * (N + 1), N, axis);
                }
                hipLaunchKernelGGL(coarse_batch_kernel, dim3(grid[kCoarseStage]), dim3(256), 0, iter_stream, d_inst,
                                   st + kCoarseStage * (N + 1), N);
                hipLaunchKernelGGL(unreached_bits_batch_kernel, dim3(grid[kBits]), dim3(256), 0, iter_stream, d_inst,
                                   st + kBits * (N + 1), N);
            }
            hipLaunchKernelGGL(chunk_flags_batch_kernel, dim3(grid[kChunk]), dim3(256), 0, iter_stream, d_inst,
                               st + kChunk * (N + 1), N);
            GSC_HIP(launch_satisfy_batch(d_inst, st + kSat * (N + 1), N, grid[kSat], static_cast<uint32_t>(opts.tile_wg),
                                         static_cast<uint32_t>(opts.tile_ib), false, iter_stream));
        }
        if (n_bits_mode) {
            hipLaunchKernelGGL(frontier_bits_batch_kernel, dim3(grid[kFrontierBits]), dim3(256), 0, iter_stream, d_inst,
                               st + kFrontierBits * (N + 1), N);
            hipLaunchKernelGGL(frontier_v_batch_kernel, dim3(grid[kFrontierV]), dim3(256), 0, iter_stream, d_inst,
                               st + kFrontierV * (N + 1), N);
            hipLaunchKernelGGL(frontier_plane_batch_kernel, dim3(grid[kFrontierPlane]), dim3(kFrontierThreads),
                               frontier_lds, iter_stream, d_inst, st + kFrontierPlane * (N + 1), N);
            GSC_HIP(launch_satisfy_batch(d_inst, st + kSatBits * (N + 1), N, grid[kSatBits],
                                         static_cast<uint32_t>(opts.tile_wg), static_cast<uint32_t>(opts.tile_ib), true,
                                         iter_stream));
        }
        hipLaunchKernelGGL(finish_batch_kernel, dim3(N), dim3(64), 0, iter_stream, d_inst, N, d_status);
        GSC_HIP(hipGetLastError());
    }
    GSC_HIP(hipMemcpyAsync(h_status, d_status, sizeof(BatchStatus) * N, hipMemcpyDeviceToHost, iter_stream));
    GSC_HIP(hipMemcpyAsync(h_records, d_records, sizeof(IterRecord) * kRing * opts.max_active, hipMemcpyDeviceToHost,
                           iter_stream));
}

// Prepare the admitted cases on the preparation stream (runs while a step is
// in flight on the iteration stream). Cases that prepared are appended to
// `prepared`; configuration errors fail their request; a case that does not
// fit the arena after all (fragmentation) goes back to the head of the queue.
void gsc_server::admit(std::vector<Request*>& admitted, std::vector<Request*>& prepared) {
    std::vector<Request*> launched;
    for (Request* r : admitted) {
        r->solver = std::make_unique<gsc_solver>();
        gsc_solver& s = *r->solver;
        copy_case(&s, r->spec->cfg);
        s.opts = r->spec->opts;
        s.arena = arena.get();
        s.stream = prep_stream;
        s.owns_stream = false;
        s.host_words = h_words + kHostWords * r->slot;
        s.owns_host_words = false;
        try {
            // Masks, valid prefix and abstraction of every admitted case are
            // enqueued before the first wait on the preparation stream.
            s.setup_launch();
            s.launch_abstraction();
            launched.push_back(r);
        } catch (const ConfigError& e) {
            std::lock_guard<std::mutex> lock(m);
            release_slot(r);
            fail(r, GSC_ERR_CONFIG, e.what());
        } catch (const DeviceError& e) {
            const bool oom = status_of(e) == GSC_ERR_OUT_OF_MEMORY;
            std::lock_guard<std::mutex> lock(m);
            r->solver.reset();  // partial device state back to the arena
            release_slot(r);
            if (oom && !(active.empty() && launched.empty())) {
                r->state = GSC_REQUEST_QUEUED;  // fragmentation: retry when others have left
                queue.push_front(r);
            } else {
                fail(r, status_of(e), e.what());
            }
        }
    }
    for (Request* r : launched) {
        gsc_solver& s = *r->solver;
        try {
            // May regrow the overflow table from the arena (more escaped boxes
            // than the first table held); admission counts the arena's bytes in
            // use, so the regrown table is charged like every other buffer.
            s.setup_check();
            s.finish_abstraction();
            s.seed_async();
            {
                std::lock_guard<std::mutex> lock(m);
                r->est_bytes = s.bytes_allocated;  // what the request holds in the arena now
            }
            prepared.push_back(r);
        } catch (const ConfigError& e) {
            std::lock_guard<std::mutex> lock(m);
            r->solver.reset();
            release_slot(r);
            fail(r, GSC_ERR_CONFIG, e.what());
        } catch (const DeviceError& e) {
            // The regrown table did not fit: like a set-up that did not fit, the
            // request waits for running cases to leave, or fails alone.
            const bool oom = status_of(e) == GSC_ERR_OUT_OF_MEMORY;
            std::lock_guard<std::mutex> lock(m);
            r->solver.reset();
            release_slot(r);
            if (oom && !active.empty()) {
                r->state = GSC_REQUEST_QUEUED;
                queue.push_front(r);
            } else {
                fail(r, status_of(e), e.what());
            }
        }
    }
    GSC_HIP(hipStreamSynchronize(prep_stream));
}

// A case stopped: results to the host, device state back to the arena.
void gsc_server::retire(Request* r, const BatchStatus& st) {
    gsc_solver& s = *r->solver;
    r->t_stop = Clock::now();
    gsc_solve_summary& cs = r->summary;
    const uint64_t pt = cs.pairs_tested, bc = cs.boxes_counted, ts = cs.test_slots;
    std::memset(&cs, 0, sizeof(cs));
    cs.iterations = st.iterations;
    cs.converged = st.done == kDoneConverged ? 1 : 0;
    std::snprintf(cs.message, sizeof(cs.message), "%s", done_message(st.done));
    cs.reachable_states = st.total_reachable;
    cs.certified_candidate_states = st.total_certified;
    cs.candidate_states = s.candidate_count;
    cs.abstraction_ms = s.abstraction_kernel_ms;
    cs.solve_ms = ms_between(r->t_admit, r->t_stop);
    cs.pairs_tested = pt;
    cs.boxes_counted = bc;
    cs.test_slots = ts;
    cs.pair_bytes_read = pt * sizeof(uint32_t);
    cs.frontier = s.frontier_mode();
    s.summary = cs;
    s.iter_stats = r->stats;
    s.timings = gsc_kernel_timings{};
    s.timings.abstraction_kernel_ms = s.abstraction_kernel_ms;
    s.solved = true;
    std::lock_guard<std::mutex> lock(m);
    r->in_retire = true;
    ++retiring;
    retire_q.push_back(r);
    cv_retire.notify_one();
}

void gsc_server::retire_loop() {
    try {
        GSC_HIP(hipSetDevice(opts.device));
    } catch (const std::exception&) {
    }
    for (;;) {
        Request* r = nullptr;
        {
            std::unique_lock<std::mutex> lock(m);
            cv_retire.wait(lock, [&] { return retire_stop || !retire_q.empty(); });
            if (retire_q.empty()) return;
            r = retire_q.front();
            retire_q.pop_front();
        }
        finish_retire(r);
    }
}

// Retire thread: the stopped case's results to host memory on the retire
// stream (the worker synchronised the step that stopped it), then its HBM back
// to the arena and the request done.
void gsc_server::finish_retire(Request* r) {
    gsc_solver& s = *r->solver;
    const uint64_t S = r->shape.state_count;
    // Pinned pool block when one is free (the copies then run at the link's
    // speed and nothing is zero-filled), pageable memory otherwise.
    uint8_t* block = result_pool ? static_cast<uint8_t*>(result_pool->allocate(11 * S)) : nullptr;
    if (block) {
        r->pool = result_pool.get();
    } else {
        r->fallback.resize(11 * S);
        block = r->fallback.data();
    }
    r->results = block;
    gsc_result_buffers out{};
    out.controller = reinterpret_cast<uint32_t*>(block);
    out.reach_step = reinterpret_cast<uint32_t*>(block) + S;
    out.valid_mask = block + 8 * S;
    out.candidate_mask = block + 9 * S;
    out.reachable_mask = block + 10 * S;
    std::memset(&r->report, 0, sizeof(r->report));
    std::string error;
    const auto t0 = Clock::now();
    try {
        s.stream = retire_stream;
        s.download(&out, &r->report);
    } catch (const std::exception& e) {
        error = e.what();
    }
    if (trace)
        std::fprintf(stderr, "[serve] t=%.3f retire request %llu: %llu states copied in %.3f ms\n",
                     ms_between(t_created, t0), static_cast<unsigned long long>(r->id), static_cast<unsigned long long>(S),
                     ms_between(t0, Clock::now()));
    r->solver.reset();  // arena pieces released
    std::lock_guard<std::mutex> lock(m);
    release_slot(r);
    r->in_retire = false;
    --retiring;
    if (!error.empty()) {
        fail(r, GSC_ERR_DEVICE, error);
    } else {
        r->state = GSC_REQUEST_DONE;
        r->t_done = Clock::now();
        ++stats.completed;
        cv_done.notify_all();
    }
    cv_work.notify_one();  // HBM and a slot came back
}

void gsc_server::step_loop() {
    for (;;) {
        // 1. Admission, in arrival order, against slots and the HBM budget.
        std::vector<Request*> admitted;
        {
            std::unique_lock<std::mutex> lock(m);
            if (active.empty() && !stopping) cv_work.wait(lock, [&] { return stopping || !queue.empty(); });
            if (stopping) {  // shutting down: finish what runs, admit nothing more
                while (!queue.empty()) {
                    fail(queue.front(), GSC_ERR_BUSY, "server shutting down");
                    queue.pop_front();
                }
                if (active.empty()) return;
            }
            uint64_t free_bytes = arena->capacity() - arena->in_use();
            while (!queue.empty() && !free_slots.empty()) {
                Request* r = queue.front();
                if (r->est_bytes > arena->capacity()) {
                    queue.pop_front();
                    fail(r, GSC_ERR_OUT_OF_MEMORY,
                         "case needs " + std::to_string(r->est_bytes) + " bytes of HBM, above the server budget of " +
                             std::to_string(arena->capacity()));
                    continue;
                }
                if (r->est_bytes > free_bytes) break;  // wait for running cases to leave
                free_bytes -= r->est_bytes;
                queue.pop_front();
                r->slot = free_slots.back();
                free_slots.pop_back();
                r->state = GSC_REQUEST_RUNNING;
                r->t_admit = Clock::now();
                admitted.push_back(r);
            }
            if (active.empty() && admitted.empty() && !queue.empty()) {
                if (retiring > 0) {
                    // Stopped cases still hold HBM and slots until their results
                    // are on the host: wait for one of them.
                    const uint32_t before = retiring;
                    cv_work.wait(lock, [&] { return stopping || retiring < before; });
                    continue;
                }
                // Nothing runs and the head still does not fit: it never will.
                Request* r = queue.front();
                queue.pop_front();
                fail(r, GSC_ERR_OUT_OF_MEMORY, "case does not fit the free HBM budget");
                continue;
            }
        }
        // 2. One step of the admitted cases on the iteration stream...
        const auto t_step = Clock::now();
        const bool stepping = !active.empty();
        if (stepping) {
            if (plan_dirty) plan();
            enqueue_step();
        }
        const auto t_enqueued = Clock::now();
        // 3. ...while the newly admitted ones are prepared on the other stream.
        std::vector<Request*> prepared;
        if (!admitted.empty()) admit(admitted, prepared);
        const auto t_admitted = Clock::now();
        // 4. Collect the step: iteration history, stopped cases leave.
        if (stepping) {
            GSC_HIP(hipStreamSynchronize(iter_stream));
            if (trace)
                std::fprintf(stderr, "[serve] t=%.3f step %llu: %zu active, plan+enqueue %.3f ms, %zu admitted %.3f ms, wait %.3f ms\n",
                             ms_between(t_created, t_step), static_cast<unsigned long long>(stats.steps), active.size(),
                             ms_between(t_step, t_enqueued),
                             admitted.size(), ms_between(t_enqueued, t_admitted), ms_between(t_admitted, Clock::now()));
            ++stats.steps;
            std::vector<Request*> still;
            for (size_t p = 0; p < active.size(); ++p) {
                Request* r = active[p];
                const BatchStatus& st = h_status[p];
                for (uint32_t it = r->seen + 1; it <= st.iterations; ++it) {
                    const IterRecord& rec = h_records[static_cast<size_t>(r->slot) * kRing + it % kRing];
                    gsc_iteration_stats s;
                    std::memset(&s, 0, sizeof(s));
                    s.iteration = rec.iteration;
                    s.newly_reachable = rec.newly_reachable;
                    s.newly_certified = rec.newly_certified;
                    s.total_reachable = rec.total_reachable;
                    s.total_certified = rec.total_certified;
                    r->stats.push_back(s);
                    r->summary.pairs_tested += rec.pairs_tested;
                    r->summary.boxes_counted += rec.boxes_counted;
                    r->summary.test_slots += rec.test_slots;
                }
                r->seen = st.iterations;
                if (st.done)
                    retire(r, st);
                else
                    still.push_back(r);
            }
            if (still.size() != active.size()) plan_dirty = true;
            active.swap(still);
        }
        // 5. Prepared cases join the next step.
        if (!prepared.empty()) {
            for (Request* r : prepared) {
                r->seen = 0;
                if (r->solver->cfg.max_iterations == 0) {
                    // No iteration at all (reachability.hpp:53): the seeds are the result.
                    BatchStatus st{};
                    st.done = kDoneMaxIterations;
                    st.total_reachable = r->solver->candidate_count;
                    retire(r, st);
                    continue;
                }
                active.push_back(r);
            }
            plan_dirty = true;
        }
        {
            std::lock_guard<std::mutex> lock(m);
            stats.max_active_seen = std::max<uint32_t>(stats.max_active_seen, static_cast<uint32_t>(active.size()));
        }
    }
}

void gsc_server::run() {
    try {
        GSC_HIP(hipSetDevice(opts.device));
        step_loop();
    } catch (const std::exception& e) {
        fail_all(std::string("server stopped: ") + e.what());
    }
}

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

void gsc_server_default_options(gsc_server_options* o) {
    if (!o) return;
    std::memset(o, 0, sizeof(*o));
    o->device = 0;
    o->max_active = 64;
    o->iterations_per_step = 4;
    o->tile_wg = 512;
    o->tile_ib = 8;
    o->result_pool_mib = 1024;
    o->hbm_budget_bytes = 64ull << 30;
}

int gsc_server_create(const gsc_server_options* opts, gsc_server** out, char* err, size_t err_len) {
    if (!out) return GSC_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    gsc_server_options o;
    if (opts)
        o = *opts;
    else
        gsc_server_default_options(&o);
    if (o.max_active < 1 || o.max_active > 1024 || o.iterations_per_step < 1 ||
        o.iterations_per_step > static_cast<int32_t>(kRing) ||
        !satisfy_variant_supported(static_cast<uint32_t>(o.tile_wg), static_cast<uint32_t>(o.tile_ib)) ||
        o.hbm_budget_bytes == 0 || o.result_pool_mib < 0) {
        put(err, err_len, "invalid server options");
        return GSC_ERR_INVALID_ARGUMENT;
    }
    auto* s = new gsc_server();
    s->opts = o;
    s->trace = std::getenv("GSC_SERVE_TRACE") != nullptr;
    try {
        GSC_HIP(hipSetDevice(o.device));
        GSC_HIP(hipStreamCreate(&s->iter_stream));
        GSC_HIP(hipStreamCreate(&s->prep_stream));
        GSC_HIP(hipStreamCreate(&s->retire_stream));
        const uint32_t A = static_cast<uint32_t>(o.max_active);
        GSC_HIP(hipMalloc(reinterpret_cast<void**>(&s->d_inst), sizeof(BatchInstance) * A));
        GSC_HIP(hipMalloc(reinterpret_cast<void**>(&s->d_starts), sizeof(uint32_t) * kStages * (A + 1)));
        GSC_HIP(hipMalloc(reinterpret_cast<void**>(&s->d_status), sizeof(BatchStatus) * A));
        GSC_HIP(hipMalloc(reinterpret_cast<void**>(&s->d_records), sizeof(IterRecord) * kRing * A));
        GSC_HIP(hipHostMalloc(reinterpret_cast<void**>(&s->h_status), sizeof(BatchStatus) * A, hipHostMallocDefault));
        GSC_HIP(hipHostMalloc(reinterpret_cast<void**>(&s->h_records), sizeof(IterRecord) * kRing * A,
                              hipHostMallocDefault));
        GSC_HIP(hipHostMalloc(reinterpret_cast<void**>(&s->h_words), kHostWords * sizeof(uint32_t) * A,
                              hipHostMallocDefault));
        std::memset(s->h_words, 0, kHostWords * sizeof(uint32_t) * A);
        std::memset(s->h_status, 0, sizeof(BatchStatus) * A);
        std::memset(s->h_records, 0, sizeof(IterRecord) * kRing * A);
        s->arena = std::make_unique<DeviceArena>(o.device, o.hbm_budget_bytes);
        if (o.result_pool_mib > 0) {
            const uint64_t bytes = static_cast<uint64_t>(o.result_pool_mib) << 20;
            GSC_HIP(hipHostMalloc(&s->pinned_results, bytes, hipHostMallocDefault));
            s->result_pool = std::make_unique<DeviceArena>(s->pinned_results, bytes);
        }
        for (int k = static_cast<int>(A) - 1; k >= 0; --k) s->free_slots.push_back(k);
        s->stats.budget_bytes = s->arena->capacity();
        s->retirer = std::thread([s] { s->retire_loop(); });
        s->worker = std::thread([s] { s->run(); });
    } catch (const std::exception& e) {
        put(err, err_len, e.what());
        const int rc = dynamic_cast<const DeviceError*>(&e) ? status_of(*dynamic_cast<const DeviceError*>(&e))
                                                            : GSC_ERR_DEVICE;
        delete s;
        return rc;
    }
    *out = s;
    return GSC_OK;
}

int gsc_server_submit(gsc_server* s, const gsc_case_desc* cfg, uint64_t* id) {
    if (!s || !cfg) return GSC_ERR_INVALID_ARGUMENT;
    auto r = std::make_unique<Request>();
    r->t_submit = Clock::now();
    r->spec = std::make_unique<gsc_solver>();
    copy_case(r->spec.get(), *cfg);
    gsc_solver_default_options(&r->spec->opts);
    r->spec->opts.device = s->opts.device;
    r->spec->opts.time_kernels = 0;
    int rc = GSC_OK;
    std::string why;
    try {
        r->shape = make_shape(r->spec->cfg);
        if (r->spec->cfg.wrap_dim < 0 || r->spec->cfg.wrap_dim > 3) throw ConfigError("wrap_dim must be 0, 1, 2 or 3");
        r->est_bytes = gsc_solver::device_bytes(r->shape, r->spec->cfg);
    } catch (const std::exception& e) {
        rc = GSC_ERR_CONFIG;
        why = e.what();
    }
    std::lock_guard<std::mutex> lock(s->m);
    if (!s->fatal.empty() || s->stopping) return GSC_ERR_DEVICE;
    r->id = s->next_id++;
    if (id) *id = r->id;
    ++s->stats.submitted;
    Request* raw = r.get();
    s->requests.emplace(raw->id, std::move(r));
    if (rc != GSC_OK) {
        s->fail(raw, rc, why);
        return GSC_OK;  // the request exists and reports its error
    }
    s->queue.push_back(raw);
    s->cv_work.notify_one();
    return GSC_OK;
}

int gsc_server_wait(gsc_server* s, uint64_t id, double timeout_ms, gsc_request_info* info) {
    if (!s) return GSC_ERR_INVALID_ARGUMENT;
    std::unique_lock<std::mutex> lock(s->m);
    auto it = s->requests.find(id);
    if (it == s->requests.end()) return GSC_ERR_INVALID_ARGUMENT;
    Request* r = it->second.get();
    auto finished = [&] { return r->state == GSC_REQUEST_DONE || r->state == GSC_REQUEST_FAILED; };
    if (timeout_ms < 0)
        s->cv_done.wait(lock, finished);
    else if (!s->cv_done.wait_for(lock, std::chrono::duration<double, std::milli>(timeout_ms), finished))
        return GSC_ERR_BUSY;
    if (info) {
        std::memset(info, 0, sizeof(*info));
        info->id = r->id;
        info->state = r->state;
        info->status = r->status;
        info->state_count = r->shape.state_count;
        info->pair_count = r->shape.pair_count;
        info->device_bytes = r->est_bytes;
        if (r->t_admit != Clock::time_point{}) info->queued_ms = ms_between(r->t_submit, r->t_admit);
        info->latency_ms = ms_between(r->t_submit, r->t_done);
        info->summary = r->summary;
        put(info->error, sizeof(info->error), r->error);
    }
    return r->state == GSC_REQUEST_DONE ? GSC_OK : r->status;
}

int gsc_server_fetch(gsc_server* s, uint64_t id, const gsc_result_buffers* out, gsc_run_report* report) {
    if (!s) return GSC_ERR_INVALID_ARGUMENT;
    std::unique_ptr<Request> r;
    {
        std::lock_guard<std::mutex> lock(s->m);
        auto it = s->requests.find(id);
        if (it == s->requests.end()) return GSC_ERR_INVALID_ARGUMENT;
        if (it->second->state != GSC_REQUEST_DONE && it->second->state != GSC_REQUEST_FAILED) return GSC_ERR_BUSY;
        r = std::move(it->second);
        s->requests.erase(it);
    }
    if (r->state == GSC_REQUEST_FAILED) {
        if (report) {
            std::memset(report, 0, sizeof(*report));
            put(report->message, sizeof(report->message), r->error);
        }
        return r->status;
    }
    if (out) {
        const size_t S = r->shape.state_count;
        if (out->valid_mask) std::memcpy(out->valid_mask, r->valid(), S);
        if (out->candidate_mask) std::memcpy(out->candidate_mask, r->cand(), S);
        if (out->reachable_mask) std::memcpy(out->reachable_mask, r->reach(), S);
        if (out->controller) std::memcpy(out->controller, r->ctrl(), S * 4);
        if (out->reach_step) std::memcpy(out->reach_step, r->rstep(), S * 4);
        if (out->iteration_stats) {
            const size_t n = std::min<size_t>(out->iteration_stats_capacity, r->stats.size());
            std::memcpy(out->iteration_stats, r->stats.data(), n * sizeof(gsc_iteration_stats));
        }
    }
    if (report) {
        *report = r->report;
        report->iteration_stats_written =
            out && out->iteration_stats ? std::min<uint32_t>(out->iteration_stats_capacity, r->stats.size()) : 0;
        report->solve_ms = r->summary.solve_ms;
        report->total_ms = ms_between(r->t_submit, r->t_done);
    }
    return GSC_OK;
}

int gsc_server_get_stats(gsc_server* s, gsc_server_stats* out) {
    if (!s || !out) return GSC_ERR_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lock(s->m);
    *out = s->stats;
    out->queued = static_cast<uint32_t>(s->queue.size());
    out->active = static_cast<uint32_t>(s->opts.max_active - s->free_slots.size());
    out->budget_in_use = s->arena->in_use();
    out->budget_peak = s->arena->peak();
    return GSC_OK;
}

void gsc_server_destroy(gsc_server* s) { delete s; }

}  // extern "C"
