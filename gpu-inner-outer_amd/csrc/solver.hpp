// solver.hpp — one case resident on one device (gsc_solver) and the host-side
// helpers shared by the single-instance, batched and serving front ends.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <map>
#include <mutex>
#include <thread>
#include <vector>

#include "kernels.hpp"

namespace gsc_amd {

struct DeviceError : std::runtime_error {
    hipError_t code;
    DeviceError(hipError_t c, const std::string& what) : std::runtime_error(what), code(c) {}
};
struct ConfigError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

#define GSC_HIP(expr)                                                                             \
    do {                                                                                          \
        hipError_t e__ = (expr);                                                                  \
        if (e__ != hipSuccess) throw DeviceError(e__, hipGetErrorString(e__));                    \
    } while (0)

constexpr uint32_t kBatch = 8;          // iterations launched per host round trip
constexpr uint32_t kEventsPerIter = 4;  // start, after prefix, after satisfy, after finish
// Host words a solver reads back: two u64 mask counts, then the overflow count.
constexpr uint32_t kHostWords = 8, kWordMaskCounts = 0, kWordOverflow = 4;
constexpr uint64_t kExportSlice = 1ull << 22;         // pairs per export slice (37.7 MB of triplet)
constexpr uint64_t kExportSliceNoAlloc = 1ull << 20;  // the same without allocation (9.4 MB)

inline double ms_since(std::chrono::steady_clock::time_point t0) {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

inline void put(char* dst, size_t len, const std::string& s) {
    if (dst && len) std::snprintf(dst, len, "%s", s.c_str());
}

inline uint32_t blocks_for(uint64_t items, uint32_t per_block) {
    const uint64_t b = (items + per_block - 1) / per_block;
    if (b > 0x7FFFFFFFull) throw ConfigError("kernel launch grid exceeds the device limit");
    return static_cast<uint32_t>(b);
}

inline const char* done_message(uint32_t done) {
    switch (done) {
        case kDoneConverged: return "all candidate states obtained a controller";
        case kDoneStalled: return "iteration stalled before candidate closure";
        default: return "maximum iterations reached";
    }
}

// build_state_grid / build_input_grid / build_prefix_layout (grid.hpp:103-168, 253-269),
// with the reference's error texts.
inline gsc_case_shape make_shape(const gsc_case_desc& c) {
    gsc_case_shape sh;
    std::memset(&sh, 0, sizeof(sh));
    uint64_t total = 1;
    for (int d = 0; d < 4; ++d) {
        if (!(c.state_eta[d] > 0.0)) throw ConfigError("state eta must be positive");
        const double width = c.state_ub[d] - c.state_lb[d];
        if (!(width > 0.0)) throw ConfigError("state bounds must satisfy ub > lb");
        sh.state_shape[d] = static_cast<uint32_t>(std::ceil(width / c.state_eta[d]));
    }
    sh.state_strides[0] = 1;
    for (int d = 1; d < 4; ++d) sh.state_strides[d] = sh.state_strides[d - 1] * sh.state_shape[d - 1];
    for (int d = 0; d < 4; ++d) total *= sh.state_shape[d];
    sh.state_count = total;
    if (total > 0xFFFFFFFFull) throw ConfigError("state grid exceeds the selected state index type");
    for (int d = 0; d < 2; ++d) {
        if (!(c.input_eta[d] > 0.0)) throw ConfigError("input eta must be positive");
        const double width = c.input_ub[d] - c.input_lb[d];
        if (!(width > 0.0)) throw ConfigError("input bounds must satisfy ub > lb");
        sh.input_shape[d] = static_cast<uint32_t>(std::ceil(width / c.input_eta[d]));
    }
    sh.input_strides[0] = 1;
    sh.input_strides[1] = sh.input_shape[0];
    sh.input_count = static_cast<uint64_t>(sh.input_shape[0]) * sh.input_shape[1];
    if (sh.input_count > 0xFFFFFFFFull) throw ConfigError("input grid exceeds uint32_t");
    sh.pair_count = sh.state_count * sh.input_count;
    uint64_t pt = 1;
    for (int d = 0; d < 4; ++d) {
        sh.prefix_shape[d] = sh.state_shape[d] + 1;
        pt *= sh.prefix_shape[d];
    }
    sh.prefix_count = pt;
    return sh;
}

// estimate_memory_budget, abstraction.hpp:211-243 (u32 state index, u64 prefix counts).
inline gsc_memory_budget make_budget(const gsc_case_shape& sh) {
    gsc_memory_budget b;
    std::memset(&b, 0, sizeof(b));
    const uint64_t P = sh.pair_count, S = sh.state_count;
    b.pair_min_bytes = P * 4;
    b.pair_max_bytes = P * 4;
    b.pair_valid_bytes = P;
    b.pair_satisfied_bytes = P;
    b.reachable_prev_bytes = S;
    b.reachable_next_bytes = S;
    b.valid_mask_bytes = S;
    b.candidate_mask_bytes = S;
    b.controller_bytes = S * 4;
    b.reach_step_bytes = S * 4;
    b.prefix_reachable_bytes = sh.prefix_count * 8;
    b.prefix_valid_bytes = sh.prefix_count * 8;
    b.total_bytes = b.pair_min_bytes + b.pair_max_bytes + b.pair_valid_bytes + b.pair_satisfied_bytes +
                    b.reachable_prev_bytes + b.reachable_next_bytes + b.valid_mask_bytes + b.candidate_mask_bytes +
                    b.controller_bytes + b.reach_step_bytes + b.prefix_reachable_bytes + b.prefix_valid_bytes;
    return b;
}

// One hipMalloc'd block of HBM carved into 256-byte aligned pieces (first fit,
// adjacent free pieces merged on release). The serving loop admits requests
// against it so that no hipMalloc/hipFree runs while other solves are on the GPU.
class DeviceArena {
  public:
    DeviceArena(int device, uint64_t bytes) : capacity_(round_up(bytes)) {
        GSC_HIP(hipSetDevice(device));
        GSC_HIP(hipMalloc(&base_, capacity_));
        free_[0] = capacity_;
    }
    // Over device memory the caller owns (256-byte aligned): nothing is
    // allocated or freed.
    DeviceArena(void* base, uint64_t bytes) : base_(base), capacity_(bytes & ~uint64_t(255)), owns_(false) {
        free_[0] = capacity_;
    }
    ~DeviceArena() {
        if (base_ && owns_) (void)hipFree(base_);
    }
    DeviceArena(const DeviceArena&) = delete;
    DeviceArena& operator=(const DeviceArena&) = delete;

    static uint64_t round_up(uint64_t b) { return (std::max<uint64_t>(b, 1) + 255) & ~uint64_t(255); }

    // nullptr when no free piece is large enough.
    void* allocate(uint64_t bytes) {
        const uint64_t need = round_up(bytes);
        std::lock_guard<std::mutex> lock(m_);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < need) continue;
            const uint64_t off = it->first, size = it->second;
            free_.erase(it);
            if (size > need) free_[off + need] = size - need;
            used_[off] = need;
            in_use_ += need;
            peak_ = std::max(peak_, in_use_);
            return static_cast<char*>(base_) + off;
        }
        return nullptr;
    }

    void release(void* p) {
        if (!p) return;
        std::lock_guard<std::mutex> lock(m_);
        const uint64_t off = static_cast<uint64_t>(static_cast<char*>(p) - static_cast<char*>(base_));
        auto u = used_.find(off);
        if (u == used_.end()) return;
        uint64_t start = off, size = u->second;
        in_use_ -= size;
        used_.erase(u);
        auto next = free_.find(start + size);
        if (next != free_.end()) {
            size += next->second;
            free_.erase(next);
        }
        auto prev = free_.lower_bound(start);
        if (prev != free_.begin()) {
            --prev;
            if (prev->first + prev->second == start) {
                start = prev->first;
                size += prev->second;
                free_.erase(prev);
            }
        }
        free_[start] = size;
    }

    uint64_t capacity() const { return capacity_; }
    uint64_t in_use() {
        std::lock_guard<std::mutex> lock(m_);
        return in_use_;
    }
    uint64_t peak() {
        std::lock_guard<std::mutex> lock(m_);
        return peak_;
    }

  private:
    void* base_ = nullptr;
    uint64_t capacity_ = 0;
    uint64_t in_use_ = 0;
    uint64_t peak_ = 0;
    bool owns_ = true;
    std::map<uint64_t, uint64_t> free_;  // offset -> size
    std::map<uint64_t, uint64_t> used_;  // offset -> size
    std::mutex m_;
};

}  // namespace gsc_amd

using namespace gsc_amd;

struct gsc_solver {
    gsc_solver_options opts{};
    std::string name;
    std::vector<gsc_rect4> obstacles, targets;
    std::vector<uint8_t> allow;  // the caller's per-state constraint mask, when given
    gsc_case_desc cfg{};
    gsc_case_shape shape{};
    DevCase dc{};
    hipStream_t stream = nullptr;
    bool owns_stream = true;
    DeviceArena* arena = nullptr;  // when set, every device buffer comes from it
    std::vector<std::pair<void*, uint64_t>> allocations;  // (pointer, bytes) from alloc()
    uint64_t bytes_allocated = 0;

    uint8_t *valid = nullptr, *cand = nullptr, *reach = nullptr;
    uint32_t *ctrl = nullptr, *rstep = nullptr;
    uint32_t* prefix_valid = nullptr;
    uint32_t* prefix_unreached[2] = {nullptr, nullptr};  // by iteration parity
    uint32_t* coarse_unreached[2] = {nullptr, nullptr};  // their coarse samples
    uint32_t* coarse_valid = nullptr;
    uint64_t coarse_count = 0;                            // entries per coarse table
    uint32_t* hull = nullptr;                             // [S] state hulls
    uint4* chunk_hull = nullptr;                          // [n_tasks] 64-cell chunk hulls
    uint8_t* chunk_flags = nullptr;                       // [planes * chunk_flag_pitch] per iteration
    uint2 *hull2 = nullptr, *hull3 = nullptr;             // θ / v hulls per cell
    uint32_t* rec = nullptr;
    OverflowEntry* overflow = nullptr;
    uint32_t overflow_capacity = 0;
    uint32_t* overflow_count = nullptr;
    double2* cs = nullptr;
    double* vt = nullptr;
    double2* inc = nullptr;
    double* inc_sums = nullptr;  // [plane][input][4], see increment_sums_kernel
    int32_t* cell_offs = nullptr;  // [plane][input][8], see increment_sums_kernel
    unsigned long long* ubits = nullptr;
    unsigned long long* probe_or_bits = nullptr;  // [n_tasks + 1], plane-aligned (see chunk_flags_body)
    unsigned long long* valid_bits = nullptr;     // [S / 64 + 1], the valid states (set-up)
    // Frontier skipping from bitmasks (frontier_plane_kernel): the states
    // reached since the previous iteration, the frontier words, the keep /
    // always words and the plane hull rectangles.
    unsigned long long* ubits_fresh = nullptr;    // [S / 64 + 1]: reached since the previous iteration
    uint32_t* fresh_v = nullptr;                  // [planes * n_y * ceil(n_x / 32)] (frontier_v_kernel)
    unsigned long long* frontier_words = nullptr;  // [n_tasks + 1], plane-aligned
    unsigned long long* frontier_keep = nullptr;   // [n_tasks]
    unsigned long long* frontier_always = nullptr;  // [n_tasks]
    int4* plane_rect = nullptr;                    // [planes]
    uint32_t* probe_base = nullptr;  // [plane][I_pad], see pair_rows_kernel
    uint32_t* row_kind = nullptr;    // [plane][I_pad], see pair_rows_kernel
    uint4* row_box = nullptr;        // [plane][input], see pair_rows_kernel
    FastRow* fast_offs = nullptr;      // [plane][I_pad], see pair_rows_kernel
    uint32_t* plane_mixed = nullptr;   // [plane], see pair_rows_kernel
    uint2 *box2 = nullptr, *box3 = nullptr;
    double* rad = nullptr;
    uint32_t* fail2d = nullptr;
    gsc_rect4 *d_obstacles = nullptr, *d_targets = nullptr;
    unsigned long long* mask_counts = nullptr;
    IterCtl* ctl = nullptr;
    IterRecord* records = nullptr;
    IterCtl* host_ctl = nullptr;        // pinned, allocated by the first solve()
    IterRecord* host_records = nullptr; // pinned, allocated by the first solve()
    uint32_t* host_words = nullptr;     // pinned scratch words (overflow count, mask counts)
    // Pinned staging of the model tables (one buffer, each table 256-B aligned),
    // allocated by the first abstraction and kept; null without allocation.
    unsigned char* table_stage = nullptr;
    size_t table_stage_bytes = 0;
    bool owns_host_words = true;        // false when a server lends them from its pool
    bool owns_host_ctl = true;          // false when host_ctl / host_records are lent
    std::vector<hipEvent_t> events;

    // Host side of the model tables, rebuilt by every upload_tables().
    struct HostTables {
        std::vector<double2> cs;
        std::vector<double> vt, rad;
        std::vector<uint2> box2, box3, hull2, hull3;
        std::vector<uint32_t> fail;
    };
    HostTables tables;

    // Abstraction export staging (export_triplet), allocated by the first export.
    struct ExportStage {
        uint32_t *dmin = nullptr, *dmax = nullptr;  // device (from alloc)
        uint8_t* dval = nullptr;
        uint32_t *hmin = nullptr, *hmax = nullptr;  // pinned host
        uint8_t* hval = nullptr;
        hipEvent_t done = nullptr;
    };
    ExportStage xstage[2];
    uint64_t export_slice = 0;
    double export_ms = 0.0;
    // No allocation after set-up by anything but the arena: the null stream, no
    // HIP events, host words lent by the caller, no pinned export staging.
    bool no_alloc = false;

    uint64_t candidate_count = 0;
    // Satisfaction tile: cells per workgroup and inputs per block of record loads.
    uint32_t tile_wg = 512, tile_ib = 8;
    uint32_t plane_pitch = 0;  // LDS row pitch of the fused prefix build; 0: separate passes
    bool abstraction_ready = false;
    bool solved = false;

    std::chrono::steady_clock::time_point abstraction_t0{};
    double abstraction_kernel_ms = 0.0;
    double abstraction_wall_ms = 0.0;
    double setup_h2d_ms = 0.0;
    std::chrono::steady_clock::time_point setup_t0{};
    gsc_solve_summary summary{};
    std::vector<gsc_iteration_stats> iter_stats;
    gsc_kernel_timings timings{};
    std::string last_error;

    template <typename T>
    T* alloc(uint64_t count) {
        void* p = nullptr;
        const uint64_t bytes = std::max<uint64_t>(count, 1) * sizeof(T);
        if (arena) {
            p = arena->allocate(bytes);
            if (!p) throw DeviceError(hipErrorOutOfMemory, "device arena exhausted");
        } else {
            GSC_HIP(hipMalloc(&p, bytes));
        }
        allocations.emplace_back(p, bytes);
        bytes_allocated += bytes;
        return static_cast<T*>(p);
    }

    // Gives one buffer from alloc() back before the solver ends.
    void release(void* p) {
        auto it = std::find_if(allocations.begin(), allocations.end(), [p](const auto& a) { return a.first == p; });
        if (it == allocations.end()) return;
        bytes_allocated -= it->second;
        allocations.erase(it);
        if (stream || no_alloc) GSC_HIP(hipStreamSynchronize(stream));
        if (arena)
            arena->release(p);
        else
            GSC_HIP(hipFree(p));
    }

    ~gsc_solver() {
        if (stream || no_alloc) (void)hipStreamSynchronize(stream);
        for (const auto& a : allocations) {
            if (arena)
                arena->release(a.first);
            else
                (void)hipFree(a.first);
        }
        for (ExportStage& st : xstage) {
            if (st.hmin) (void)hipHostFree(st.hmin);
            if (st.hmax) (void)hipHostFree(st.hmax);
            if (st.hval) (void)hipHostFree(st.hval);
            if (st.done) (void)hipEventDestroy(st.done);
        }
        if (table_stage) (void)hipHostFree(table_stage);
        if (host_ctl && owns_host_ctl) (void)hipHostFree(host_ctl);
        if (host_records && owns_host_ctl) (void)hipHostFree(host_records);
        if (host_words && owns_host_words) (void)hipHostFree(host_words);
        for (hipEvent_t e : events) (void)hipEventDestroy(e);
        if (stream && owns_stream) (void)hipStreamDestroy(stream);
    }

    // Device bytes setup() allocates for a case of this shape (the arena
    // admission estimate; the overflow table may grow later).
    static uint64_t device_bytes(const gsc_case_shape& sh, const gsc_case_desc& cfg);
    // The OR probe bitmask is built (by the chunk flag pass, which runs where
    // frontier skipping does) and probed (by the stored-record satisfaction
    // kernels). Fixed for the solver's life: its options do not change.
    bool probe_or_used() const { return !opts.sweep_all_pairs && !opts.recompute_abstraction && dc.hull_enabled; }
    // Frontier skipping on bitmasks instead of prefix tables: where the OR
    // bitmask is used, every pair's box is settled by row scans of the
    // bitmasks — no record escaped and every compact box has at most sixteen
    // rows in at most two bitmask slices along θ and along v (other boxes are counted on the
    // unreached prefix table, which this mode does not build) — and a plane
    // fits the frontier kernel's LDS.
    // The frontier skipping a solve of this case runs (GSC_FRONTIER_*).
    uint32_t frontier_mode() const {
        if (opts.sweep_all_pairs || !dc.hull_enabled) return GSC_FRONTIER_NONE;
        return bits_frontier() ? GSC_FRONTIER_BITS : GSC_FRONTIER_PREFIX;
    }
    // The bitmask frontier's kernel arguments for this case.
    FrontierArgs frontier_args() const;
    bool bits_frontier() const {
        return probe_or_used() && overflow_used == 0 && ext1_used < 16u && slices_max <= 2u &&
               frontier_lds_bytes(dc) <= kFrontierLdsMax;
    }
    uint32_t overflow_used = 0;  // escaped records of the current abstraction
    uint32_t ext1_used = 0;      // largest y extent of its compact records
    uint32_t slices_max = 0;     // most bitmask slices along θ or v a compact box needs (upload_tables)

    void build_tables(std::vector<double2>& h_cs, std::vector<double>& h_vt, std::vector<uint2>& h_box2,
                      std::vector<uint2>& h_box3, std::vector<double>& h_rad, std::vector<uint32_t>& h_fail);
    void setup();         // setup_launch() + setup_check()
    void setup_launch();  // buffers, masks, valid prefix: enqueued on `stream`
    void setup_check();   // waits, then checks candidate inside valid
    // Table of a & !b; with ctl, into P[k & 1] for the coming iteration k.
    void build_prefix(const uint8_t* a, const uint8_t* b, uint32_t* P0, uint32_t* P1, const IterCtl* ctl_or_null);
    // Coarse sample of P0 / P1 into C0 / C1 (by iteration parity with ctl).
    void build_coarse(const uint32_t* P0, const uint32_t* P1, uint32_t* C0, uint32_t* C1, const IterCtl* ctl_or_null);
    void upload_tables();         // model tables, host and device, enqueued on `stream`
    void build_abstraction();     // launch_abstraction() + finish_abstraction()
    void launch_abstraction(bool rerun = false);  // enqueue on `stream`, no wait
    void finish_abstraction();    // wait; rerun once with a larger overflow table if needed
    void seed_async();            // fixpoint seeds and control reset, enqueued on `stream`
    void solve();
    void download(const gsc_result_buffers* out, gsc_run_report* report);
    void export_triplet(const gsc_result_buffers* out);
};

inline int status_of(const DeviceError& e) {
    return e.code == hipErrorOutOfMemory || e.code == hipErrorMemoryAllocation ? GSC_ERR_OUT_OF_MEMORY : GSC_ERR_DEVICE;
}

inline void copy_case(gsc_solver* s, const gsc_case_desc& c) {
    s->cfg = c;
    s->name = c.name ? c.name : "";
    s->obstacles.assign(c.obstacles, c.obstacles + (c.obstacles ? c.obstacle_count : 0));
    s->targets.assign(c.extra_targets, c.extra_targets + (c.extra_targets ? c.extra_target_count : 0));
    s->allow.clear();
    if (c.state_allow) {
        // The caller's mask holds one byte per state; a grid that has no valid
        // shape fails again, with its message, in setup().
        try {
            const gsc_case_shape sh = make_shape(c);
            s->allow.assign(c.state_allow, c.state_allow + sh.state_count);
        } catch (const ConfigError&) {
        }
    }
    s->cfg.state_allow = s->allow.empty() ? nullptr : s->allow.data();
    s->cfg.name = s->name.c_str();
    s->cfg.obstacles = s->obstacles.empty() ? nullptr : s->obstacles.data();
    s->cfg.obstacle_count = static_cast<uint32_t>(s->obstacles.size());
    s->cfg.extra_targets = s->targets.empty() ? nullptr : s->targets.data();
    s->cfg.extra_target_count = static_cast<uint32_t>(s->targets.size());
}
