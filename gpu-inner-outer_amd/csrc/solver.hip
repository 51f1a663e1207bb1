// solver.hip — host orchestration of the gfx950 backend and the C ABI.
//
// A gsc_solver owns one case on one device: its tables, masks, pair records,
// prefix tables and iteration state, all resident in HBM. The fixpoint loop
// (kernels.cu:654-767 / reachability.hpp:30-155) is launched in batches of
// iterations with no host synchronisation inside a batch: every kernel reads a
// device-side `done` flag and returns at once after the stop test fired, so a
// batch past the last iteration costs a few empty launches and the stop
// iteration and message are exactly the reference's.
#include "solver.hpp"


// Per cell of one axis (θ or v), the union of the box ranges over the inputs of
// the other axis: table[cell * n_in + k] -> one (lo, hi) range covering them all
// (lo > hi wraps, on the wrap axis only). Rejected ranges add nothing.
static std::vector<uint2> cell_hulls(const std::vector<uint2>& table, uint32_t n, uint32_t n_in, bool wraps) {
    std::vector<uint2> out(n, uint2{0, n - 1});
    std::vector<uint8_t> cov(n);
    for (uint32_t cell = 0; cell < n; ++cell) {
        std::fill(cov.begin(), cov.end(), 0);
        for (uint32_t k = 0; k < n_in; ++k) {
            const uint2 r = table[static_cast<size_t>(cell) * n_in + k];
            if (r.x == kBoxReject) continue;
            if (r.x <= r.y) {
                for (uint32_t t = r.x; t <= r.y; ++t) cov[t] = 1;
            } else {  // wrapped range
                for (uint32_t t = r.x; t < n; ++t) cov[t] = 1;
                for (uint32_t t = 0; t <= r.y; ++t) cov[t] = 1;
            }
        }
        uint32_t first = n, last = 0, count = 0;
        for (uint32_t t = 0; t < n; ++t)
            if (cov[t]) {
                first = std::min(first, t);
                last = t;
                ++count;
            }
        if (count == 0 || count == n) continue;  // nothing (never tested) or everything: full range
        if (!wraps) {
            out[cell] = uint2{first, last};
            continue;
        }
        // Wrapping axis: the complement of the longest circular run of uncovered cells.
        uint32_t best_len = 0, best_start = 0;
        for (uint32_t t = 0; t < n; ++t) {
            if (cov[t] || cov[(t + n - 1) % n] == 0) continue;  // start of an uncovered run
            uint32_t len = 0;
            while (len < n && !cov[(t + len) % n]) ++len;
            if (len > best_len) {
                best_len = len;
                best_start = t;
            }
        }
        out[cell] = uint2{(best_start + best_len) % n, (best_start + n - 1) % n};
    }
    return out;
}

// Host tables. The θ trajectory depends on (θ cell, ω) only and the v trajectory
// on (v cell, a) only (model.hpp:35-45), so cos θ_k, sin θ_k and v_k are filled
// here once per combination with the host libm, in the reference's order:
//   θ_{k+1} = wrap_angle(θ_k + dt*ω),  v_{k+1} = v_k + dt*a.
// The θ and v cell ranges of the successor box follow from θ_K, v_K and the
// constant radii h2 + T*dw2, h3 + T*dw3 (model.hpp:74-75, 84-89, 107-111).
void gsc_solver::build_tables(std::vector<double2>& h_cs, std::vector<double>& h_vt, std::vector<uint2>& h_box2,
                              std::vector<uint2>& h_box3, std::vector<double>& h_rad, std::vector<uint32_t>& h_fail) {
    const uint32_t n2 = dc.shape[2], n3 = dc.shape[3], na = dc.n_a, nw = dc.n_w, K = dc.K;
    const double T = cfg.sample_time;
    const double dt = dc.dt;
    const double h[4] = {0.5 * cfg.state_eta[0], 0.5 * cfg.state_eta[1], 0.5 * cfg.state_eta[2],
                         0.5 * cfg.state_eta[3]};
    h_cs.assign(static_cast<size_t>(n2) * nw * K, double2{0.0, 0.0});
    h_box2.assign(static_cast<size_t>(n2) * nw, uint2{0, 0});
    const double r2 = h[2] + T * cfg.disturbance_half_width[2];
    for (uint32_t iz = 0; iz < n2; ++iz) {
        const double th0 = cell_center(cfg.state_lb[2], iz, cfg.state_eta[2]);
        for (uint32_t iw = 0; iw < nw; ++iw) {
            const double omega = cell_center(cfg.input_lb[1], iw, cfg.input_eta[1]);
            double th = th0;
            double2* row = &h_cs[(static_cast<size_t>(iz) * nw + iw) * K];
            for (uint32_t k = 0; k < K; ++k) {
                row[k].x = std::cos(th);
                row[k].y = std::sin(th);
                th = wrap_angle(th + dt * omega);
            }
            uint32_t lo = 0, hi = 0;
            if (dim_to_cells(dc, 2, th - r2, th + r2, &lo, &hi))
                h_box2[static_cast<size_t>(iz) * nw + iw] = uint2{lo, hi};
            else
                h_box2[static_cast<size_t>(iz) * nw + iw] = uint2{kBoxReject, kBoxReject};
        }
    }
    h_vt.assign(static_cast<size_t>(n3) * na * K, 0.0);
    h_box3.assign(static_cast<size_t>(n3) * na, uint2{0, 0});
    h_rad.assign(static_cast<size_t>(n3) * 2, 0.0);
    const double r3 = h[3] + T * cfg.disturbance_half_width[3];
    for (uint32_t iv = 0; iv < n3; ++iv) {
        const double v0 = cell_center(cfg.state_lb[3], iv, cfg.state_eta[3]);
        const double abs_v = std::fabs(v0);
        const double jfx0 = h[0] + T * abs_v * h[2] + T * h[3];
        const double jfx1 = h[1] + T * abs_v * h[2] + T * h[3];
        h_rad[2 * iv] = jfx0 + T * cfg.disturbance_half_width[0];
        h_rad[2 * iv + 1] = jfx1 + T * cfg.disturbance_half_width[1];
        for (uint32_t ia = 0; ia < na; ++ia) {
            const double acc = cell_center(cfg.input_lb[0], ia, cfg.input_eta[0]);
            double v = v0;
            double* row = &h_vt[(static_cast<size_t>(iv) * na + ia) * K];
            for (uint32_t k = 0; k < K; ++k) {
                row[k] = v;
                v += dt * acc;
            }
            uint32_t lo = 0, hi = 0;
            if (dim_to_cells(dc, 3, v - r3, v + r3, &lo, &hi))
                h_box3[static_cast<size_t>(iv) * na + ia] = uint2{lo, hi};
            else
                h_box3[static_cast<size_t>(iv) * na + ia] = uint2{kBoxReject, kBoxReject};
        }
    }
    h_fail.clear();
    if (cfg.constraint_type != GSC_CONSTRAINT_NONE) {
        const uint32_t n0 = dc.shape[0], n1 = dc.shape[1];
        const size_t pitch = size_t{1} << dc.fail_shift;
        h_fail.assign((n1 + 1) * pitch, 0u);
        for (uint32_t iy = 0; iy < n1; ++iy) {
            const double yc = cell_center(cfg.state_lb[1], iy, cfg.state_eta[1]);
            uint32_t row = 0;
            for (uint32_t ix = 0; ix < n0; ++ix) {
                const double xc = cell_center(cfg.state_lb[0], ix, cfg.state_eta[0]);
                const bool ok = constraint_ok(cfg.constraint_type, cfg.hyperbolic_a, cfg.hyperbolic_b,
                                              cfg.hyperbolic_c, cfg.elliptic_a, cfg.elliptic_b, xc, yc);
                row += ok ? 0u : 1u;
                h_fail[static_cast<size_t>(iy + 1) * pitch + ix + 1] = h_fail[static_cast<size_t>(iy) * pitch + ix + 1] + row;
            }
        }
    }
}

// Shape, device buffers, masks and the valid prefix, enqueued on `stream`; the
// candidate-inside-valid check (abstraction.hpp:261-265) waits in setup_check.
void gsc_solver::setup_launch() {
    if (cfg.wrap_dim < 0 || cfg.wrap_dim > 3) throw ConfigError("wrap_dim must be 0, 1, 2 or 3");
    shape = make_shape(cfg);
    if (shape.prefix_count > 0xFFFFFFFFull) throw ConfigError("state grid too large for the 32-bit prefix layout");
    if (shape.input_count > (1u << 19)) throw ConfigError("input grid larger than 524288 inputs is not supported");
    if (shape.state_count * 4 >= 0xFFFFFFFFull) throw ConfigError("state grid too large for one launch");

    dc = DevCase{};
    for (int d = 0; d < 4; ++d) {
        dc.lb[d] = cfg.state_lb[d];
        dc.ub[d] = cfg.state_ub[d];
        dc.eta[d] = cfg.state_eta[d];
        dc.lb_minus_eps[d] = cfg.state_lb[d] - kEps;
        dc.ub_plus_eps[d] = cfg.state_ub[d] + kEps;
        dc.shape[d] = shape.state_shape[d];
        dc.stride[d] = static_cast<uint32_t>(shape.state_strides[d]);
    }
    dc.pstride[0] = 1;
    for (int d = 1; d < 4; ++d) dc.pstride[d] = dc.pstride[d - 1] * shape.prefix_shape[d - 1];
    dc.wrap_dim = cfg.wrap_dim;
    dc.wrap_span_minus_eps = (cfg.state_ub[cfg.wrap_dim] - cfg.state_lb[cfg.wrap_dim]) - kEps;
    dc.S = static_cast<uint32_t>(shape.state_count);
    dc.I = static_cast<uint32_t>(shape.input_count);
    dc.I_pad = (dc.I + kInputBlock - 1) / kInputBlock * kInputBlock;
    dc.PQ = shape.state_shape[0] * shape.state_shape[1];
    if (static_cast<uint64_t>(dc.I_pad) * dc.PQ >= 0xFFFFFFFFull)
        throw ConfigError("input count x plane size exceeds the 32-bit record offset");
    dc.chunks_per_plane = (dc.PQ + 63) / 64;
    {
        const uint64_t tasks = static_cast<uint64_t>(dc.chunks_per_plane) * shape.state_shape[2] * shape.state_shape[3];
        if (tasks * 64 >= 0xFFFFFFFFull) throw ConfigError("state grid too large for one launch");
        // Probe rows are 32-bit bit offsets over both probe bitmasks.
        if (((static_cast<uint64_t>(dc.S) + 63) / 64 + 2 + tasks) * 64 >= 0xFFFFFFFFull)
            throw ConfigError("state grid too large for one launch");
        dc.n_tasks = static_cast<uint32_t>(tasks);
        dc.groups_per_plane = (dc.PQ + 255) / 256;
        dc.n_groups = dc.groups_per_plane * shape.state_shape[2] * shape.state_shape[3];
    }
    dc.n_a = shape.input_shape[0];
    dc.n_w = shape.input_shape[1];
    dc.K = cfg.integration_substeps;
    dc.fail_shift = 0;
    while ((1u << dc.fail_shift) < dc.shape[0] + 1) ++dc.fail_shift;
    if ((static_cast<uint64_t>(dc.shape[1]) + 1) << (dc.fail_shift + 2) > 0xFFFFFFFFull)
        throw ConfigError("x/y grid too large for the 32-bit constraint table offsets");
    dc.dt = cfg.sample_time / static_cast<double>(cfg.integration_substeps);
    dc.constraint_type = cfg.constraint_type;
    dc.hyp_a = cfg.hyperbolic_a;
    dc.hyp_b = cfg.hyperbolic_b;
    dc.hyp_c = cfg.hyperbolic_c;
    dc.ell_a = cfg.elliptic_a;
    dc.ell_b = cfg.elliptic_b;
    dc.candidate = cfg.candidate;
    dc.map = cfg.map;
    dc.band_enabled = cfg.target_band_enabled;
    dc.band_c = cfg.target_band_c;
    dc.band_w = cfg.target_band_w;
    // Hull skipping needs x and y boxes that never wrap.
    dc.hull_enabled = (cfg.wrap_dim == 2 || cfg.wrap_dim == 3) ? 1 : 0;
    dc.cshape[0] = (dc.shape[0] + kCoarse - 1) / kCoarse + 1;
    dc.cshape[1] = (dc.shape[1] + kCoarse - 1) / kCoarse + 1;
    dc.cstride[0] = 1;
    dc.cstride[1] = dc.cshape[0];
    dc.cstride[2] = dc.cshape[0] * dc.cshape[1];
    dc.cstride[3] = dc.cstride[2] * (dc.shape[2] + 1);
    coarse_count = static_cast<uint64_t>(dc.cstride[3]) * (dc.shape[3] + 1);
    dc.n_obstacles = static_cast<uint32_t>(obstacles.size());
    dc.n_targets = static_cast<uint32_t>(targets.size());

    GSC_HIP(hipSetDevice(opts.device));
    if (no_alloc) {
        // The null stream, no events, no pinned memory: nothing is created.
        if (!arena || !host_words || !host_ctl || !host_records)
            throw ConfigError("a solver without allocation needs an arena and lent host words");
        stream = nullptr;
        owns_stream = false;
        opts.time_kernels = 0;
    } else if (!stream) {
        GSC_HIP(hipStreamCreate(&stream));
        owns_stream = true;
    }
    const uint64_t S = shape.state_count, PT = shape.prefix_count;
    valid = alloc<uint8_t>(S);
    cand = alloc<uint8_t>(S);
    reach = alloc<uint8_t>(S);
    ctrl = alloc<uint32_t>(S);
    rstep = alloc<uint32_t>(S);
    prefix_valid = alloc<uint32_t>(PT);
    prefix_unreached[0] = alloc<uint32_t>(PT);
    prefix_unreached[1] = alloc<uint32_t>(PT);
    coarse_unreached[0] = alloc<uint32_t>(coarse_count);
    coarse_unreached[1] = alloc<uint32_t>(coarse_count);
    coarse_valid = alloc<uint32_t>(coarse_count);
    hull = alloc<uint32_t>(S);
    chunk_hull = alloc<uint4>(dc.n_tasks);
    const uint64_t n_planes = static_cast<uint64_t>(shape.state_shape[2]) * shape.state_shape[3];
    chunk_flags = alloc<uint8_t>(n_planes * chunk_flag_pitch(dc));
    // Recompute mode keeps no per-pair state at all.
    if (!opts.recompute_abstraction) rec = alloc<uint32_t>(n_planes * dc.I_pad * dc.PQ);
    probe_base = alloc<uint32_t>(n_planes * dc.I_pad);
    row_kind = alloc<uint32_t>(n_planes * dc.I_pad);
    // The plain probe bitmask (+1 word: the row probe reads one word past),
    // then the plane-aligned OR bitmask (one word per chunk, +1), then the
    // valid bitmask (as the plain one).
    ubits = alloc<unsigned long long>(2ull * probe_bits_words(dc) + dc.n_tasks + 1);
    probe_or_bits = ubits + probe_bits_words(dc);
    valid_bits = probe_or_bits + dc.n_tasks + 1;
    ubits_fresh = alloc<unsigned long long>(probe_bits_words(dc));
    fresh_v = alloc<uint32_t>(n_planes * dc.shape[1] * ((dc.shape[0] + 31u) / 32u));
    frontier_words = alloc<unsigned long long>(static_cast<uint64_t>(dc.n_tasks) + 1);
    frontier_keep = alloc<unsigned long long>(dc.n_tasks);
    frontier_always = alloc<unsigned long long>(dc.n_tasks);
    plane_rect = alloc<int4>(n_planes);
    overflow_capacity = opts.recompute_abstraction ? 1u : 1u << 16;
    overflow = alloc<OverflowEntry>(overflow_capacity);
    overflow_count = alloc<uint32_t>(2);  // escaped records, then the tallest compact box's y extent
    mask_counts = alloc<unsigned long long>(2);
    ctl = alloc<IterCtl>(1);
    records = alloc<IterRecord>(kRing);
    if (!host_words) {
        GSC_HIP(hipHostMalloc(reinterpret_cast<void**>(&host_words), kHostWords * sizeof(uint32_t), hipHostMallocDefault));
        owns_host_words = true;
    }
    std::memset(host_words, 0, kHostWords * sizeof(uint32_t));
    if (!no_alloc) {
        events.resize(opts.time_kernels ? kBatch * kEventsPerIter + 2 : 2);
        for (auto& e : events) GSC_HIP(hipEventCreate(&e));
    }

    // The model tables (trigonometry, θ / v boxes, radii, constraint table,
    // trajectory increments) are filled by every launch_abstraction: they are
    // part of the abstraction's work, which the reference does inside its
    // timed kernel (kernels.cu:164-217, 609-641). Sizes as in build_tables.
    const uint64_t n2 = dc.shape[2], n3 = dc.shape[3];
    cs = alloc<double2>(n2 * dc.n_w * dc.K);
    vt = alloc<double>(n3 * dc.n_a * dc.K);
    box2 = alloc<uint2>(n2 * dc.n_w);
    box3 = alloc<uint2>(n3 * dc.n_a);
    hull2 = alloc<uint2>(n2);
    hull3 = alloc<uint2>(n3);
    rad = alloc<double>(2 * n3);
    if (cfg.constraint_type != GSC_CONSTRAINT_NONE)
        fail2d = alloc<uint32_t>((static_cast<uint64_t>(dc.shape[1]) + 1) << dc.fail_shift);
    const uint64_t n_inc = static_cast<uint64_t>(dc.shape[2]) * dc.shape[3] * dc.I * dc.K;
    if (n_inc * sizeof(double2) > (4ull << 30)) throw ConfigError("trajectory increment table exceeds 4 GiB");
    inc = alloc<double2>(n_inc);
    inc_sums = alloc<double>(4ull * dc.shape[2] * dc.shape[3] * dc.I);
    cell_offs = alloc<int32_t>(8ull * dc.shape[2] * dc.shape[3] * dc.I);
    row_box = alloc<uint4>(static_cast<uint64_t>(dc.shape[2]) * dc.shape[3] * dc.I);
    fast_offs = alloc<FastRow>(static_cast<uint64_t>(dc.shape[2]) * dc.shape[3] * dc.I_pad);
    plane_mixed = alloc<uint32_t>(static_cast<uint64_t>(dc.shape[2]) * dc.shape[3]);
    d_obstacles = alloc<gsc_rect4>(obstacles.size());
    d_targets = alloc<gsc_rect4>(targets.size());
    uint8_t* d_allow = allow.empty() ? nullptr : alloc<uint8_t>(S);

    const auto t0 = std::chrono::steady_clock::now();
    if (!obstacles.empty())
        GSC_HIP(hipMemcpyAsync(d_obstacles, obstacles.data(), obstacles.size() * sizeof(gsc_rect4),
                               hipMemcpyHostToDevice, stream));
    if (!targets.empty())
        GSC_HIP(hipMemcpyAsync(d_targets, targets.data(), targets.size() * sizeof(gsc_rect4), hipMemcpyHostToDevice,
                               stream));
    if (d_allow) GSC_HIP(hipMemcpyAsync(d_allow, allow.data(), S, hipMemcpyHostToDevice, stream));
    // Padding rows of both prefix tables stay zero for the life of the solver.
    GSC_HIP(hipMemsetAsync(prefix_valid, 0, PT * sizeof(uint32_t), stream));
    GSC_HIP(hipMemsetAsync(prefix_unreached[0], 0, PT * sizeof(uint32_t), stream));
    GSC_HIP(hipMemsetAsync(prefix_unreached[1], 0, PT * sizeof(uint32_t), stream));
    GSC_HIP(hipMemsetAsync(coarse_unreached[0], 0, coarse_count * sizeof(uint32_t), stream));
    GSC_HIP(hipMemsetAsync(coarse_unreached[1], 0, coarse_count * sizeof(uint32_t), stream));
    // The padding bytes of each plane's chunk-flag row stay zero.
    GSC_HIP(hipMemsetAsync(chunk_flags, 0, static_cast<uint64_t>(n_planes) * chunk_flag_pitch(dc), stream));
    {
        // The fused prefix build keeps one plane's 2D prefix in LDS.
        const uint32_t pitch = (dc.shape[0] + 63u) / 64u * 64u;
        plane_pitch = static_cast<uint64_t>(pitch) * dc.shape[1] * sizeof(uint32_t) <= kPlaneLdsMax ? pitch : 0u;
    }
    GSC_HIP(hipMemsetAsync(mask_counts, 0, 2 * sizeof(unsigned long long), stream));

    // Masks on the device (the reference builds them in a serial host loop even in
    // GPU mode, abstraction.hpp:257-265).
    hipLaunchKernelGGL(masks_kernel, dim3(std::min<uint64_t>(blocks_for(S, 256), 2048)), dim3(256), 0, stream, dc, d_obstacles, d_targets, d_allow, valid,
                       cand, mask_counts);
    GSC_HIP(hipGetLastError());
    unsigned long long* counts = reinterpret_cast<unsigned long long*>(host_words + kWordMaskCounts);
    GSC_HIP(hipMemcpyAsync(counts, mask_counts, 2 * sizeof(unsigned long long), hipMemcpyDeviceToHost, stream));
    // The valid prefix and bitmask do not wait for the check: a failed check
    // discards them.
    build_prefix(valid, nullptr, prefix_valid, prefix_valid, nullptr);
    hipLaunchKernelGGL(valid_bits_kernel, dim3(blocks_for(S, 256)), dim3(256), 0, stream, static_cast<uint32_t>(S),
                       valid, valid_bits);
    GSC_HIP(hipGetLastError());
    build_coarse(prefix_valid, prefix_valid, coarse_valid, coarse_valid, nullptr);
    setup_t0 = t0;
}

void gsc_solver::setup_check() {
    GSC_HIP(hipStreamSynchronize(stream));
    setup_h2d_ms = ms_since(setup_t0);
    const unsigned long long* counts = reinterpret_cast<const unsigned long long*>(host_words + kWordMaskCounts);
    if (counts[1] != 0) throw ConfigError("candidate set contains invalid states");
    candidate_count = counts[0];
}

void gsc_solver::setup() {
    setup_launch();
    setup_check();
}

uint64_t gsc_solver::device_bytes(const gsc_case_shape& sh, const gsc_case_desc& cfg) {
    auto r = [](uint64_t b) { return DeviceArena::round_up(b); };
    const uint32_t substeps = cfg.integration_substeps;
    const uint64_t S = sh.state_count, PT = sh.prefix_count;
    const uint64_t n0 = sh.state_shape[0], n1 = sh.state_shape[1], n2 = sh.state_shape[2], n3 = sh.state_shape[3];
    const uint64_t na = sh.input_shape[0], nw = sh.input_shape[1], I = sh.input_count;
    const uint64_t I_pad = (I + kInputBlock - 1) / kInputBlock * kInputBlock;
    const uint64_t planes = n2 * n3, K = substeps;
    uint64_t b = 3 * r(S) + 3 * r(4 * S) + 3 * r(4 * PT) + r(8 * n2) + r(8 * n3);
    b += r(16 * planes * ((n0 * n1 + 63) / 64));
    b += r(planes * (((n0 * n1 + 63) / 64 + 15) / 16 * 16));
    b += 3 * r(4 * ((n0 + kCoarse - 1) / kCoarse + 1) * ((n1 + kCoarse - 1) / kCoarse + 1) * (n2 + 1) * (n3 + 1));
    b += r(4 * planes * I_pad * n0 * n1) + 2 * r(4 * planes * I_pad);  // records, probe bases, row kinds
    b += r(8 * (2 * ((S + 63) / 64 + 1) + planes * ((n0 * n1 + 63) / 64) + 1));  // the three bitmasks
    // Bitmask frontier: previous unreached bitmask, frontier, keep and always words, plane rectangles.
    b += r(8 * ((S + 63) / 64 + 1)) + r(4 * planes * n1 * ((n0 + 31) / 32)) + r(8 * (planes * ((n0 * n1 + 63) / 64) + 1)) +
         2 * r(8 * planes * ((n0 * n1 + 63) / 64)) + r(16 * planes);
    b += r(sizeof(OverflowEntry) * (1u << 16)) + r(4) + r(16) + r(sizeof(IterCtl)) + r(sizeof(IterRecord) * kRing);
    b += r(16 * n2 * nw * K) + r(8 * n3 * na * K) + r(8 * n2 * nw) + r(8 * n3 * na) + r(16 * n3);
    b += r(4 * (n0 + 1) * (n1 + 1)) + r(16 * planes * I * K) + r(32 * planes * I) + 2 * r(64 * 64);
    b += r(32 * planes * I) + r(16 * planes * I);  // cell offsets, row boxes
    b += r(sizeof(FastRow) * planes * I_pad) + r(4 * planes);  // fast rows, plane flags
    if (cfg.state_allow) b += r(S);
    return b;
}

void gsc_solver::seed_async() {
    hipLaunchKernelGGL(seed_kernel, dim3(blocks_for(dc.S, 256)), dim3(256), 0, stream, dc.S, cand, reach, ctrl, rstep);
    GSC_HIP(hipGetLastError());
    hipLaunchKernelGGL(reset_ctl_kernel, dim3(1), dim3(256), 0, stream, ctl, static_cast<unsigned long long>(candidate_count));
    GSC_HIP(hipGetLastError());
}

void gsc_solver::build_coarse(const uint32_t* P0, const uint32_t* P1, uint32_t* C0, uint32_t* C1, const IterCtl* c) {
    hipLaunchKernelGGL(coarse_kernel, dim3(blocks_for(coarse_count, 256)), dim3(256), 0, stream, dc, P0, P1, C0, C1, c);
    GSC_HIP(hipGetLastError());
}

void gsc_solver::build_prefix(const uint8_t* a, const uint8_t* b, uint32_t* P0, uint32_t* P1, const IterCtl* c) {
    const uint64_t lines0 = static_cast<uint64_t>(dc.shape[1]) * dc.shape[2] * dc.shape[3];
    hipLaunchKernelGGL(prefix_rows_kernel, dim3(blocks_for(lines0, 4)), dim3(256), 0, stream, dc, a, b, P0, P1, c);
    GSC_HIP(hipGetLastError());
    for (int axis = 1; axis <= 3; ++axis) {
        const int o1 = axis == 1 ? 2 : 1;
        const int o2 = axis == 3 ? 2 : 3;
        const uint64_t lines = static_cast<uint64_t>(dc.shape[0]) * dc.shape[o1] * dc.shape[o2];
        hipLaunchKernelGGL(prefix_axis_kernel, dim3(blocks_for(lines, 256)), dim3(256), 0, stream, dc, axis, P0, P1,
                           c);
        GSC_HIP(hipGetLastError());
    }
}

void gsc_solver::build_abstraction() {
    // In recompute mode this pass writes the state hulls only (4 B per state);
    // the boxes themselves are recomputed inside every satisfaction launch.
    launch_abstraction();
    finish_abstraction();
}

// The model tables of the abstraction: built on the host (trigonometry with
// the host libm, θ / v boxes, radii, constraint table), uploaded, then the
// per-(plane, input) tables on the device (probe offsets, trajectory
// increments and their sums). The host vectors stay alive in `tables` until
// the next call, past the asynchronous copies.
void gsc_solver::upload_tables() {
    HostTables& t = tables;
    build_tables(t.cs, t.vt, t.box2, t.box3, t.rad, t.fail);
    t.hull2 = cell_hulls(t.box2, dc.shape[2], dc.n_w, dc.wrap_dim == 2);
    t.hull3 = cell_hulls(t.box3, dc.shape[3], dc.n_a, dc.wrap_dim == 3);
    {
        // The most bitmask slices a compact box's θ or v range needs in the
        // satisfaction's row scan: with the OR bitmask (both ranges two cells
        // or longer) one slice per cell beyond the first, else one per cell. Ranges a compact record cannot have (rejected, or
        // inverted off the wrap axis) are left out.
        const auto lengths = [](const std::vector<uint2>& tab, uint32_t n, bool wraps) {
            std::vector<uint32_t> out;
            for (const uint2& r : tab) {
                if (r.x == kBoxReject || (r.x > r.y && !wraps)) continue;
                const uint32_t len = r.y >= r.x ? r.y - r.x + 1u : r.y + n - r.x + 1u;
                if (std::find(out.begin(), out.end(), len) == out.end()) out.push_back(len);
            }
            return out;
        };
        const std::vector<uint32_t> lz = lengths(t.box2, dc.shape[2], dc.wrap_dim == 2);
        const std::vector<uint32_t> lw = lengths(t.box3, dc.shape[3], dc.wrap_dim == 3);
        slices_max = 0;
        for (const uint32_t a : lz)
            for (const uint32_t b : lw) {
                const bool orb = probe_or_used() && a >= 2u && b >= 2u;
                slices_max = std::max(slices_max, std::max(orb ? a - 1u : a, orb ? b - 1u : b));
            }
    }
    {
        // Filter constants of the x / y fast path (set_cell_filter): every cell
        // centre lies within [lb, lb + n eta + eta], every increment within
        // |dt| max|v| (times three roundings), every radius within the table's.
        double vmax = 0.0, rmax[2] = {0.0, 0.0};
        for (const double v : t.vt) vmax = std::max(vmax, std::fabs(v));
        for (size_t k = 0; k + 1 < t.rad.size(); k += 2) {
            rmax[0] = std::max(rmax[0], std::fabs(t.rad[k]));
            rmax[1] = std::max(rmax[1], std::fabs(t.rad[k + 1]));
        }
        const double abs_sum = static_cast<double>(dc.K) * std::fabs(dc.dt) * vmax * (1.0 + 1e-12);
        for (int d = 0; d < 2; ++d) {
            const double top = dc.lb[d] + static_cast<double>(dc.shape[d]) * dc.eta[d];
            const double x_abs = std::max(std::fabs(dc.lb[d]), std::fabs(top)) + std::fabs(dc.eta[d]);
            set_cell_filter(dc, d, dc.K, x_abs, abs_sum, rmax[d]);
        }
    }
    // The uploads. From pageable memory every copy is staged by the runtime
    // while the host waits, so the tables are packed into one pinned buffer
    // first (kept for later abstractions; the previous one's copies have
    // ended, since finish_abstraction synchronises the stream). Without
    // allocation, and for a solver that draws on a server's arena (whose
    // admission path must not allocate host memory either), they go straight
    // from the host vectors.
    struct Upload {
        void* dst;
        const void* src;
        size_t bytes;
    };
    const Upload ups[] = {{cs, t.cs.data(), t.cs.size() * sizeof(double2)},
                          {vt, t.vt.data(), t.vt.size() * sizeof(double)},
                          {box2, t.box2.data(), t.box2.size() * sizeof(uint2)},
                          {box3, t.box3.data(), t.box3.size() * sizeof(uint2)},
                          {hull2, t.hull2.data(), t.hull2.size() * sizeof(uint2)},
                          {hull3, t.hull3.data(), t.hull3.size() * sizeof(uint2)},
                          {rad, t.rad.data(), t.rad.size() * sizeof(double)},
                          {fail2d, t.fail.data(), fail2d ? t.fail.size() * sizeof(uint32_t) : 0}};
    const auto aligned = [](size_t b) { return (b + 255) & ~static_cast<size_t>(255); };
    size_t total = 0;
    for (const Upload& u : ups) total += aligned(u.bytes);
    if (!no_alloc && arena == nullptr && table_stage_bytes < total) {
        if (table_stage) GSC_HIP(hipHostFree(table_stage));
        table_stage = nullptr;
        table_stage_bytes = 0;
        GSC_HIP(hipHostMalloc(reinterpret_cast<void**>(&table_stage), total, hipHostMallocDefault));
        table_stage_bytes = total;
    }
    size_t off = 0;
    for (const Upload& u : ups) {
        if (u.bytes == 0 || u.dst == nullptr) continue;
        const void* src = u.src;
        if (table_stage) {
            std::memcpy(table_stage + off, u.src, u.bytes);
            src = table_stage + off;
            off += aligned(u.bytes);
        }
        GSC_HIP(hipMemcpyAsync(u.dst, src, u.bytes, hipMemcpyHostToDevice, stream));
    }
    const uint64_t n_planes = static_cast<uint64_t>(dc.shape[2]) * dc.shape[3];
    const uint64_t n_inc = n_planes * dc.I * dc.K;
    if (n_inc) {
        hipLaunchKernelGGL(increments_kernel, dim3(blocks_for(n_inc, 256)), dim3(256), 0, stream, dc, cs, vt, inc);
        GSC_HIP(hipGetLastError());
    }
    // Also with no substeps (K = 0): the sums are then zero.
    const uint64_t n_rows = n_planes * dc.I;
    hipLaunchKernelGGL(increment_sums_kernel, dim3(blocks_for(n_rows, 256)), dim3(256), 0, stream, dc, inc, rad,
                       inc_sums, cell_offs);
    GSC_HIP(hipGetLastError());
    hipLaunchKernelGGL(pair_rows_kernel, dim3(blocks_for(n_planes, 4)), dim3(256), 0, stream, dc, box2, box3,
                       cell_offs, probe_or_used() ? 1u : 0u, probe_base, row_kind, row_box, fast_offs,
                       plane_mixed, overflow_count + 1);
    GSC_HIP(hipGetLastError());
}

void gsc_solver::launch_abstraction(bool rerun) {
    // A rerun (overflow regrowth) keeps the first run's start, so the timings
    // cover both runs.
    if (!rerun) {
        abstraction_t0 = std::chrono::steady_clock::now();
        if (!events.empty()) GSC_HIP(hipEventRecord(events[0], stream));
    }
    // The escaped-record count and the tallest box's extent start at 0 (the
    // per-row pass of upload_tables and the abstraction kernels add to them).
    // A fill kernel, not hipMemsetAsync, so a caller may capture this sequence.
    hipLaunchKernelGGL(fill_u32_kernel, dim3(1), dim3(64), 0, stream, overflow_count, 2u, 0u);
    GSC_HIP(hipGetLastError());
    upload_tables();
    AbstractionArgs A{};
    A.c = dc;
    A.inc = inc;
    A.inc_sums = inc_sums;
    A.cell_offs = cell_offs;
    A.row_kind = row_kind;
    A.row_box = row_box;
    A.fast_offs = fast_offs;
    A.plane_mixed = plane_mixed;
    A.rad = rad;
    A.fail2d = fail2d;
    A.prefix_valid = prefix_valid;
    A.rec = opts.recompute_abstraction ? nullptr : rec;
    A.overflow = overflow;
    A.overflow_count = overflow_count;
    A.overflow_capacity = overflow_capacity;
    A.ext1_max = overflow_count + 1;
    A.hull = hull;
    A.chunk_hull = chunk_hull;
    GSC_HIP(launch_abstraction_kernels(A, stream));
    if (probe_or_used()) {
        // What the bitmask frontier needs of the state hulls (used by solve()
        // when no record escaped).
        hipLaunchKernelGGL(frontier_setup_kernel, dim3(dc.shape[2] * dc.shape[3]), dim3(256), 0, stream, dc, hull,
                           frontier_keep, frontier_always, plane_rect);
        GSC_HIP(hipGetLastError());
    }
    if (!events.empty()) GSC_HIP(hipEventRecord(events[1], stream));
    GSC_HIP(hipMemcpyAsync(host_words + kWordOverflow, overflow_count, 2 * sizeof(uint32_t), hipMemcpyDeviceToHost,
                           stream));
}

void gsc_solver::finish_abstraction() {
    GSC_HIP(hipStreamSynchronize(stream));
    uint32_t used = host_words[kWordOverflow];
    if (used > overflow_capacity) {
        // More escaped boxes than the table holds: grow it and run once more.
        if (used >= kRecIndexMask) throw ConfigError("too many irregular successor boxes");
        // The old table goes first, so an arena can reuse its bytes.
        release(overflow);
        overflow = nullptr;
        overflow_capacity = used;
        overflow = alloc<OverflowEntry>(overflow_capacity);
        launch_abstraction(true);
        GSC_HIP(hipStreamSynchronize(stream));
        used = host_words[kWordOverflow];
        if (used > overflow_capacity) throw ConfigError("successor-box overflow table did not converge");
    }
    overflow_used = used;
    ext1_used = host_words[kWordOverflow + 1];
    abstraction_wall_ms = ms_since(abstraction_t0);
    if (events.empty()) {
        abstraction_kernel_ms = abstraction_wall_ms;  // no events without allocation: wall time
    } else {
        float ms = 0.f;
        GSC_HIP(hipEventElapsedTime(&ms, events[0], events[1]));
        abstraction_kernel_ms = ms;
    }
    abstraction_ready = true;
}

FrontierArgs gsc_solver::frontier_args() const {
    FrontierArgs FA{};
    FA.c = dc;
    FA.valid = valid;
    FA.reach = reach;
    FA.valid_bits = valid_bits;
    FA.cur = ubits;
    FA.fresh = ubits_fresh;
    FA.fresh_v = fresh_v;
    FA.keep = frontier_keep;
    FA.always = frontier_always;
    FA.plane_rect = plane_rect;
    FA.hull2 = hull2;
    FA.hull3 = hull3;
    FA.frontier = frontier_words;
    FA.or_bits = probe_or_bits;
    FA.ctl = ctl;
    FA.row_words = (dc.shape[0] + 31u) / 32u;
    return FA;
}

void gsc_solver::solve() {
    if (!abstraction_ready) build_abstraction();
    if (!host_ctl) {
        GSC_HIP(hipHostMalloc(reinterpret_cast<void**>(&host_ctl), sizeof(IterCtl), hipHostMallocDefault));
        std::memset(host_ctl, 0, sizeof(IterCtl));
    }
    if (!host_records) {
        GSC_HIP(hipHostMalloc(reinterpret_cast<void**>(&host_records), sizeof(IterRecord) * kRing, hipHostMallocDefault));
        std::memset(host_records, 0, sizeof(IterRecord) * kRing);
    }
    if (no_alloc) {
        std::memset(host_ctl, 0, sizeof(IterCtl));
        std::memset(host_records, 0, sizeof(IterRecord) * kRing);
    }
    const uint32_t S = dc.S;
    hipLaunchKernelGGL(seed_kernel, dim3(blocks_for(S, 256)), dim3(256), 0, stream, S, cand, reach, ctrl, rstep);
    GSC_HIP(hipGetLastError());
    hipLaunchKernelGGL(reset_ctl_kernel, dim3(1), dim3(256), 0, stream, ctl, static_cast<unsigned long long>(candidate_count));
    GSC_HIP(hipGetLastError());
    GSC_HIP(hipStreamSynchronize(stream));

    iter_stats.clear();
    std::memset(&summary, 0, sizeof(summary));
    double prefix_ms = 0.0, sat_ms = 0.0, red_ms = 0.0;
    const bool timed = opts.time_kernels != 0;
    SatisfyArgs A{};
    A.c = dc;
    A.prefix_unreached[0] = prefix_unreached[0];
    A.prefix_unreached[1] = prefix_unreached[1];
    A.coarse_unreached[0] = coarse_unreached[0];
    A.coarse_unreached[1] = coarse_unreached[1];
    A.coarse_valid = coarse_valid;
    A.hull = hull;
    A.chunk_hull = chunk_hull;
    A.chunk_flags = chunk_flags;
    A.hull2 = hull2;
    A.hull3 = hull3;
    A.prefix_valid = prefix_valid;
    A.unreached_bits = ubits;
    A.probe_base = probe_base;
    A.probe_or_bits = probe_or_used() ? probe_or_bits : nullptr;
    A.valid_bits = valid_bits;
    const bool bits_mode = bits_frontier();
    A.frontier_bits = bits_mode ? frontier_words : nullptr;
    const FrontierArgs FA = frontier_args();
    const bool chunk_flags_used = !opts.sweep_all_pairs && dc.hull_enabled;
    A.rec = rec;
    A.overflow = overflow;
    A.box2 = box2;
    A.box3 = box3;
    A.valid = valid;
    A.cand = cand;
    A.reach = reach;
    A.ctrl = ctrl;
    A.rstep = rstep;
    A.ctl = ctl;
    A.sweep_all = opts.sweep_all_pairs ? 1u : 0u;
    RecomputeArgs RA{};
    RA.sat = A;
    RA.inc = inc;
    RA.inc_sums = inc_sums;
    RA.cell_offs = cell_offs;
    RA.rad = rad;
    RA.fail2d = fail2d;
    const uint32_t max_it = cfg.max_iterations;
    uint32_t done = 0;
    uint32_t finished = 0;

    const auto t0 = std::chrono::steady_clock::now();
    for (uint32_t first = 1; first <= max_it && !done; first += kBatch) {
        const uint32_t last = std::min<uint64_t>(static_cast<uint64_t>(first) + kBatch - 1, max_it);
        for (uint32_t it = first; it <= last; ++it) {
            hipEvent_t* ev = &events[(it - first) * kEventsPerIter];
            if (timed) GSC_HIP(hipEventRecord(ev[0], stream));
            if (bits_mode) {
                // The unreached bitmask (keeping the previous one), then the
                // frontier words and the OR probe bitmask.
                hipLaunchKernelGGL(frontier_bits_kernel, dim3(blocks_for((S + 7u) / 8u, 256)), dim3(256), 0, stream, FA,
                                   it);
                GSC_HIP(hipGetLastError());
                GSC_HIP(launch_frontier_plane(FA, stream));
            } else if (plane_pitch) {
                // Plane kernel (x, y and the bitmask), θ pass, v pass with the coarse sample.
                hipLaunchKernelGGL(prefix_plane_kernel, dim3(dc.shape[2] * dc.shape[3]), dim3(kPlaneThreads),
                                   plane_lds_bytes(plane_pitch, dc.shape[1], dc.PQ), stream, dc, valid, reach,
                                   prefix_unreached[0], prefix_unreached[1], ubits, plane_pitch, ctl);
                GSC_HIP(hipGetLastError());
                const uint64_t lines2 = static_cast<uint64_t>(dc.shape[0]) * dc.shape[1] * dc.shape[3];
                hipLaunchKernelGGL(prefix_axis_kernel, dim3(blocks_for(lines2, 256)), dim3(256), 0, stream, dc, 2,
                                   prefix_unreached[0], prefix_unreached[1], ctl);
                GSC_HIP(hipGetLastError());
                const uint64_t lines3 = static_cast<uint64_t>(dc.shape[0]) * dc.shape[1] * dc.shape[2];
                hipLaunchKernelGGL(prefix_last_axis_kernel, dim3(blocks_for(lines3, 256)), dim3(256), 0, stream, dc,
                                   prefix_unreached[0], prefix_unreached[1], coarse_unreached[0], coarse_unreached[1],
                                   ctl);
                GSC_HIP(hipGetLastError());
            } else {
                build_prefix(valid, reach, prefix_unreached[0], prefix_unreached[1], ctl);
                build_coarse(prefix_unreached[0], prefix_unreached[1], coarse_unreached[0], coarse_unreached[1], ctl);
                hipLaunchKernelGGL(unreached_bits_kernel, dim3(blocks_for(S, 256)), dim3(256), 0, stream, S, valid,
                                   reach, ubits, ctl);
                GSC_HIP(hipGetLastError());
            }
            // Which 64-cell chunks saw a change in their hull, for the
            // satisfaction's frontier skipping (both abstraction modes).
            if (chunk_flags_used && !bits_mode) GSC_HIP(launch_chunk_flags(A, chunk_flags, it, stream));
            if (timed) GSC_HIP(hipEventRecord(ev[1], stream));
            if (opts.recompute_abstraction) {
                hipLaunchKernelGGL(satisfy_recompute_kernel, dim3(satisfy_blocks(dc, 256)), dim3(256), 0, stream, RA);
                GSC_HIP(hipGetLastError());
            } else {
                GSC_HIP(launch_satisfy(A, it, tile_wg, tile_ib, stream));
            }
            if (timed) GSC_HIP(hipEventRecord(ev[2], stream));
            hipLaunchKernelGGL(finish_kernel, dim3(1), dim3(64), 0, stream, ctl, records, max_it);
            GSC_HIP(hipGetLastError());
            if (timed) GSC_HIP(hipEventRecord(ev[3], stream));
        }
        GSC_HIP(hipMemcpyAsync(host_ctl, ctl, sizeof(IterCtl), hipMemcpyDeviceToHost, stream));
        GSC_HIP(hipMemcpyAsync(host_records, records, sizeof(IterRecord) * kRing, hipMemcpyDeviceToHost, stream));
        GSC_HIP(hipStreamSynchronize(stream));
        done = host_ctl->done;
        const uint32_t upto = host_ctl->iterations;
        for (uint32_t it = first; it <= upto && it <= last; ++it) {
            const IterRecord& r = host_records[it % kRing];
            gsc_iteration_stats st;
            std::memset(&st, 0, sizeof(st));
            st.iteration = r.iteration;
            st.newly_reachable = r.newly_reachable;
            st.newly_certified = r.newly_certified;
            st.total_reachable = r.total_reachable;
            st.total_certified = r.total_certified;
            if (timed) {
                hipEvent_t* ev = &events[(it - first) * kEventsPerIter];
                float a = 0, b = 0, c = 0, d = 0;
                GSC_HIP(hipEventElapsedTime(&a, ev[0], ev[1]));
                GSC_HIP(hipEventElapsedTime(&b, ev[1], ev[2]));
                GSC_HIP(hipEventElapsedTime(&c, ev[2], ev[3]));
                GSC_HIP(hipEventElapsedTime(&d, ev[0], ev[3]));
                st.prefix_build_ms = a;
                st.satisfaction_check_ms = b;
                st.reduction_ms = c;
                st.iteration_ms = d;
                prefix_ms += a;
                sat_ms += b;
                red_ms += c;
            }
            summary.pairs_tested += r.pairs_tested;
            summary.boxes_counted += r.boxes_counted;
            summary.test_slots += r.test_slots;
            iter_stats.push_back(st);
            if (opts.verbose && (it <= 5 || it % 10 == 0 || it == max_it))
                std::printf("GPU: Iteration %u/%u, newly reachable: %llu, newly certified: %llu, total certified: "
                            "%llu/%llu (%g ms) [prefix: %g ms, satisfaction: %g ms, reduction: %g ms]\n",
                            it, max_it, r.newly_reachable, r.newly_certified, r.total_certified,
                            static_cast<unsigned long long>(candidate_count), st.iteration_ms, st.prefix_build_ms,
                            st.satisfaction_check_ms, st.reduction_ms);
        }
        finished = upto;
    }
    summary.solve_ms = ms_since(t0);
    summary.iterations = finished;
    summary.converged = done == kDoneConverged ? 1 : 0;
    std::snprintf(summary.message, sizeof(summary.message), "%s", done_message(done));
    summary.reachable_states = host_ctl->total_reachable;
    summary.certified_candidate_states = host_ctl->total_certified;
    summary.candidate_states = candidate_count;
    summary.abstraction_ms = abstraction_kernel_ms;
    summary.prefix_ms = prefix_ms;
    summary.satisfaction_ms = sat_ms;
    summary.reduction_ms = red_ms;
    summary.pair_bytes_read = summary.pairs_tested * sizeof(uint32_t);
    summary.frontier = frontier_mode();
    if (max_it == 0) {
        summary.iterations = 0;
        summary.converged = 0;
        std::snprintf(summary.message, sizeof(summary.message), "%s", "maximum iterations reached");
        summary.reachable_states = candidate_count;
    }
    timings = gsc_kernel_timings{};
    timings.abstraction_kernel_ms = abstraction_kernel_ms;
    timings.abstraction_memcpy_h2d_ms = setup_h2d_ms;
    timings.prefix_build_total_ms = prefix_ms;
    timings.pair_satisfaction_ms = sat_ms;
    timings.reduce_inputs_ms = red_ms;
    solved = true;
}

// Host copy of n bytes split over a few threads: the staging buffer is pinned
// and the destination is the caller's pageable memory, so one thread alone is
// bound by its own memcpy bandwidth, well below the link's.
static void parallel_copy(void* dst, const void* src, uint64_t n) {
    constexpr unsigned kThreads = 4;
    if (n < (8u << 20)) {
        std::memcpy(dst, src, n);
        return;
    }
    const uint64_t part = (n / kThreads + 4095) & ~uint64_t(4095);
    std::thread t[kThreads - 1];
    for (unsigned k = 1; k < kThreads; ++k) {
        const uint64_t b = std::min(n, part * k), e = std::min(n, part * (k + 1));
        t[k - 1] = std::thread([=] { std::memcpy(static_cast<char*>(dst) + b, static_cast<const char*>(src) + b, e - b); });
    }
    std::memcpy(dst, src, std::min(n, part));
    for (auto& th : t) th.join();
}

// The reference's abstraction triplet (kernels.cu:806-824) for every pair, in
// slices: export_kernel expands a slice of records on the device, an async
// copy moves it into one of two pinned host buffers, and while the host copies
// that buffer into the caller's arrays the device already works on the next
// slice into the other pair of buffers. Staging is allocated once per solver
// and kept for later calls. A solver without allocation (no_alloc) takes
// its device staging from its arena and copies each slice straight into the
// caller's arrays.
void gsc_solver::export_triplet(const gsc_result_buffers* out) {
    const uint64_t P = shape.pair_count;
    const bool want[3] = {out->abstraction_min_flat != nullptr, out->abstraction_max_flat != nullptr,
                          out->abstraction_valid != nullptr};
    if (!export_slice) {
        export_slice = std::min<uint64_t>(P, no_alloc ? kExportSliceNoAlloc : kExportSlice);
        for (int k = 0; k < (no_alloc ? 1 : 2); ++k) {
            ExportStage& st = xstage[k];
            st.dmin = alloc<uint32_t>(export_slice);
            st.dmax = alloc<uint32_t>(export_slice);
            st.dval = alloc<uint8_t>(export_slice);
            if (no_alloc) continue;
            GSC_HIP(hipHostMalloc(reinterpret_cast<void**>(&st.hmin), export_slice * 4, hipHostMallocDefault));
            GSC_HIP(hipHostMalloc(reinterpret_cast<void**>(&st.hmax), export_slice * 4, hipHostMallocDefault));
            GSC_HIP(hipHostMalloc(reinterpret_cast<void**>(&st.hval), export_slice, hipHostMallocDefault));
            GSC_HIP(hipEventCreateWithFlags(&st.done, hipEventDisableTiming));
        }
    }
    const uint64_t n = (P + export_slice - 1) / export_slice;
    auto expand = [&](uint64_t k, ExportStage& st) {
        const uint64_t b = k * export_slice, e = std::min(P, b + export_slice);
        hipLaunchKernelGGL(export_kernel, dim3(blocks_for(e - b, 256)), dim3(256), 0, stream, dc, rec, box2, box3,
                           overflow, b, e, st.dmin, st.dmax, st.dval);
        GSC_HIP(hipGetLastError());
    };
    if (no_alloc) {
        for (uint64_t k = 0; k < n; ++k) {
            const uint64_t b = k * export_slice, e = std::min(P, b + export_slice);
            expand(k, xstage[0]);
            if (want[0])
                GSC_HIP(hipMemcpyAsync(out->abstraction_min_flat + b, xstage[0].dmin, (e - b) * 4,
                                       hipMemcpyDeviceToHost, stream));
            if (want[1])
                GSC_HIP(hipMemcpyAsync(out->abstraction_max_flat + b, xstage[0].dmax, (e - b) * 4,
                                       hipMemcpyDeviceToHost, stream));
            if (want[2])
                GSC_HIP(hipMemcpyAsync(out->abstraction_valid + b, xstage[0].dval, e - b, hipMemcpyDeviceToHost, stream));
            GSC_HIP(hipStreamSynchronize(stream));
        }
        return;
    }
    auto enqueue = [&](uint64_t k) {
        ExportStage& st = xstage[k & 1];
        const uint64_t m = std::min(P, (k + 1) * export_slice) - k * export_slice;
        expand(k, st);
        if (want[0]) GSC_HIP(hipMemcpyAsync(st.hmin, st.dmin, m * 4, hipMemcpyDeviceToHost, stream));
        if (want[1]) GSC_HIP(hipMemcpyAsync(st.hmax, st.dmax, m * 4, hipMemcpyDeviceToHost, stream));
        if (want[2]) GSC_HIP(hipMemcpyAsync(st.hval, st.dval, m, hipMemcpyDeviceToHost, stream));
        GSC_HIP(hipEventRecord(st.done, stream));
    };
    enqueue(0);
    if (n > 1) enqueue(1);
    for (uint64_t k = 0; k < n; ++k) {
        ExportStage& st = xstage[k & 1];
        GSC_HIP(hipEventSynchronize(st.done));
        const uint64_t b = k * export_slice, m = std::min(P, b + export_slice) - b;
        if (want[0]) parallel_copy(out->abstraction_min_flat + b, st.hmin, m * 4);
        if (want[1]) parallel_copy(out->abstraction_max_flat + b, st.hmax, m * 4);
        if (want[2]) parallel_copy(out->abstraction_valid + b, st.hval, m);
        // This pair of buffers is free again: slice k + 2 may overwrite it.
        if (k + 2 < n) enqueue(k + 2);
    }
}

void gsc_solver::download(const gsc_result_buffers* out, gsc_run_report* report) {
    const auto t0 = std::chrono::steady_clock::now();
    const uint64_t S = shape.state_count;
    if (out) {
        if (out->valid_mask) GSC_HIP(hipMemcpyAsync(out->valid_mask, valid, S, hipMemcpyDeviceToHost, stream));
        if (out->candidate_mask) GSC_HIP(hipMemcpyAsync(out->candidate_mask, cand, S, hipMemcpyDeviceToHost, stream));
        if (solved) {
            if (out->reachable_mask)
                GSC_HIP(hipMemcpyAsync(out->reachable_mask, reach, S, hipMemcpyDeviceToHost, stream));
            if (out->controller)
                GSC_HIP(hipMemcpyAsync(out->controller, ctrl, S * 4, hipMemcpyDeviceToHost, stream));
            if (out->reach_step) GSC_HIP(hipMemcpyAsync(out->reach_step, rstep, S * 4, hipMemcpyDeviceToHost, stream));
        }
        GSC_HIP(hipStreamSynchronize(stream));
        if (abstraction_ready && (out->abstraction_min_flat || out->abstraction_max_flat || out->abstraction_valid)) {
            if (opts.recompute_abstraction)
                throw ConfigError("the pair abstraction is not stored in recompute mode");
            const auto tx = std::chrono::steady_clock::now();
            export_triplet(out);
            export_ms = ms_since(tx);
        }
        if (out->iteration_stats) {
            const uint32_t n = std::min<uint32_t>(out->iteration_stats_capacity, static_cast<uint32_t>(iter_stats.size()));
            std::memcpy(out->iteration_stats, iter_stats.data(), n * sizeof(gsc_iteration_stats));
        }
    }
    timings.result_memcpy_d2h_ms = ms_since(t0);
    timings.total_kernel_ms = timings.abstraction_kernel_ms + timings.pair_satisfaction_ms + timings.reduce_inputs_ms +
                              timings.prefix_build_total_ms;
    timings.total_memcpy_ms = timings.abstraction_memcpy_h2d_ms + timings.iteration_memcpy_ms + timings.result_memcpy_d2h_ms;
    timings.total_compute_ms = timings.total_kernel_ms + timings.total_memcpy_ms;
    if (report) {
        report->compiled_with_gpu = 1;
        report->runtime_available = 1;
        report->executed = solved ? 1 : 0;
        report->converged = summary.converged;
        put(report->result_message, sizeof(report->result_message), summary.message);
        report->abstraction_ms = abstraction_wall_ms;
        report->solve_ms = summary.solve_ms;
        report->total_ms = abstraction_wall_ms + summary.solve_ms;
        report->iterations = summary.iterations;
        report->iteration_stats_written =
            out && out->iteration_stats ? std::min<uint32_t>(out->iteration_stats_capacity, static_cast<uint32_t>(iter_stats.size())) : 0;
        report->reachable_states = summary.reachable_states;
        report->certified_candidate_states = summary.certified_candidate_states;
        report->candidate_states = candidate_count;
        report->pair_count = shape.pair_count;
        report->shape = shape;
        report->budget = make_budget(shape);
        report->kernel_timings = timings;
        report->memory_stats.total_allocated_bytes = bytes_allocated;
        report->memory_stats.abstraction_bytes = shape.pair_count * 4 + uint64_t(overflow_capacity) * sizeof(OverflowEntry);
        report->memory_stats.prefix_bytes = shape.prefix_count * 4 * 2;
        report->memory_stats.iteration_bytes = shape.state_count * (3 + 8) + sizeof(IterCtl) + sizeof(IterRecord) * kRing;
        size_t free_b = 0, total_b = 0;
        if (hipMemGetInfo(&free_b, &total_b) == hipSuccess) report->memory_stats.peak_usage_bytes = total_b - free_b;
    }
}

// ===========================================================================
// C ABI
// ===========================================================================

extern "C" {

int gsc_backend_compiled(void) { return 1; }

int gsc_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int gsc_case_shape_of(const gsc_case_desc* cfg, gsc_case_shape* out, char* err, size_t err_len) {
    if (!cfg || !out) return GSC_ERR_INVALID_ARGUMENT;
    try {
        *out = make_shape(*cfg);
        return GSC_OK;
    } catch (const std::exception& e) {
        put(err, err_len, e.what());
        return GSC_ERR_CONFIG;
    }
}

void gsc_solver_default_options(gsc_solver_options* o) {
    if (!o) return;
    std::memset(o, 0, sizeof(*o));
    o->device = 0;
    o->keep_export = 0;
    o->time_kernels = 1;
    o->sweep_all_pairs = 0;
    o->verbose = 0;
}

int gsc_solver_create(const gsc_case_desc* cfg, const gsc_solver_options* opts, gsc_solver** out, char* err,
                      size_t err_len) {
    if (!cfg || !out) return GSC_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    auto* s = new gsc_solver();
    if (opts)
        s->opts = *opts;
    else
        gsc_solver_default_options(&s->opts);
    // Default tile: 512 cells, where the tested states are sparse enough that a
    // group's four waves share a few batches; 256 under sweep_all_pairs, where
    // every state is tested and larger groups would run their batches in turn.
    const int32_t wg = s->opts.tile_wg ? s->opts.tile_wg : s->opts.sweep_all_pairs ? 256 : 512;
    const int32_t ib = s->opts.tile_ib ? s->opts.tile_ib : 8;
    if (wg < 0 || ib < 0 || !satisfy_variant_supported(static_cast<uint32_t>(wg), static_cast<uint32_t>(ib))) {
        put(err, err_len, "tile_wg must be 128, 256, 512 or 1024 and tile_ib 4 or 8");
        delete s;
        return GSC_ERR_INVALID_ARGUMENT;
    }
    s->tile_wg = static_cast<uint32_t>(wg);
    s->tile_ib = static_cast<uint32_t>(ib);
    copy_case(s, *cfg);
    try {
        s->setup();
    } catch (const ConfigError& e) {
        put(err, err_len, e.what());
        delete s;
        return GSC_ERR_CONFIG;
    } catch (const DeviceError& e) {
        put(err, err_len, e.what());
        const int st = status_of(e);
        delete s;
        return st;
    } catch (const std::exception& e) {
        put(err, err_len, e.what());
        delete s;
        return GSC_ERR_DEVICE;
    }
    *out = s;
    return GSC_OK;
}

int gsc_solver_shape(const gsc_solver* s, gsc_case_shape* out) {
    if (!s || !out) return GSC_ERR_INVALID_ARGUMENT;
    *out = s->shape;
    return GSC_OK;
}

#define GSC_GUARD(s, body)                                   \
    try {                                                    \
        body;                                                \
        return GSC_OK;                                       \
    } catch (const ConfigError& e) {                         \
        (s)->last_error = e.what();                          \
        return GSC_ERR_CONFIG;                               \
    } catch (const DeviceError& e) {                         \
        (s)->last_error = e.what();                          \
        return status_of(e);                                 \
    } catch (const std::exception& e) {                      \
        (s)->last_error = e.what();                          \
        return GSC_ERR_DEVICE;                               \
    }

int gsc_solver_build_abstraction(gsc_solver* s, double* kernel_ms) {
    if (!s) return GSC_ERR_INVALID_ARGUMENT;
    GSC_GUARD(s, {
        s->build_abstraction();
        if (kernel_ms) *kernel_ms = s->abstraction_kernel_ms;
    })
}

int gsc_solver_solve(gsc_solver* s, gsc_solve_summary* out) {
    if (!s) return GSC_ERR_INVALID_ARGUMENT;
    GSC_GUARD(s, {
        s->solve();
        if (out) *out = s->summary;
    })
}

int gsc_solver_download(gsc_solver* s, const gsc_result_buffers* out, gsc_run_report* report) {
    if (!s) return GSC_ERR_INVALID_ARGUMENT;
    GSC_GUARD(s, s->download(out, report))
}

const char* gsc_solver_last_error(const gsc_solver* s) { return s ? s->last_error.c_str() : ""; }

void gsc_solver_destroy(gsc_solver* s) { delete s; }

int gsc_run_case_u32(const gsc_case_desc* cfg, const gsc_result_buffers* out, gsc_run_report* report) {
    if (!cfg || !report) return GSC_ERR_INVALID_ARGUMENT;
    std::memset(report, 0, sizeof(*report));
    report->compiled_with_gpu = 1;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
        put(report->message, sizeof(report->message), "no HIP device found");
        return GSC_ERR_NO_DEVICE;
    }
    report->runtime_available = 1;
    gsc_solver_options opts;
    gsc_solver_default_options(&opts);
    opts.verbose = cfg->verbose;
    gsc_solver* s = nullptr;
    const int st = gsc_solver_create(cfg, &opts, &s, report->message, sizeof(report->message));
    if (st != GSC_OK) return st;
    int rc = gsc_solver_build_abstraction(s, nullptr);
    if (rc == GSC_OK) rc = gsc_solver_solve(s, nullptr);
    if (rc == GSC_OK) rc = gsc_solver_download(s, out, report);
    if (rc != GSC_OK) {
        put(report->message, sizeof(report->message), s->last_error);
        report->executed = 0;
    }
    gsc_solver_destroy(s);
    return rc;
}

}  // extern "C"

// ===========================================================================
// gsc_run_case_u32_static: the same run with every device buffer carved out of
// one file-scope device array, so the call allocates nothing.
// ===========================================================================
namespace {

// Each device that loads this library holds its own copy of the array until
// the process ends. In host code the name is only a stand-in for the symbol;
// the device address comes from hipGetSymbolAddress, once per device.
static __device__ __align__(256) unsigned char scratch[GSC_STATIC_SCRATCH_BYTES];

constexpr int kMaxStaticDevices = 64;

// Per device: the symbol's address on that device, a lock that serialises the
// calls sharing its array, and the host words the fixpoint loop reads back
// (pageable; nothing pinned is allocated).
struct StaticDevice {
    std::mutex m;
    void* base = nullptr;
    IterCtl ctl{};
    IterRecord records[kRing]{};
    uint32_t words[kHostWords]{};
};
StaticDevice g_static[kMaxStaticDevices];

}  // namespace

extern "C" int gsc_run_case_u32_static(const gsc_case_desc* cfg, const gsc_result_buffers* out,
                                       gsc_run_report* report) {
    if (!cfg || !report) return GSC_ERR_INVALID_ARGUMENT;
    std::memset(report, 0, sizeof(*report));
    report->compiled_with_gpu = 1;
    int n = 0, dev = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0 || hipGetDevice(&dev) != hipSuccess) {
        put(report->message, sizeof(report->message), "no HIP device found");
        return GSC_ERR_NO_DEVICE;
    }
    if (dev < 0 || dev >= kMaxStaticDevices) {
        put(report->message, sizeof(report->message), "device ordinal beyond the static scratch table");
        return GSC_ERR_INVALID_ARGUMENT;
    }
    report->runtime_available = 1;
    StaticDevice& sd = g_static[dev];
    std::lock_guard<std::mutex> lock(sd.m);
    if (!sd.base) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            put(report->message, sizeof(report->message),
                std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
            return GSC_ERR_DEVICE;
        }
        sd.base = p;
    }

    // The size check: every buffer of the case plus the export staging.
    gsc_case_shape sh{};
    try {
        sh = make_shape(*cfg);
    } catch (const ConfigError& e) {
        put(report->message, sizeof(report->message), e.what());
        return GSC_ERR_CONFIG;
    }
    const bool exporting = out && (out->abstraction_min_flat || out->abstraction_max_flat || out->abstraction_valid);
    const uint64_t stage = exporting ? 3 * DeviceArena::round_up(std::min<uint64_t>(sh.pair_count, kExportSliceNoAlloc) * 4)
                                     : 0;
    const uint64_t need = gsc_solver::device_bytes(sh, *cfg) + stage;
    if (need > GSC_STATIC_SCRATCH_BYTES) {
        char msg[GSC_MESSAGE_LEN];
        std::snprintf(msg, sizeof(msg),
                      "case needs %llu bytes of device memory; gsc_run_case_u32_static holds at most %llu "
                      "(GSC_STATIC_SCRATCH_BYTES)",
                      static_cast<unsigned long long>(need), static_cast<unsigned long long>(GSC_STATIC_SCRATCH_BYTES));
        put(report->message, sizeof(report->message), msg);
        return GSC_ERR_OUT_OF_MEMORY;
    }

    DeviceArena arena(sd.base, GSC_STATIC_SCRATCH_BYTES);
    gsc_solver s;
    gsc_solver_default_options(&s.opts);
    s.opts.device = dev;
    s.opts.time_kernels = 0;
    s.opts.verbose = cfg->verbose;
    copy_case(&s, *cfg);
    s.no_alloc = true;
    s.arena = &arena;
    s.host_words = sd.words;
    s.owns_host_words = false;
    s.host_ctl = &sd.ctl;
    s.host_records = sd.records;
    s.owns_host_ctl = false;
    int rc = GSC_OK;
    try {
        s.setup();
        s.build_abstraction();
        s.solve();
        s.download(out, report);
    } catch (const ConfigError& e) {
        s.last_error = e.what();
        rc = GSC_ERR_CONFIG;
    } catch (const DeviceError& e) {
        s.last_error = e.what();
        rc = status_of(e);
    } catch (const std::exception& e) {
        s.last_error = e.what();
        rc = GSC_ERR_DEVICE;
    }
    if (rc != GSC_OK) {
        put(report->message, sizeof(report->message), s.last_error);
        report->executed = 0;
    }
    return rc;
}
