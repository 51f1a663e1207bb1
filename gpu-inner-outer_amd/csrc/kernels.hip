// kernels.hip — gfx950 kernels of the grid symbolic-control hot path.
//
//   masks_kernel         valid / candidate masks + fixpoint seeds
//                        (abstraction.hpp:65-106, kernels.cu:576-598)
//   prefix_rows_kernel   4D prefix sums, x axis: one wave per (y,z,w) line,
//   prefix_axis_kernel   then one coalesced pass per remaining axis
//                        (prefix_sum.hpp:15-63, kernels.cu:219-308)
//   abstraction_kernel   successor cell box of every (state, input) pair
//                        (kernels.cu:164-217, abstraction.hpp:161-209)
//   satisfy_kernel       pair satisfaction fused with the first-input reduction
//                        (kernels.cu:310-393, reachability.hpp:59-109)
//   finish_kernel        per-iteration totals and the stop test
//                        (kernels.cu:722-766, reachability.hpp:420-452)
//   export_kernel        reference abstraction triplet from the pair records
//
// No MFMA anywhere: nothing here is a dense contraction. The satisfaction pass
// streams one 4-byte record per pair (coalesced) and gathers prefix corners;
// the abstraction pass is scalar FP64 with the trigonometry read from
// host-built tables.
#include "kernels.hpp"

namespace gsc_amd {

namespace {

__device__ inline uint32_t lane_id() { return threadIdx.x & 63u; }

__device__ inline uint64_t lanes_below_mask() {
    const uint32_t l = lane_id();
    return l == 0 ? 0ull : (~0ull >> (64u - l));
}

// Inclusive prefix sum across a 64-lane wave.
__device__ inline uint32_t wave_inclusive_scan(uint32_t v) {
    const uint32_t l = lane_id();
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t n = __shfl_up(v, off, 64);
        if (l >= static_cast<uint32_t>(off)) v += n;
    }
    return v;
}

__device__ inline uint64_t wave_sum_u64(uint64_t v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__device__ inline bool loop_done(const IterCtl* ctl) {
    return __atomic_load_n(&ctl->done, __ATOMIC_RELAXED) != 0;
}

__device__ inline void unflatten(const DevCase& c, uint32_t s, uint32_t idx[4]) {
    // grid.hpp:60-66 with x fastest.
    uint32_t t = s / c.shape[0];
    idx[0] = s - t * c.shape[0];
    uint32_t t2 = t / c.shape[1];
    idx[1] = t - t2 * c.shape[1];
    idx[3] = t2 / c.shape[2];
    idx[2] = t2 - idx[3] * c.shape[2];
}

}  // namespace

// ---------------------------------------------------------------------------
// Masks and seeds. Thread per state.
__global__ __launch_bounds__(256) void masks_kernel(DevCase c, const gsc_rect4* __restrict__ obstacles,
                                                     const gsc_rect4* __restrict__ targets,
                                                     uint8_t* __restrict__ valid, uint8_t* __restrict__ cand,
                                                     unsigned long long* __restrict__ counts) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    bool is_cand = false, bad = false;
    if (s < c.S) {
        uint32_t idx[4];
        unflatten(c, s, idx);
        double x[4];
        for (int d = 0; d < 4; ++d) x[d] = cell_center(c.lb[d], idx[d], c.eta[d]);
        bool ok = point_in_rect(c.map, x);
        if (ok)
            for (uint32_t k = 0; k < c.n_obstacles; ++k)
                if (point_in_rect(obstacles[k], x)) {
                    ok = false;
                    break;
                }
        if (ok && c.constraint_type != GSC_CONSTRAINT_NONE)
            ok = constraint_ok(c.constraint_type, c.hyp_a, c.hyp_b, c.hyp_c, c.ell_a, c.ell_b, x[0], x[1]);
        is_cand = point_in_rect(c.candidate, x);
        if (is_cand && c.band_enabled) is_cand = fabs(x[0] * x[1] - c.band_c) <= c.band_w;
        if (!is_cand)
            for (uint32_t k = 0; k < c.n_targets; ++k)
                if (point_in_rect(targets[k], x)) {
                    is_cand = true;
                    break;
                }
        valid[s] = ok ? 1 : 0;
        cand[s] = is_cand ? 1 : 0;
        bad = is_cand && !ok;
    }
    const uint64_t nc = __popcll(__ballot(is_cand));
    const uint64_t nb = __popcll(__ballot(bad));
    if (lane_id() == 0) {
        if (nc) atomicAdd(&counts[0], static_cast<unsigned long long>(nc));
        if (nb) atomicAdd(&counts[1], static_cast<unsigned long long>(nb));
    }
}

// Seeds of the fixpoint (kernels.cu:576-598): reachable = candidate,
// controller = invalid, reach_step = 0 on candidates else unreachable.
__global__ __launch_bounds__(256) void seed_kernel(uint32_t S, const uint8_t* __restrict__ cand,
                                                    uint8_t* __restrict__ reach, uint32_t* __restrict__ ctrl,
                                                    uint32_t* __restrict__ rstep) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= S) return;
    const uint8_t c = cand[s];
    reach[s] = c;
    ctrl[s] = GSC_INVALID_INPUT;
    rstep[s] = c ? 0u : GSC_UNREACHABLE_STEP;
}

__global__ void reset_ctl_kernel(IterCtl* ctl, unsigned long long candidate_count) {
    const uint32_t t = threadIdx.x;
    if (t == 0) {
        ctl->done = 0;
        ctl->iterations = 0;
        ctl->candidate_count = candidate_count;
        ctl->total_reachable = candidate_count;
        ctl->total_certified = 0;
    }
    for (uint32_t k = t; k < kRing; k += blockDim.x) {
        ctl->ring[k].new_reachable = 0;
        ctl->ring[k].new_certified = 0;
        ctl->ring[k].pairs_tested = 0;
        ctl->ring[k].pad = 0;
    }
}

// ---------------------------------------------------------------------------
// Prefix table of (a & !b) over the padded layout. Padding entries (index 0 on
// any axis) are zero from allocation and never written.
// Pass 1: one wave per interior (y,z,w) line; wave scan along x.
__global__ __launch_bounds__(256) void prefix_rows_kernel(DevCase c, const uint8_t* __restrict__ a,
                                                           const uint8_t* __restrict__ b, uint32_t* __restrict__ P,
                                                           const IterCtl* ctl) {
    if (ctl && loop_done(ctl)) return;
    const uint32_t line = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const uint32_t lines = c.shape[1] * c.shape[2] * c.shape[3];
    if (line >= lines) return;
    const uint32_t y = line % c.shape[1];
    const uint32_t t = line / c.shape[1];
    const uint32_t z = t % c.shape[2];
    const uint32_t w = t / c.shape[2];
    const uint32_t src = y * c.stride[1] + z * c.stride[2] + w * c.stride[3];
    const uint32_t dst = (y + 1) * c.pstride[1] + (z + 1) * c.pstride[2] + (w + 1) * c.pstride[3] + 1;
    uint32_t carry = 0;
    for (uint32_t x0 = 0; x0 < c.shape[0]; x0 += 64) {
        const uint32_t x = x0 + lane_id();
        uint32_t v = 0;
        if (x < c.shape[0]) v = (a[src + x] != 0 && !(b && b[src + x] != 0)) ? 1u : 0u;
        v = wave_inclusive_scan(v) + carry;
        if (x < c.shape[0]) P[dst + x] = v;
        carry = __shfl(v, 63, 64);
    }
}

// Pass 2..4: running sum along `axis` (1, 2 or 3); one thread per line, lines
// enumerated with x fastest so neighbouring threads touch neighbouring words.
__global__ __launch_bounds__(256) void prefix_axis_kernel(DevCase c, int axis, uint32_t* __restrict__ P,
                                                           const IterCtl* ctl) {
    if (ctl && loop_done(ctl)) return;
    const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
    // The two other non-x axes.
    const int o1 = axis == 1 ? 2 : 1;
    const int o2 = axis == 3 ? 2 : 3;
    const uint32_t n0 = c.shape[0];
    const uint32_t lines = n0 * c.shape[o1] * c.shape[o2];
    if (tid >= lines) return;
    const uint32_t x = tid % n0;
    const uint32_t r = tid / n0;
    const uint32_t i1 = r % c.shape[o1];
    const uint32_t i2 = r / c.shape[o1];
    uint32_t off = (x + 1) + (i1 + 1) * c.pstride[o1] + (i2 + 1) * c.pstride[o2] + c.pstride[axis];
    const uint32_t step = c.pstride[axis];
    const uint32_t n = c.shape[axis];
    uint32_t acc = 0;
    uint32_t k = 0;
    for (; k + 4 <= n; k += 4) {
        const uint32_t v0 = P[off], v1 = P[off + step], v2 = P[off + 2 * step], v3 = P[off + 3 * step];
        acc += v0;
        P[off] = acc;
        acc += v1;
        P[off + step] = acc;
        acc += v2;
        P[off + 2 * step] = acc;
        acc += v3;
        P[off + 3 * step] = acc;
        off += 4 * step;
    }
    for (; k < n; ++k) {
        acc += P[off];
        P[off] = acc;
        off += step;
    }
}

// ---------------------------------------------------------------------------
// Abstraction. A workgroup owns a tile of 64 consecutive states; lane l of
// every wave is state s0+l, and the four waves take interleaved inputs, so the
// trajectory tables are read at wave-uniform addresses. Records are staged in
// LDS as [state][input] and written out as one contiguous block.
//
// Per pair, the arithmetic is the reference's in the reference's order:
//   centre           grid.hpp:72
//   nominal x, y     model.hpp:35-42: x += dt*(v_k*cos θ_k), y += dt*(v_k*sin θ_k)
//                    with cos θ_k, sin θ_k, v_k from host tables
//   radius           model.hpp:71-73, 84-89, 107-111 (per v cell, host table)
//   cells            grid.hpp:211-281 (dims 0,1 here; dims 2,3 from box2/box3)
//   constraint scan  abstraction.hpp:110-159 as a 2D prefix query over x,y fails
__global__ __launch_bounds__(256) void abstraction_kernel(AbstractionArgs A) {
    extern __shared__ uint32_t tile[];
    const DevCase& c = A.c;
    const uint32_t lane = lane_id();
    const uint32_t wave = threadIdx.x >> 6;
    const uint32_t s0 = blockIdx.x * 64u;
    const uint32_t s = s0 + lane;
    const bool live = s < c.S;
    const uint32_t n_live = min(64u, c.S - s0);

    uint32_t idx[4] = {0, 0, 0, 0};
    if (live) unflatten(c, s, idx);
    const double x0 = cell_center(c.lb[0], idx[0], c.eta[0]);
    const double y0 = cell_center(c.lb[1], idx[1], c.eta[1]);
    const uint32_t fpitch = c.shape[0] + 1;
    bool state_ok = live;
    if (live && c.constraint_type != GSC_CONSTRAINT_NONE) {
        // abstraction.hpp:177-181: the state's own centre must satisfy the constraint.
        const uint32_t* F = A.fail2d;
        const uint32_t f = F[(idx[1] + 1) * fpitch + idx[0] + 1] - F[(idx[1] + 1) * fpitch + idx[0]] -
                           F[idx[1] * fpitch + idx[0] + 1] + F[idx[1] * fpitch + idx[0]];
        state_ok = f == 0;
    }
    const double rad0 = A.rad[2 * idx[3]];
    const double rad1 = A.rad[2 * idx[3] + 1];
    const int wd = c.wrap_dim;

    for (uint32_t i0 = 0; i0 < c.I; i0 += A.in_chunk) {
        const uint32_t ic = min(A.in_chunk, c.I - i0);
        for (uint32_t ii = wave; ii < ic; ii += 4) {
            const uint32_t i = i0 + ii;
            uint32_t r = kRecInvalid;
            uint32_t lo[4], hi[4];
            bool ok = state_ok;
            const uint32_t ia = i % c.n_a;
            const uint32_t iw = i / c.n_a;
            if (ok) {
                const uint2 b2 = A.box2[idx[2] * c.n_w + iw];
                const uint2 b3 = A.box3[idx[3] * c.n_a + ia];
                ok = b2.x != kBoxReject && b3.x != kBoxReject;
                lo[2] = b2.x;
                hi[2] = b2.y;
                lo[3] = b3.x;
                hi[3] = b3.y;
            }
            if (ok) {
                const double2* __restrict__ cs = A.cs + static_cast<size_t>(idx[2] * c.n_w + iw) * c.K;
                const double* __restrict__ vt = A.vt + static_cast<size_t>(idx[3] * c.n_a + ia) * c.K;
                double px = x0, py = y0;
                for (uint32_t k = 0; k < c.K; ++k) {
                    const double2 t = cs[k];
                    const double vk = vt[k];
                    const double dx = vk * t.x;
                    const double dy = vk * t.y;
                    px = px + c.dt * dx;
                    py = py + c.dt * dy;
                }
                ok = dim_to_cells(c, 0, px - rad0, px + rad0, &lo[0], &hi[0]) &&
                     dim_to_cells(c, 1, py - rad1, py + rad1, &lo[1], &hi[1]);
            }
            if (ok && c.constraint_type != GSC_CONSTRAINT_NONE) {
                // Every loop of abstraction.hpp:124-156 must be non-empty for a cell to be
                // tested; the test reads x and y only, so it reduces to "no failing (x, y)
                // cell in the x/y rectangle". The θ walk is empty only when its range is
                // inverted without a wrap (abstraction.hpp:125-135).
                const bool has_wrap = lo[wd] > hi[wd];
                const bool empty = lo[3] > hi[3] || (lo[2] > hi[2] && !has_wrap) || lo[1] > hi[1] || lo[0] > hi[0];
                if (!empty) {
                    const uint32_t* F = A.fail2d;
                    const uint32_t f = F[(hi[1] + 1) * fpitch + hi[0] + 1] - F[(hi[1] + 1) * fpitch + lo[0]] -
                                       F[lo[1] * fpitch + hi[0] + 1] + F[lo[1] * fpitch + lo[0]];
                    ok = f == 0;
                }
            }
            if (ok) {
                // Inverted ranges on a non-wrap axis or wrapped x/y ranges take the escape path.
                const bool regular = lo[0] <= hi[0] && lo[1] <= hi[1] && (lo[2] <= hi[2] || wd == 2) &&
                                     (lo[3] <= hi[3] || wd == 3);
                const int off0 = static_cast<int>(lo[0]) - static_cast<int>(idx[0]);
                const int off1 = static_cast<int>(lo[1]) - static_cast<int>(idx[1]);
                const bool fits = regular && off0 >= kOffMin && off0 <= kOffMax && off1 >= kOffMin &&
                                  off1 <= kOffMax && hi[0] - lo[0] <= kExtMax && hi[1] - lo[1] <= kExtMax;
                if (fits) {
                    const uint32_t vc = box_count_wrap_u32(A.prefix_valid, c.pstride, wd, c.shape[wd], lo, hi);
                    r = rec_pack(off0, hi[0] - lo[0], off1, hi[1] - lo[1], vc == 0);
                } else {
                    const int64_t vc = box_count_i64(A.prefix_valid, c.pstride, wd, c.shape[wd], lo, hi);
                    const uint32_t e = atomicAdd(A.overflow_count, 1u);
                    if (e < A.overflow_capacity) {
                        OverflowEntry o;
                        o.min_flat = lo[0] * c.stride[0] + lo[1] * c.stride[1] + lo[2] * c.stride[2] + lo[3] * c.stride[3];
                        o.max_flat = hi[0] * c.stride[0] + hi[1] * c.stride[1] + hi[2] * c.stride[2] + hi[3] * c.stride[3];
                        o.dead = vc == 0 ? 1u : 0u;
                        o.reserved = 0;
                        A.overflow[e] = o;
                        r = kRecEscape | e;
                    } else {
                        r = kRecInvalid;  // capacity exceeded: the host sees the count and reruns
                    }
                }
            }
            if (live) tile[lane * ic + ii] = r;
        }
        __syncthreads();
        const uint64_t base = static_cast<uint64_t>(s0) * c.I + i0;
        if (ic == c.I) {
            const uint32_t n = n_live * ic;
            for (uint32_t t = threadIdx.x; t < n; t += blockDim.x) A.rec[base + t] = tile[t];
        } else {
            const uint32_t n = n_live * ic;
            for (uint32_t t = threadIdx.x; t < n; t += blockDim.x) {
                const uint32_t l = t / ic;
                A.rec[base + static_cast<uint64_t>(l) * c.I + (t - l * ic)] = tile[t];
            }
        }
        __syncthreads();
    }
}

// Reference abstraction triplet (min_flat, max_flat, valid) from the records;
// zeros for invalid pairs (kernels.cu:189-216). Thread per pair.
__global__ __launch_bounds__(256) void export_kernel(DevCase c, const uint32_t* __restrict__ rec,
                                                      const uint2* __restrict__ box2, const uint2* __restrict__ box3,
                                                      const OverflowEntry* __restrict__ overflow, uint64_t pair_begin,
                                                      uint64_t pair_end, uint32_t* __restrict__ out_min,
                                                      uint32_t* __restrict__ out_max, uint8_t* __restrict__ out_valid) {
    const uint64_t p = pair_begin + static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (p >= pair_end) return;
    const uint64_t q = p - pair_begin;
    const uint32_t r = rec[p];
    if (r == kRecInvalid) {
        out_min[q] = 0;
        out_max[q] = 0;
        out_valid[q] = 0;
        return;
    }
    if (r & kRecEscape) {
        const OverflowEntry o = overflow[r & kRecIndexMask];
        out_min[q] = o.min_flat;
        out_max[q] = o.max_flat;
        out_valid[q] = 1;
        return;
    }
    const uint32_t s = static_cast<uint32_t>(p / c.I);
    const uint32_t i = static_cast<uint32_t>(p - static_cast<uint64_t>(s) * c.I);
    uint32_t idx[4];
    unflatten(c, s, idx);
    const uint2 b2 = box2[idx[2] * c.n_w + i / c.n_a];
    const uint2 b3 = box3[idx[3] * c.n_a + i % c.n_a];
    const uint32_t lo0 = idx[0] + rec_off0(r), lo1 = idx[1] + rec_off1(r);
    out_min[q] = lo0 * c.stride[0] + lo1 * c.stride[1] + b2.x * c.stride[2] + b3.x * c.stride[3];
    out_max[q] = (lo0 + rec_ext0(r)) * c.stride[0] + (lo1 + rec_ext1(r)) * c.stride[1] + b2.y * c.stride[2] +
                 b3.y * c.stride[3];
    out_valid[q] = 1;
}

// ---------------------------------------------------------------------------
// Pair satisfaction + first-input reduction, one workgroup per chunk of
// consecutive states.
//   1. load state flags, compact the states to test into LDS (wave ballots);
//   2. test their pairs in flat (state, input) order — coalesced 4-byte record
//      loads — with `sat = !dead && count(unreached valid cells in box) == 0`,
//      which equals the reference's `valid && cntV > 0 && cntR == cntV`
//      because reachable ⊆ valid; each wave stores its 64-pair ballot as one
//      word of an LDS bitmap;
//   3. per unsettled state, the lowest set bit of its input range is the
//      chosen input (reachability.hpp:106 / kernels.cu:369-374); update the
//      state in place (safe: the prefix snapshot was built before this launch,
//      and only this workgroup touches these states).
__global__ __launch_bounds__(256) void satisfy_kernel(SatisfyArgs A) {
    if (loop_done(A.ctl)) return;
    const DevCase& c = A.c;
    extern __shared__ uint64_t bits[];
    __shared__ uint32_t s_state[256];
    __shared__ uint32_t s_idx[4][256];
    __shared__ uint8_t s_use[256];
    __shared__ uint32_t s_wave_count[4];
    __shared__ unsigned long long s_counts[3];

    const uint32_t lane = lane_id();
    const uint32_t wave = threadIdx.x >> 6;
    const uint32_t s0 = blockIdx.x * A.chunk_states;
    const uint32_t t = threadIdx.x;
    const uint32_t s = s0 + t;
    if (t < 3) s_counts[t] = 0;

    bool use = false, eval = false;
    uint32_t idx[4] = {0, 0, 0, 0};
    if (t < A.chunk_states && s < c.S) {
        if (A.valid[s]) {
            const bool is_cand = A.cand[s] != 0;
            const bool reached = A.reach[s] != 0;
            const bool certified = A.ctrl[s] != GSC_INVALID_INPUT;
            use = !((!is_cand && reached) || (is_cand && certified));
        }
        eval = A.sweep_all ? true : use;
        if (eval) unflatten(c, s, idx);
    }
    const uint64_t m = __ballot(eval);
    if (lane == 0) s_wave_count[wave] = static_cast<uint32_t>(__popcll(m));
    __syncthreads();
    uint32_t base_pos = 0;
    for (uint32_t w = 0; w < wave; ++w) base_pos += s_wave_count[w];
    const uint32_t n_eval = s_wave_count[0] + s_wave_count[1] + s_wave_count[2] + s_wave_count[3];
    if (eval) {
        const uint32_t pos = base_pos + static_cast<uint32_t>(__popcll(m & lanes_below_mask()));
        s_state[pos] = s;
        s_use[pos] = use ? 1 : 0;
        s_idx[0][pos] = idx[0];
        s_idx[1][pos] = idx[1];
        s_idx[2][pos] = idx[2];
        s_idx[3][pos] = idx[3];
    }
    __syncthreads();
    if (n_eval == 0) return;

    const uint32_t I = c.I;
    const uint32_t n_pairs = n_eval * I;
    const int wd = c.wrap_dim;
    for (uint32_t pbase = wave * 64u; pbase < n_pairs; pbase += blockDim.x) {
        const uint32_t q = pbase + lane;
        bool sat = false;
        if (q < n_pairs) {
            const uint32_t a = q / I;
            const uint32_t i = q - a * I;
            const uint32_t st = s_state[a];
            const uint32_t r = A.rec[static_cast<uint64_t>(st) * I + i];
            if (r != kRecInvalid) {
                if (r & kRecEscape) {
                    const OverflowEntry o = A.overflow[r & kRecIndexMask];
                    if (!o.dead) {
                        uint32_t lo[4], hi[4];
                        unflatten(c, o.min_flat, lo);
                        unflatten(c, o.max_flat, hi);
                        sat = box_count_i64(A.prefix_unreached, c.pstride, wd, c.shape[wd], lo, hi) == 0;
                    }
                } else if (!(r & kRecDead)) {
                    const uint2 b2 = A.box2[s_idx[2][a] * c.n_w + i / c.n_a];
                    const uint2 b3 = A.box3[s_idx[3][a] * c.n_a + i % c.n_a];
                    uint32_t lo[4], hi[4];
                    lo[0] = s_idx[0][a] + rec_off0(r);
                    hi[0] = lo[0] + rec_ext0(r);
                    lo[1] = s_idx[1][a] + rec_off1(r);
                    hi[1] = lo[1] + rec_ext1(r);
                    lo[2] = b2.x;
                    hi[2] = b2.y;
                    lo[3] = b3.x;
                    hi[3] = b3.y;
                    sat = box_count_wrap_u32(A.prefix_unreached, c.pstride, wd, c.shape[wd], lo, hi) == 0;
                }
            }
        }
        const uint64_t b = __ballot(sat);
        if (lane == 0) bits[pbase >> 6] = b;
    }
    __syncthreads();

    uint64_t nr = 0, nc = 0;
    if (t < n_eval && s_use[t]) {
        const uint32_t first = t * I;
        const uint32_t last = first + I;  // exclusive
        uint32_t chosen = GSC_INVALID_INPUT;
        for (uint32_t wpos = first >> 6; (wpos << 6) < last; ++wpos) {
            uint64_t word = bits[wpos];
            const uint32_t wstart = wpos << 6;
            if (wstart < first) word &= ~0ull << (first - wstart);
            if (wstart + 64 > last) word &= ~0ull >> (wstart + 64 - last);
            if (word) {
                chosen = wstart + static_cast<uint32_t>(__ffsll(static_cast<long long>(word)) - 1) - first;
                break;
            }
        }
        if (chosen != GSC_INVALID_INPUT) {
            const uint32_t st = s_state[t];
            const bool is_cand = A.cand[st] != 0;
            const bool reached = A.reach[st] != 0;
            const bool certified = A.ctrl[st] != GSC_INVALID_INPUT;
            if (!reached) {
                A.reach[st] = 1;
                nr = 1;
            }
            if (!certified) {
                A.ctrl[st] = chosen;
                if (is_cand) {
                    nc = 1;
                } else if (A.rstep[st] == GSC_UNREACHABLE_STEP) {
                    A.rstep[st] = A.iter;
                }
            }
        }
    }
    nr = wave_sum_u64(nr);
    nc = wave_sum_u64(nc);
    if (lane == 0) {
        if (nr) atomicAdd(&s_counts[0], static_cast<unsigned long long>(nr));
        if (nc) atomicAdd(&s_counts[1], static_cast<unsigned long long>(nc));
    }
    __syncthreads();
    if (t == 0) {
        IterCounters& slot = A.ctl->ring[A.iter % kRing];
        if (s_counts[0]) atomicAdd(&slot.new_reachable, s_counts[0]);
        if (s_counts[1]) atomicAdd(&slot.new_certified, s_counts[1]);
        atomicAdd(&slot.pairs_tested, static_cast<unsigned long long>(n_pairs));
    }
}

// Totals and the stop test of one iteration (kernels.cu:722-766). One thread.
__global__ void finish_kernel(IterCtl* ctl, IterRecord* records, uint32_t iter, uint32_t max_iterations) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    if (ctl->done) return;
    IterCounters& slot = ctl->ring[iter % kRing];
    const unsigned long long nr = slot.new_reachable;
    const unsigned long long nc = slot.new_certified;
    ctl->total_reachable += nr;
    ctl->total_certified += nc;
    ctl->iterations = iter;
    IterRecord& rec = records[iter % kRing];
    rec.iteration = iter;
    rec.newly_reachable = nr;
    rec.newly_certified = nc;
    rec.total_reachable = ctl->total_reachable;
    rec.total_certified = ctl->total_certified;
    rec.pairs_tested = slot.pairs_tested;
    slot.new_reachable = 0;
    slot.new_certified = 0;
    slot.pairs_tested = 0;
    uint32_t done = 0;
    if (ctl->total_certified == ctl->candidate_count)
        done = kDoneConverged;
    else if (nr == 0 && nc == 0)
        done = kDoneStalled;
    else if (iter >= max_iterations)
        done = kDoneMaxIterations;
    __atomic_store_n(&ctl->done, done, __ATOMIC_RELAXED);
}

}  // namespace gsc_amd
