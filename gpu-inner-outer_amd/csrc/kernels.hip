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
//                        (kernels.cu:722-766, reachability.hpp:111-152)
//   export_kernel        reference abstraction triplet from the pair records
//
// No MFMA anywhere: nothing here is a dense contraction. The satisfaction pass
// streams one 4-byte record per pair (coalesced) and gathers prefix corners;
// the abstraction pass is scalar FP64 with the trigonometry read from
// host-built tables.
#include <type_traits>

#include "kernels.hpp"

namespace gsc_amd {

namespace {

__device__ inline uint32_t lane_id() { return threadIdx.x & 63u; }
// Number of set bits of the wave mask m below this lane (mbcnt: two
// instructions, against a 64-bit lane mask, an and and two popcounts).
__device__ inline uint32_t lanes_below(uint64_t m) {
    return __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(m >> 32),
                                     __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(m), 0u));
}

// Inclusive prefix sum across a 64-lane wave.
// DPP steps (data moves between lanes inside the VALU, no LDS round trip):
// shifts by 1, 2, 4, 8 within each 16-lane row (lanes shifted in from outside
// the row read 0), then row 0's and row 2's last lane broadcast into rows 1
// and 3, then lane 31 into rows 2 and 3. Rows a step leaves out keep 0 to add.
__device__ inline uint32_t wave_inclusive_scan(uint32_t v) {
    v += static_cast<uint32_t>(__builtin_amdgcn_update_dpp(0, static_cast<int>(v), 0x111, 0xf, 0xf, true));  // row_shr:1
    v += static_cast<uint32_t>(__builtin_amdgcn_update_dpp(0, static_cast<int>(v), 0x112, 0xf, 0xf, true));  // row_shr:2
    v += static_cast<uint32_t>(__builtin_amdgcn_update_dpp(0, static_cast<int>(v), 0x114, 0xf, 0xf, true));  // row_shr:4
    v += static_cast<uint32_t>(__builtin_amdgcn_update_dpp(0, static_cast<int>(v), 0x118, 0xf, 0xf, true));  // row_shr:8
    v += static_cast<uint32_t>(__builtin_amdgcn_update_dpp(0, static_cast<int>(v), 0x142, 0xa, 0xf, false));  // row_bcast:15
    v += static_cast<uint32_t>(__builtin_amdgcn_update_dpp(0, static_cast<int>(v), 0x143, 0xc, 0xf, false));  // row_bcast:31
    return v;
}

// Value of lane 63 (wave-uniform).
__device__ inline uint32_t last_lane(uint32_t v) {
    return static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(v), 63));
}

__device__ inline uint64_t wave_sum_u64(uint64_t v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Pointer into the constant address space: loads through it at wave-uniform
// addresses compile to scalar (s_load) instructions.
template <typename T>
using ConstPtr = const __attribute__((address_space(4))) T*;

template <typename T>
__device__ inline ConstPtr<T> to_const(const T* p) {
    return reinterpret_cast<ConstPtr<T>>(reinterpret_cast<uintptr_t>(p));
}

// Whether any of bits [row, row + ext] (ext <= 31) is set, given the two
// 32-bit words w0 = word row / 32 and w1 = the next: a 32-bit funnel shift
// brings them to bit 0, and shifting left by 31 - ext drops the bits above.
__device__ __forceinline__ bool bits_any(uint32_t w0, uint32_t w1, uint32_t row, uint32_t ext) {
    const uint32_t window = __builtin_amdgcn_alignbit(w1, w0, row & 31u);
    return (window << (31u - ext)) != 0;
}
// The same with the shift 31 - ext given (a record's rec_shift0).
__device__ __forceinline__ bool bits_any_sh(uint32_t w0, uint32_t w1, uint32_t row, uint32_t sh) {
    const uint32_t window = __builtin_amdgcn_alignbit(w1, w0, row & 31u);
    return (window << sh) != 0;
}

__device__ inline bool loop_done(const IterCtl* ctl) {
    return __atomic_load_n(&ctl->done, __ATOMIC_RELAXED) != 0;
}

// `done` and `iterations` are adjacent words of IterCtl: one 64-bit load gives
// both (whether to run, and which iteration this is).
__device__ __forceinline__ uint64_t loop_state(const IterCtl* ctl) {
    static_assert(offsetof(IterCtl, iterations) == offsetof(IterCtl, done) + 4, "done, iterations adjacent");
    return __atomic_load_n(reinterpret_cast<const unsigned long long*>(&ctl->done), __ATOMIC_RELAXED);
}
__device__ __forceinline__ bool state_done(uint64_t st) { return static_cast<uint32_t>(st) != 0; }
__device__ __forceinline__ uint32_t state_next_iteration(uint64_t st) { return static_cast<uint32_t>(st >> 32) + 1u; }

__device__ inline void unflatten(const DevCase& c, uint32_t s, uint32_t idx[4]) {
    // grid.hpp:30-36 with x fastest.
    uint32_t t = s / c.shape[0];
    idx[0] = s - t * c.shape[0];
    uint32_t t2 = t / c.shape[1];
    idx[1] = t - t2 * c.shape[1];
    idx[3] = t2 / c.shape[2];
    idx[2] = t2 - idx[3] * c.shape[2];
}

}  // namespace

// ---------------------------------------------------------------------------
// Masks and seeds. Grid-stride over the states (coalesced byte stores per
// pass); each thread counts its candidates and bad candidates, and each
// workgroup adds its sums with one atomic per counter (per-wave atomics on the
// two counters serialised the kernel: 0.44 ms at 4M states).
// `allow` (may be null) removes states a caller-side pointwise constraint rejects.
__global__ __launch_bounds__(256) void masks_kernel(DevCase c, const gsc_rect4* __restrict__ obstacles,
                                                     const gsc_rect4* __restrict__ targets,
                                                     const uint8_t* __restrict__ allow,
                                                     uint8_t* __restrict__ valid, uint8_t* __restrict__ cand,
                                                     unsigned long long* __restrict__ counts) {
    uint64_t n_cand = 0, n_bad = 0;
    for (uint32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < c.S; s += gridDim.x * blockDim.x) {
        bool is_cand = false, bad = false;
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
        if (ok && allow) ok = allow[s] != 0;
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
        n_cand += is_cand ? 1u : 0u;
        n_bad += bad ? 1u : 0u;
    }
    __shared__ unsigned long long s_sum[2][4];
    n_cand = wave_sum_u64(n_cand);
    n_bad = wave_sum_u64(n_bad);
    const uint32_t wave = threadIdx.x >> 6;
    if (lane_id() == 0) {
        s_sum[0][wave] = n_cand;
        s_sum[1][wave] = n_bad;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long a = 0, b = 0;
        for (uint32_t w = 0; w < blockDim.x / 64u; ++w) {
            a += s_sum[0][w];
            b += s_sum[1][w];
        }
        if (a) atomicAdd(&counts[0], a);
        if (b) atomicAdd(&counts[1], b);
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

// p[0..n) = value (a kernel rather than hipMemsetAsync, so callers may capture it).
__global__ void fill_u32_kernel(uint32_t* __restrict__ p, uint32_t n, uint32_t value) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = value;
}

// Control block back to "iteration 0, not done", counters zeroed (by a kernel,
// never hipMemset, so the reset can sit in a captured graph).
__device__ __forceinline__ void reset_ctl_body(IterCtl* ctl, unsigned long long candidate_count) {
    const uint32_t t = threadIdx.x;
    if (t == 0) {
        ctl->done = 0;
        ctl->iterations = 0;
        ctl->candidate_count = candidate_count;
        ctl->total_reachable = candidate_count;
        ctl->total_certified = 0;
    }
    for (uint32_t k = t; k < kRing * kCounterSlots; k += blockDim.x) {
        const uint32_t r = k / kCounterSlots, j = k % kCounterSlots;
        ctl->ring[r].new_reachable[j] = 0;
        ctl->ring[r].new_certified[j] = 0;
        ctl->ring[r].pairs_tested[j] = 0;
        ctl->ring[r].boxes_counted[j] = 0;
        ctl->ring[r].test_slots[j] = 0;
    }
}

__global__ void reset_ctl_kernel(IterCtl* ctl, unsigned long long candidate_count) {
    reset_ctl_body(ctl, candidate_count);
}

// ---------------------------------------------------------------------------
// Prefix table of (a & !b) over the padded layout. Padding entries (index 0 on
// any axis) are zero from allocation and never written.
// Pass 1: one wave per interior (y,z,w) line; wave scan along x. `block` covers
// 4 lines.
__device__ __forceinline__ void prefix_rows_body(const DevCase& c, const uint8_t* __restrict__ a,
                                                 const uint8_t* __restrict__ b, uint32_t* __restrict__ P,
                                                 uint32_t block) {
    const uint32_t line = block * 4u + (threadIdx.x >> 6);
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
        if (x < c.shape[0]) {
            // Both masks loaded unconditionally: a short-circuit would load b only
            // after a had arrived, two memory round trips instead of one.
            const uint32_t av = a[src + x], bv = b ? b[src + x] : 0u;
            v = ((av != 0) & (bv == 0)) ? 1u : 0u;
        }
        v = wave_inclusive_scan(v) + carry;
        if (x < c.shape[0]) P[dst + x] = v;
        carry = last_lane(v);
    }
}

__global__ __launch_bounds__(256) void prefix_rows_kernel(DevCase c, const uint8_t* __restrict__ a,
                                                           const uint8_t* __restrict__ b, uint32_t* P0, uint32_t* P1,
                                                           const IterCtl* ctl) {
    uint32_t* P = P0;
    if (ctl) {
        const uint64_t st = loop_state(ctl);
        if (state_done(st)) return;
        P = (state_next_iteration(st) & 1u) ? P1 : P0;
    }
    prefix_rows_body(c, a, b, P, blockIdx.x);
}

// Bitmask of valid, unreached states (bit s of word s / 64), for the probe in
// satisfy_kernel. Thread per state (256 per block); lane 0 of each wave stores
// the wave's ballot.
__device__ __forceinline__ void unreached_bits_body(uint32_t S, const uint8_t* __restrict__ valid,
                                                    const uint8_t* __restrict__ reach,
                                                    unsigned long long* __restrict__ bits, uint32_t block) {
    const uint32_t s = block * 256u + threadIdx.x;
    uint32_t vv = 0, rv = 0;
    if (s < S) {
        vv = valid[s];  // both loads issued together (no short-circuit)
        rv = reach[s];
    }
    const bool u = (vv != 0) & (rv == 0);
    const uint64_t b = __ballot(u);
    if (lane_id() == 0 && s < S) bits[s >> 6] = b;
}

__global__ __launch_bounds__(256) void unreached_bits_kernel(uint32_t S, const uint8_t* __restrict__ valid,
                                                              const uint8_t* __restrict__ reach,
                                                              unsigned long long* __restrict__ bits, const IterCtl* ctl) {
    if (ctl && loop_done(ctl)) return;
    unreached_bits_body(S, valid, reach, bits, blockIdx.x);
}

__global__ __launch_bounds__(256) void valid_bits_kernel(uint32_t S, const uint8_t* __restrict__ valid,
                                                          unsigned long long* __restrict__ bits) {
    const uint32_t s = blockIdx.x * 256u + threadIdx.x;
    const uint64_t b = __ballot(s < S && valid[s] != 0);
    if (lane_id() == 0 && s < S) bits[s >> 6] = b;
}

// Pass 2..4: running sum along `axis` (1, 2 or 3); one thread per line (256 per
// block), lines enumerated with x fastest so neighbouring threads touch
// neighbouring words.
// Values of a prefix line loaded together in the axis passes.
constexpr uint32_t kLineBatch = 16;

__device__ __forceinline__ void prefix_axis_body(const DevCase& c, int axis, uint32_t* __restrict__ P,
                                                 uint32_t block) {
    const uint32_t tid = block * 256u + threadIdx.x;
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
    // kLineBatch values of the line are loaded before any is stored, so that
    // one memory round trip covers them (a store could alias a later load of
    // the same line, which would otherwise keep every load behind it).
    uint32_t acc = 0;
    for (uint32_t k = 0; k < n; k += kLineBatch) {
        uint32_t v[kLineBatch];
#pragma unroll
        for (uint32_t j = 0; j < kLineBatch; ++j) v[j] = k + j < n ? P[off + j * step] : 0u;
#pragma unroll
        for (uint32_t j = 0; j < kLineBatch; ++j) {
            acc += v[j];
            if (k + j < n) P[off + j * step] = acc;
        }
        off += kLineBatch * step;
    }
}

__global__ __launch_bounds__(256) void prefix_axis_kernel(DevCase c, int axis, uint32_t* P0, uint32_t* P1,
                                                           const IterCtl* ctl) {
    uint32_t* P = P0;
    if (ctl) {
        const uint64_t st = loop_state(ctl);
        if (state_done(st)) return;
        P = (state_next_iteration(st) & 1u) ? P1 : P0;
    }
    prefix_axis_body(c, axis, P, blockIdx.x);
}

// ---------------------------------------------------------------------------
// Fused prefix passes (planes that fit in LDS): three launches per table build
// instead of rows, three axes, coarse sample and bitmask.
//
// prefix_plane_kernel: one 1024-thread workgroup per (θ, v) plane.
//  1. Stage: the workgroup reads the masks of states [lo, hi) as 4-byte
//     words, every load issued before the first is used, and keeps one bit
//     per state (valid ∧ ¬reached) in LDS; lo is the plane's first state
//     rounded down to 64, hi covers the plane and every unreached-bitmask
//     word that starts in it.
//  2. The plane's words of the unreached bitmask (word k belongs to the plane
//     holding state 64k, including any states of the next plane it covers)
//     are copied out of the staged bits.
//  3. x: a wave per row, 64 cells at a time; a cell's running count is the
//     row's carry plus the set bits below it in the chunk's ballot (mbcnt).
//  4. y: a wave per column in LDS, then the plane's 2D prefix goes into the
//     padded table; the θ and v passes follow (prefix_axis_kernel with axis 2,
//     prefix_last_axis_kernel).
// LDS layout: element (x, y) at y * pitch + (x ^ (y & 63)), pitch a multiple
// of 64, so that rows (lanes along x) and columns (lanes along y) both hit 64
// different banks; the staged bits follow (plane_stage_words).
__device__ __forceinline__ void prefix_plane_body(const DevCase& c, const uint8_t* __restrict__ a,
                                                  const uint8_t* __restrict__ b, uint32_t* __restrict__ P,
                                                  unsigned long long* __restrict__ bits, uint32_t pitch, uint32_t plane) {
    extern __shared__ uint32_t sm[];
    constexpr uint32_t kW = kPlaneThreads / 64;
    const uint32_t lane = lane_id();
    const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const uint32_t n0 = c.shape[0], n1 = c.shape[1], PQ = c.PQ, S = c.S;
    const uint32_t base = plane * PQ;
    uint32_t* const cb = sm + n1 * pitch;
    const uint32_t k_begin = (base + 63u) / 64u;
    const uint32_t k_end = min((base + PQ + 63u) / 64u, (S + 63u) / 64u);
    const uint32_t lo = base & ~63u;
    const uint32_t hi = max(base + PQ, k_end * 64u);  // states past S stage as 0
    const uint32_t nd = (hi - lo + 3u) / 4u;         // staged 4-state words
    // 1. Stage. A plane holds at most kPlaneLdsMax / 4 states, so nd fits in
    // kStageRounds rounds of the workgroup.
    constexpr uint32_t kStageRounds = (kPlaneLdsMax / 4u + 127u + 3u) / 4u / kPlaneThreads + 1u;
    uint32_t av[kStageRounds], bv[kStageRounds];
#pragma unroll
    for (uint32_t r = 0; r < kStageRounds; ++r) {
        const uint32_t s = lo + 4u * (r * kPlaneThreads + threadIdx.x);
        // Whole words in the masks load as one; others read word 0 and are
        // replaced below (S >= 4: word 0 is in the masks).
        const uint32_t at = (s + 4u <= S) ? s : 0u;
        av[r] = S >= 4u ? *reinterpret_cast<const uint32_t*>(a + at) : 0u;
        bv[r] = S >= 4u && b ? *reinterpret_cast<const uint32_t*>(b + at) : 0u;
    }
#pragma unroll
    for (uint32_t r = 0; r < kStageRounds; ++r) {
        const uint32_t i = r * kPlaneThreads + threadIdx.x;
        const uint32_t s = lo + 4u * i;
        uint32_t va = av[r], vb = bv[r];
        if (s + 4u > S) {
            // The masks end inside (or before) this word: its states below S
            // byte by byte, the rest 0.
            va = 0;
            vb = 0;
            for (uint32_t j = 0; s + j < S && j < 4u; ++j) {
                va |= static_cast<uint32_t>(a[s + j]) << (8u * j);
                if (b) vb |= static_cast<uint32_t>(b[s + j]) << (8u * j);
            }
        }
        // Bit 7 of each byte: the byte is non-zero (its low seven bits carry
        // into bit 7, or bit 7 is set). The four cell bits a ≠ 0 ∧ b = 0
        // sit at bits 7, 15, 23, 31; the product moves them to bits 21..24
        // (all partial products land on distinct bits).
        const uint32_t nz_a = (((va & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | va) & 0x80808080u;
        const uint32_t nz_b = (((vb & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | vb) & 0x80808080u;
        const uint32_t m = (nz_a & ~nz_b) >> 7;
        uint32_t nib = ((m * 0x204081u) >> 21) & 0xFu;
        if (i >= nd) nib = 0;
        // Eight consecutive words of states (eight lanes) make one LDS word.
        uint32_t w = nib << (4u * (lane & 7u));
        w |= static_cast<uint32_t>(__shfl_xor(static_cast<int>(w), 1, 64));
        w |= static_cast<uint32_t>(__shfl_xor(static_cast<int>(w), 2, 64));
        w |= static_cast<uint32_t>(__shfl_xor(static_cast<int>(w), 4, 64));
        if ((lane & 7u) == 0 && i < nd) cb[i / 8u] = w;
    }
    __syncthreads();
    // 2. Bitmask words, straight from the staged bits (64k - lo is a multiple of 64).
    if (bits) {
        for (uint32_t k = k_begin + threadIdx.x; k < k_end; k += kPlaneThreads) {
            const uint32_t wi = (k * 64u - lo) / 32u;
            bits[k] = static_cast<unsigned long long>(cb[wi]) | (static_cast<unsigned long long>(cb[wi + 1]) << 32);
        }
    }
    // 3. x: a wave per row, 64 cells at a time: each lane reads its cell's
    // staged bit, the ballot of the chunk gives each lane the count of set
    // bits below it (mbcnt), and the row's carry adds the chunks before.
    for (uint32_t y = wave; y < n1; y += kW) {
        const uint32_t sw = y & 63u;
        const uint32_t row_bit = base + y * n0 - lo;
        uint32_t carry = 0;
        for (uint32_t x0 = 0; x0 < n0; x0 += 64u) {
            const uint32_t x = x0 + lane;
            const uint32_t o = row_bit + x;
            const bool mine = (x < n0) & (((cb[o >> 5] >> (o & 31u)) & 1u) != 0u);
            const uint64_t m = __ballot(mine);
            const uint32_t v = carry + lanes_below(m) + (mine ? 1u : 0u);
            if (x < n0) sm[y * pitch + (x ^ sw)] = v;
            carry += static_cast<uint32_t>(__popcll(m));
        }
    }
    __syncthreads();
    // y: a wave per column, lanes along y.
    for (uint32_t x = wave; x < n0; x += kW) {
        uint32_t carry = 0;
        for (uint32_t y0 = 0; y0 < n1; y0 += 64u) {
            const uint32_t y = y0 + lane;
            const uint32_t at = y * pitch + (x ^ (y & 63u));
            uint32_t v = y < n1 ? sm[at] : 0u;
            v = wave_inclusive_scan(v) + carry;
            if (y < n1) sm[at] = v;
            carry = last_lane(v);
        }
    }
    __syncthreads();
    const uint32_t z = plane % c.shape[2], w = plane / c.shape[2];
    const uint32_t zw = (z + 1) * c.pstride[2] + (w + 1) * c.pstride[3];
    for (uint32_t y = wave; y < n1; y += kW) {
        const uint32_t row = zw + (y + 1) * c.pstride[1] + 1;
        for (uint32_t x = lane; x < n0; x += 64u) P[row + x] = sm[y * pitch + (x ^ (y & 63u))];
    }
}

__global__ __launch_bounds__(kPlaneThreads) void prefix_plane_kernel(DevCase c, const uint8_t* __restrict__ a,
                                                                      const uint8_t* __restrict__ b, uint32_t* P0,
                                                                      uint32_t* P1,
                                                                      unsigned long long* __restrict__ bits,
                                                                      uint32_t pitch, const IterCtl* ctl) {
    uint32_t* P = P0;
    if (ctl) {
        const uint64_t st = loop_state(ctl);
        if (state_done(st)) return;
        P = (state_next_iteration(st) & 1u) ? P1 : P0;
    }
    prefix_plane_body(c, a, b, P, bits, pitch, blockIdx.x);
}

// The last (v) axis pass with the coarse sample fused: the thread of line
// (x, y, θ) writes C[X, Y, θ + 1, v + 1] for every v when x + 1 and y + 1 are
// coarse sample positions (multiples of kCoarse, or the last index); the
// coarse entries on the padding planes stay zero from set-up.
__device__ __forceinline__ void prefix_last_axis_body(const DevCase& c, uint32_t* __restrict__ P,
                                                      uint32_t* __restrict__ C, uint32_t tid) {
    const uint32_t n0 = c.shape[0], n1 = c.shape[1];
    const uint32_t lines = n0 * n1 * c.shape[2];
    if (tid >= lines) return;
    const uint32_t x = tid % n0;
    const uint32_t r = tid / n0;
    const uint32_t y = r % n1;
    const uint32_t z = r / n1;
    const uint32_t step = c.pstride[3];
    uint32_t off = (x + 1) + (y + 1) * c.pstride[1] + (z + 1) * c.pstride[2] + step;
    const bool sx = (x + 1) % kCoarse == 0 || x + 1 == n0;
    const bool sy = (y + 1) % kCoarse == 0 || y + 1 == n1;
    const bool sample = sx && sy;
    uint32_t coff = 0;
    if (sample)
        coff = (x + kCoarse) / kCoarse + (y + kCoarse) / kCoarse * c.cstride[1] + (z + 1) * c.cstride[2] + c.cstride[3];
    const uint32_t n = c.shape[3];
    uint32_t acc = 0;
    for (uint32_t k = 0; k < n; k += kLineBatch) {
        uint32_t v[kLineBatch];
#pragma unroll
        for (uint32_t j = 0; j < kLineBatch; ++j) v[j] = k + j < n ? P[off + j * step] : 0u;
#pragma unroll
        for (uint32_t j = 0; j < kLineBatch; ++j) {
            acc += v[j];
            if (k + j < n) {
                P[off + j * step] = acc;
                if (sample) C[coff + j * c.cstride[3]] = acc;
            }
        }
        off += kLineBatch * step;
        coff += kLineBatch * c.cstride[3];
    }
}

__global__ __launch_bounds__(256) void prefix_last_axis_kernel(DevCase c, uint32_t* P0, uint32_t* P1, uint32_t* C0,
                                                                uint32_t* C1, const IterCtl* ctl) {
    uint32_t* P = P0;
    uint32_t* C = C0;
    if (ctl) {
        const uint64_t st = loop_state(ctl);
        if (state_done(st)) return;
        const bool odd = (state_next_iteration(st) & 1u) != 0;
        P = odd ? P1 : P0;
        C = odd ? C1 : C0;
    }
    prefix_last_axis_body(c, P, C, blockIdx.x * 256u + threadIdx.x);
}

// Coarse sample: C[X, Y, z, w] = P[min(kCoarse X, n_x), min(kCoarse Y, n_y), z, w]
// over the padded θ / v indices. Thread per coarse entry.
__device__ __forceinline__ void coarse_body(const DevCase& c, const uint32_t* __restrict__ P, uint32_t* __restrict__ C,
                                            uint32_t t) {
    const uint32_t cx = c.cshape[0], cy = c.cshape[1];
    const uint32_t n = cx * cy * (c.shape[2] + 1) * (c.shape[3] + 1);
    if (t >= n) return;
    const uint32_t X = t % cx;
    uint32_t r = t / cx;
    const uint32_t Y = r % cy;
    r /= cy;
    const uint32_t z = r % (c.shape[2] + 1);
    const uint32_t w = r / (c.shape[2] + 1);
    const uint32_t x = min(X * kCoarse, c.shape[0]);
    const uint32_t y = min(Y * kCoarse, c.shape[1]);
    C[t] = P[x + y * c.pstride[1] + z * c.pstride[2] + w * c.pstride[3]];
}

__global__ __launch_bounds__(256) void coarse_kernel(DevCase c, const uint32_t* P0, const uint32_t* P1, uint32_t* C0,
                                                      uint32_t* C1, const IterCtl* ctl) {
    bool odd = false;
    if (ctl) {
        const uint64_t st = loop_state(ctl);
        if (state_done(st)) return;
        odd = (state_next_iteration(st) & 1u) != 0;
    }
    coarse_body(c, odd ? P1 : P0, odd ? C1 : C0, blockIdx.x * 256u + threadIdx.x);
}

// ---------------------------------------------------------------------------
// Pair records are stored plane-major, in 8-input blocks (rec_offset) with
// plane = θ cell + n_θ * v cell and q = x + n_x * y (PQ = n_x * n_y cells per
// plane). A wave owns 64 consecutive cells q of one plane and walks the
// inputs in order, so
//   - record loads and stores are 64 consecutive words,
//   - the θ / v cell ranges (box2, box3) and the trajectory tables are
//     wave-uniform (scalar loads),
//   - the 16 prefix corners of neighbouring lanes are neighbouring words.
__device__ inline void wave_task(const DevCase& c, uint32_t task, uint32_t* plane, uint32_t* q0) {
    *plane = task / c.chunks_per_plane;
    *q0 = (task - *plane * c.chunks_per_plane) * 64u;
}

// Per-step position increments of the nominal trajectory, model.hpp:36-42:
//   inc[(plane * I + input) * K + k] = (dt*(v_k*cos θ_k), dt*(v_k*sin θ_k)).
// They depend on (θ cell, v cell, input) only, so they are computed once here
// and every pair of a plane adds the same increments to its own x0, y0 — the
// reference's `out[0] += dt * dx` with identical operands and rounding.
__global__ __launch_bounds__(256) void increments_kernel(DevCase c, const double2* __restrict__ cs,
                                                          const double* __restrict__ vt, double2* __restrict__ inc) {
    const uint64_t t = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    const uint64_t total = static_cast<uint64_t>(c.shape[2]) * c.shape[3] * c.I * c.K;
    if (t >= total) return;
    const uint32_t k = static_cast<uint32_t>(t % c.K);
    const uint64_t r = t / c.K;
    const uint32_t i = static_cast<uint32_t>(r % c.I);
    const uint32_t plane = static_cast<uint32_t>(r / c.I);
    const uint32_t iz = plane % c.shape[2];
    const uint32_t iv = plane / c.shape[2];
    const uint32_t iw = i / c.n_a;
    const uint32_t ia = i - iw * c.n_a;
    const double2 t2 = cs[static_cast<size_t>(iz * c.n_w + iw) * c.K + k];
    const double vk = vt[static_cast<size_t>(iv * c.n_a + ia) * c.K + k];
    const double dx = vk * t2.x;
    const double dy = vk * t2.y;
    inc[t] = double2{c.dt * dx, c.dt * dy};
}

// Per (plane, input), for the filtered fast path of pair_box
// (dim_to_cells_fast): with S the sum of the K increments in step order and
// rad the plane's radius, (S - rad) / eta and (S + rad) / eta for x and for y.
// Thread per (plane, input).
__global__ __launch_bounds__(256) void increment_sums_kernel(DevCase c, const double2* __restrict__ inc,
                                                              const double* __restrict__ rad, double* __restrict__ sums,
                                                              int32_t* __restrict__ offs) {
    const uint64_t t = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= static_cast<uint64_t>(c.shape[2]) * c.shape[3] * c.I) return;
    const uint32_t plane = static_cast<uint32_t>(t / c.I);
    const uint32_t iv = plane / c.shape[2];
    const double2* row = inc + t * c.K;
    double sx = 0.0, sy = 0.0;
    for (uint32_t k = 0; k < c.K; ++k) {
        const double2 v = row[k];
        sx = sx + v.x;
        sy = sy + v.y;
    }
    const double r0 = rad[2 * iv], r1 = rad[2 * iv + 1];
    const double q[4] = {(sx - r0) * c.inv_eta[0], (sx + r0) * c.inv_eta[0], (sy - r1) * c.inv_eta[1],
                         (sy + r1) * c.inv_eta[1]};
    for (int k = 0; k < 4; ++k) sums[4 * t + k] = q[k];
    // The same edges as cell offsets shared by every cell of the plane
    // (cell_offsets), or kOffsetsUndecided.
    for (int d = 0; d < 2; ++d) {
        CellOffsets o;
        if (!cell_offsets(c, d, q[2 * d], q[2 * d + 1], &o)) o = CellOffsets{kOffsetsUndecided, 0, 0, 0};
        offs[8 * t + 4 * d] = o.off_lo;
        offs[8 * t + 4 * d + 1] = o.off_hi;
        offs[8 * t + 4 * d + 2] = o.rej_lo;
        offs[8 * t + 4 * d + 3] = o.rej_hi;
    }
}

// Per (plane, input) of the padded input range, once per abstraction, after
// increment_sums_kernel:
//   probe_base  θ/v part of the probe row: the bit where the plane
//               lo2 + n_θ * lo3 of the pair's successor box starts in the probe
//               bitmasks. A box whose θ range holds lo2 and the next θ cell (the
//               range is longer than one cell, or wraps) and whose v range holds
//               lo3 and the next v cell probes the OR bitmask, which covers those
//               four slices, when it is built (use_or). 0 for padding inputs and
//               rejected ranges, whose records are invalid and never probed.
//   row_kind    how the abstraction kernel builds the row's records (kRow*).
//   row_box     the θ and v ranges (inputs below I only).
//   fast_offs   FastRow of the row: its cell offsets, and the cells whose box
//               stays in the domain as a start and a span; a row with no such
//               cell (an invalid row, a padding input) gets start kFastNoCell
//               and span 0, so that no cell passes.
// and per plane, plane_mixed: 0 when the abstraction kernel can build every
// row of the plane as a fast row, with the constraint rectangle test on every
// row when the case has a constraint (the rows are kRowFastChecked or
// kRowInvalid), without it otherwise (kRowFast or kRowInvalid); else 1.
// A row is kRowFast when every one of its pairs follows from the row's cell
// offsets alone: the offsets decided both axes, the x/y wrap is off, the θ/v
// ranges are ordered or wrap on their own axis, and the unclamped offsets and
// extents fit the record fields. Clamping a cell's range to the grid moves
// each edge offset toward 0 and shortens the extent, so the clamped box of
// every cell fits as well. Under a constraint the row also needs the rectangle
// test unless the reference's constraint walk is empty for it (an inverted v
// range, or an inverted θ range that does not wrap), then kRowFastChecked.
// One wave per plane.
__global__ __launch_bounds__(256) void pair_rows_kernel(DevCase c, const uint2* __restrict__ box2,
                                                         const uint2* __restrict__ box3,
                                                         const int32_t* __restrict__ offs, uint32_t use_or,
                                                         uint32_t* __restrict__ probe_base,
                                                         uint32_t* __restrict__ row_kind, uint4* __restrict__ row_box,
                                                         FastRow* __restrict__ fast_offs,
                                                         uint32_t* __restrict__ plane_mixed,
                                                         uint32_t* __restrict__ ext1_max) {
    const uint32_t plane = blockIdx.x * 4u + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    if (plane >= c.shape[2] * c.shape[3]) return;
    const uint32_t lane = lane_id();
    const uint32_t iz = plane % c.shape[2], iv = plane / c.shape[2];
    const int wd = c.wrap_dim;
    const bool constrained = c.constraint_type != GSC_CONSTRAINT_NONE;
    bool mixed = false;
    uint32_t ext1 = 0;  // the fast rows' largest unclamped y extent
    for (uint32_t i = lane; i < c.I_pad; i += 64) {
        uint32_t v = 0, kind = kRowInvalid;
        FastRow f{{0, 0}, {0, 0}, {kFastNoCell, kFastNoCell}, {0, 0}};
        if (i < c.I) {
            const uint32_t iw = i / c.n_a;
            const uint2 b2 = box2[iz * c.n_w + iw];
            const uint2 b3 = box3[iv * c.n_a + (i - iw * c.n_a)];
            row_box[static_cast<size_t>(plane) * c.I + i] = make_uint4(b2.x, b2.y, b3.x, b3.y);
            if (b2.x != kBoxReject && b3.x != kBoxReject) {
                const uint32_t pz = b2.x + c.shape[2] * b3.x;
                const bool span = b2.y != b2.x && b3.y != b3.x;
                v = use_or && span ? probe_or_bit0(c) + pz * probe_or_pitch(c) : pz * c.PQ;
                const int32_t* const o = offs + (static_cast<size_t>(plane) * c.I + i) * 8;
                const bool inv2 = b2.x > b2.y, inv3 = b3.x > b3.y;
                bool fast = wd >= 2 && (!inv2 || wd == 2) && (!inv3 || wd == 3);
                int64_t m[2] = {0, 0};  // largest |offset| per axis
                for (int d = 0; d < 2; ++d) {
                    const int32_t lo = o[4 * d], hi = o[4 * d + 1];
                    fast = fast && lo != kOffsetsUndecided && lo <= hi && hi - lo <= static_cast<int32_t>(kExtMax);
                    m[d] = max(lo < 0 ? -static_cast<int64_t>(lo) : lo, hi < 0 ? -static_cast<int64_t>(hi) : hi);
                }
                // Every clamped box's middle-row offset fits the record: its x
                // offset and its middle y offset (between its y offsets) are
                // bounded by the unclamped ones.
                fast = fast && m[0] + m[1] * c.shape[0] <= kRelMax;
                const bool empty = inv3 || (inv2 && wd != 2);
                kind = !fast ? kRowGeneral : constrained && !empty ? kRowFastChecked : kRowFast;
                if (fast) {
                    ext1 = max(ext1, static_cast<uint32_t>(o[5] - o[4]));
                    // Cells i with rej_lo <= i <= rej_hi, within [0, n - 1].
                    bool any_cell = true;
                    for (int d = 0; d < 2; ++d) {
                        const int32_t n = static_cast<int32_t>(c.shape[d]);
                        const int32_t a = max(o[4 * d + 2], 0), b = min(o[4 * d + 3], n - 1);
                        f.off_lo[d] = o[4 * d];
                        f.off_hi[d] = o[4 * d + 1];
                        f.start[d] = a;
                        f.span[d] = static_cast<uint32_t>(b - a);
                        any_cell = any_cell && a <= b;
                    }
                    if (!any_cell) f = FastRow{{0, 0}, {0, 0}, {kFastNoCell, kFastNoCell}, {0, 0}};
                }
            }
        }
        const size_t t = static_cast<size_t>(plane) * c.I_pad + i;
        probe_base[t] = v;
        row_kind[t] = kind;
        fast_offs[t] = f;
        mixed |= kind == kRowGeneral || (constrained && kind == kRowFast);
    }
    const bool any = __ballot(mixed) != 0;
    if (lane == 0) plane_mixed[plane] = any ? 1u : 0u;
    // Clamping to the grid only shortens an extent, so this bounds every
    // compact box of the plane's fast rows.
    for (int off = 32; off > 0; off >>= 1) ext1 = max(ext1, static_cast<uint32_t>(__shfl_xor(ext1, off, 64)));
    if (lane == 0 && ext1 != 0) atomicMax(ext1_max, ext1);
}

// The kernel's own argument block, addressed as ordinary (global) memory, so
// that out-of-line helpers can read the case from it without the caller
// keeping its fields in registers.
template <typename T>
__device__ inline const T* kernarg_block() {
    return reinterpret_cast<const T*>(reinterpret_cast<uintptr_t>(__builtin_amdgcn_kernarg_segment_ptr()));
}

// Exact x / y cells of one pair: the reference's step-by-step sum
// (model.hpp:40-41, additions strictly in step order) and dim_to_cells.
// Out of line on purpose: it runs only where the filtered fast path cannot
// decide, and keeping its constants out of the callers' input loops leaves the
// scalar registers to the fast path. `cg` points at the case in memory.
struct ExactCells {
    uint32_t lo0, hi0, lo1, hi1;
    uint32_t ok;
};

__device__ __noinline__ ExactCells exact_cells(const DevCase* __restrict__ cg, const double2* __restrict__ row,
                                               double x0, double y0, double rad0, double rad1) {
    const DevCase& c = *cg;
    double px = x0, py = y0;
    for (uint32_t k = 0; k < c.K; ++k) {
        const double2 v = row[k];
        px = px + v.x;
        py = py + v.y;
    }
    ExactCells e{};
    e.ok = (dim_to_cells(c, 0, px - rad0, px + rad0, &e.lo0, &e.hi0) &&
            dim_to_cells(c, 1, py - rad1, py + rad1, &e.lo1, &e.hi1))
               ? 1u
               : 0u;
    return e;
}

// The x/y cells of a pair by the per-cell filter (dim_to_cells_fast), and by
// the exact step-by-step sum where the filter cannot decide: out of line, as
// the (plane, input) cell offsets decide almost every pair, and the case is
// read from the argument block (`cg`) rather than held in registers.
__device__ __noinline__ ExactCells filtered_cells(const DevCase* __restrict__ cg, const double2* __restrict__ inc,
                                                  const double* __restrict__ sums, double x0, double y0, double rad0,
                                                  double rad1) {
    const DevCase& c = *cg;
    ExactCells e{0, 0, 0, 0, 0};
    int rx = kCellsUndecided, ry = kCellsUndecided;
    if (c.wrap_dim >= 2) {
        rx = dim_to_cells_fast(c, 0, (x0 - c.lb[0]) * c.inv_eta[0], sums[0], sums[1], &e.lo0, &e.hi0);
        ry = dim_to_cells_fast(c, 1, (y0 - c.lb[1]) * c.inv_eta[1], sums[2], sums[3], &e.lo1, &e.hi1);
    }
    if (rx == kCellsReject || ry == kCellsReject) {
        e.ok = 0;
        return e;
    }
    if (rx == kCellsUndecided || ry == kCellsUndecided) return exact_cells(cg, inc, x0, y0, rad0, rad1);
    e.ok = 1;
    return e;
}

// Successor cell box of one pair, the reference's arithmetic in the reference's order:
//   centre           grid.hpp:42 (x0, y0 passed in)
//   nominal x, y     model.hpp:35-42: x += dt*(v_k*cos θ_k), y += dt*(v_k*sin θ_k),
//                    the increments read from increments_kernel's table
//   radius           model.hpp:71-73, 84-89, 107-111 (per v cell, host table)
//   cells            grid.hpp:181-251 (dims 0,1 here; dims 2,3 from box2/box3)
//   constraint scan  abstraction.hpp:110-159 as a 2D prefix query over x,y fails
// Returns false for an invalid pair (abstraction.hpp:187-205). `inc` points at
// the pair's (plane, input) row of K increments and must be wave-uniform.
__device__ __forceinline__ bool pair_box(const DevCase& c, const DevCase* cg, const double2* inc, const double* sums,
                                         const int32_t offs[8], uint2 b2, uint2 b3, uint32_t ix, uint32_t iy, double x0,
                                         double y0, double rad0, double rad1, const uint32_t* __restrict__ fail2d,
                                         uint32_t lo[4], uint32_t hi[4]) {
    lo[2] = b2.x;
    hi[2] = b2.y;
    lo[3] = b3.x;
    hi[3] = b3.y;
    if (b2.x == kBoxReject || b3.x == kBoxReject) {
        lo[0] = hi[0] = lo[1] = hi[1] = 0;
        return false;
    }
    // The branches below are wave-uniform (`inc` and the offsets belong to one
    // (plane, input)); the per-lane tests are selects, so a wave spends no
    // exec-mask bookkeeping per pair. Fast path (x and y are plain intervals
    // when the wrap axis is θ or v): the (plane, input)'s cell offsets, integer
    // arithmetic only, when they decided; else filtered_cells.
    bool ok;
    if (c.wrap_dim >= 2 && offs[0] != kOffsetsUndecided && offs[4] != kOffsetsUndecided) {
        const CellOffsets ox{offs[0], offs[1], offs[2], offs[3]};
        const CellOffsets oy{offs[4], offs[5], offs[6], offs[7]};
        const bool ok_x = cells_from_offsets(ox, static_cast<int32_t>(ix), c.shape[0], &lo[0], &hi[0]);
        const bool ok_y = cells_from_offsets(oy, static_cast<int32_t>(iy), c.shape[1], &lo[1], &hi[1]);
        ok = ok_x & ok_y;
    } else {
        const ExactCells e = filtered_cells(cg, inc, sums, x0, y0, rad0, rad1);
        ok = e.ok != 0;
        // A rejected pair's cells are unspecified: zero them, so the table
        // reads below stay in bounds.
        lo[0] = ok ? e.lo0 : 0u;
        hi[0] = ok ? e.hi0 : 0u;
        lo[1] = ok ? e.lo1 : 0u;
        hi[1] = ok ? e.hi1 : 0u;
    }
    if (c.constraint_type != GSC_CONSTRAINT_NONE) {
        // Every loop of abstraction.hpp:124-156 must be non-empty for a cell to be
        // tested; the test reads x and y only, so it reduces to "no failing (x, y)
        // cell in the x/y rectangle". The θ walk is empty only when its range is
        // inverted without a wrap (abstraction.hpp:125-135). Cells lie in
        // [0, shape - 1], so the four reads are in bounds on every lane.
        // Bitwise, not short-circuit, so the θ / v terms stay scalar.
        const int wd = c.wrap_dim;
        const bool inv2 = b2.x > b2.y, inv3 = b3.x > b3.y;
        const bool has_wrap = ((wd == 2) & inv2) | ((wd == 3) & inv3) | ((wd == 1) & (lo[1] > hi[1])) |
                              ((wd == 0) & (lo[0] > hi[0]));
        const bool empty = inv3 | (inv2 & !has_wrap) | (lo[1] > hi[1]) | (lo[0] > hi[0]);
        // 32-bit byte offsets (the table is under 4 GiB) against the table's base.
        const uint32_t sh = c.fail_shift + 2;
        const uint32_t r0 = lo[1] << sh, r1 = (hi[1] + 1) << sh;
        const uint32_t c0 = lo[0] << 2, c1 = (hi[0] + 1) << 2;
        const char* const fb = reinterpret_cast<const char*>(fail2d);
        const auto at = [fb](uint32_t off) { return *reinterpret_cast<const uint32_t*>(fb + off); };
        const uint32_t f = at(r1 + c1) - at(r1 + c0) - at(r0 + c1) + at(r0 + c0);
        ok = ok & (empty | (f == 0));
    }
    return ok;
}

// The state's own centre must satisfy the constraint (abstraction.hpp:177-181).
__device__ __forceinline__ bool state_constraint_ok(const DevCase& c, const uint32_t* __restrict__ fail2d, uint32_t ix,
                                                    uint32_t iy) {
    if (c.constraint_type == GSC_CONSTRAINT_NONE) return true;
    const uint32_t r0 = iy << c.fail_shift, r1 = (iy + 1) << c.fail_shift;
    const uint32_t f = fail2d[r1 + ix + 1] - fail2d[r1 + ix] - fail2d[r0 + ix + 1] + fail2d[r0 + ix];
    return f == 0;
}

// A box that does not fit the compact record: an overflow entry with its flat
// corners and the dead bit settled now; the record is kRecEscape | entry, or
// kRecInvalid when the overflow table is full (the host sees the count and
// reruns).
__device__ __forceinline__ uint32_t escape_record(const AbstractionArgs* __restrict__ Ag, uint32_t lo0, uint32_t lo1,
                                               uint32_t lo2, uint32_t lo3, uint32_t hi0, uint32_t hi1, uint32_t hi2,
                                               uint32_t hi3) {
    const AbstractionArgs& A = *Ag;
    const DevCase& c = A.c;
    const uint32_t lo[4] = {lo0, lo1, lo2, lo3}, hi[4] = {hi0, hi1, hi2, hi3};
    const int wd = c.wrap_dim;
    const int64_t vc = box_count_i64(A.prefix_valid, c.pstride, wd, c.shape[wd], lo, hi);
    const uint32_t e = atomicAdd(A.overflow_count, 1u);
    if (e >= A.overflow_capacity) return kRecInvalid;
    OverflowEntry o;
    o.min_flat = lo[0] * c.stride[0] + lo[1] * c.stride[1] + lo[2] * c.stride[2] + lo[3] * c.stride[3];
    o.max_flat = hi[0] * c.stride[0] + hi[1] * c.stride[1] + hi[2] * c.stride[2] + hi[3] * c.stride[3];
    o.dead = vc == 0 ? 1u : 0u;
    o.reserved = 0;
    A.overflow[e] = o;
    return kRecEscape | e;
}

// Median of three signed integers in one instruction (a is wave-uniform).
__device__ __forceinline__ int32_t med3_i32(int32_t a, int32_t b, int32_t c) {
    int32_t r;
    asm("v_med3_i32 %0, %1, %2, %3" : "=v"(r) : "s"(a), "v"(b), "v"(c));
    return r;
}

// A lane's running hull of its state's valid boxes (see gsc_device.hpp).
struct LaneHull {
    int x0, x1, y0, y1;
    bool any_valid, unrepresentable;
    int e1;  // largest y extent of the lane's compact records
};

// A lane's cell as the fast rows use it: a range clamped to the grid and taken
// relative to the cell is clamp(i + off, 0, n - 1) - i = med3(off, -i, n - 1 - i).
struct FastLane {
    int32_t ix, iy, nix, tix, niy, tiy, n0;
    uint32_t fail_row_shift;
    const char* fail_bytes;
    bool state_ok;
};

__device__ __forceinline__ FastLane fast_lane(const DevCase& c, const uint32_t* fail2d, uint32_t ix, uint32_t iy,
                                              bool state_ok) {
    FastLane L;
    L.ix = static_cast<int32_t>(ix);
    L.iy = static_cast<int32_t>(iy);
    L.nix = -L.ix;
    L.tix = static_cast<int32_t>(c.shape[0]) - 1 - L.ix;
    L.niy = -L.iy;
    L.tiy = static_cast<int32_t>(c.shape[1]) - 1 - L.iy;
    L.n0 = static_cast<int32_t>(c.shape[0]);
    L.fail_row_shift = c.fail_shift + 2;
    L.fail_bytes = reinterpret_cast<const char*>(fail2d);
    L.state_ok = state_ok;
    return L;
}

// The record of one pair of a fast row `o` (kRowFast, or kRowFastChecked with
// `checked`): pair_box's offsets path and the record packing, reduced to the
// clamps, as every box of the row fits a record; the hull takes valid boxes.
// A row no cell qualifies for (FastRow) gives kRecInvalid on every lane.
__device__ __forceinline__ uint32_t fast_record(const ConstPtr<int32_t> o, bool checked, const FastLane& L,
                                                LaneHull& H) {
    // o: off_lo x, y, off_hi x, y, start x, y, span x, y.
    const int32_t d0 = med3_i32(o[0], L.nix, L.tix), e0 = med3_i32(o[2], L.nix, L.tix);
    const int32_t d1 = med3_i32(o[1], L.niy, L.tiy), e1 = med3_i32(o[3], L.niy, L.tiy);
    bool ok = L.state_ok & (static_cast<uint32_t>(L.ix - o[4]) <= static_cast<uint32_t>(o[6])) &
              (static_cast<uint32_t>(L.iy - o[5]) <= static_cast<uint32_t>(o[7]));
    if (checked) {
        // pair_box's constraint rectangle test; the clamped cells keep the
        // four reads in bounds on every lane.
        const uint32_t r0 = static_cast<uint32_t>(L.iy + d1) << L.fail_row_shift;
        const uint32_t r1 = static_cast<uint32_t>(L.iy + e1 + 1) << L.fail_row_shift;
        const uint32_t c0 = static_cast<uint32_t>(L.ix + d0) << 2;
        const uint32_t c1 = static_cast<uint32_t>(L.ix + e0 + 1) << 2;
        const char* const fb = L.fail_bytes;
        const auto at = [fb](uint32_t off) { return *reinterpret_cast<const uint32_t*>(fb + off); };
        const uint32_t f = at(r1 + c1) - at(r1 + c0) - at(r0 + c1) + at(r0 + c0);
        ok = ok & (f == 0);
    }
    H.any_valid |= ok;
    if (ok) {
        H.x0 = min(H.x0, d0);
        H.x1 = max(H.x1, e0);
        H.y0 = min(H.y0, d1);
        H.y1 = max(H.y1, e1);
    }
    const uint32_t ext1 = static_cast<uint32_t>(e1 - d1);
    // 24-bit product: the row's offsets fit kRelMax (pair_rows_kernel).
    const int rel = d0 + __mul24(d1 + static_cast<int32_t>(ext1 >> 1), L.n0);
    return ok ? rec_pack(rel, static_cast<uint32_t>(e0 - d0), ext1, false) : kRecInvalid;
}

// One lane's record block (inputs i0 .. i0 + 7 of cell q) as two 16-byte stores.
constexpr uint32_t kAbsGroup = 8;
__device__ __forceinline__ void store_block(uint32_t* rec_plane, uint32_t PQ, uint32_t q, uint32_t i0,
                                            const uint32_t (&rr)[kAbsGroup]) {
    uint4* const dst = reinterpret_cast<uint4*>(rec_plane + rec_offset(PQ, q, i0));
    dst[0] = make_uint4(rr[0], rr[1], rr[2], rr[3]);
    dst[1] = make_uint4(rr[4], rr[5], rr[6], rr[7]);
}

// A plane whose rows all build as fast rows (plane_mixed 0): no per-row kind,
// every row of the padded range through fast_record, kChecked for all of them
// under a constraint.
template <bool kChecked>
__device__ __forceinline__ void fast_plane(ConstPtr<int32_t> rows, uint32_t I, bool put, uint32_t* rec_plane,
                                           uint32_t PQ, uint32_t q, const FastLane& L, LaneHull& H) {
    for (uint32_t i0 = 0; i0 < I; i0 += kAbsGroup, rows += 8 * kAbsGroup) {
        uint32_t rr[kAbsGroup];
#pragma unroll
        for (uint32_t j = 0; j < kAbsGroup; ++j) rr[j] = fast_record(rows + 8 * j, kChecked, L, H);
        if (put) store_block(rec_plane, PQ, q, i0, rr);
    }
}

// Abstraction: every pair's box, packed into its record (see gsc_device.hpp).
// The `dead` bit of compact records is left 0 here and settled lazily by the
// satisfaction pass (see satisfy_kernel); escaped pairs get it eagerly.
// Two instantiations, launched one after the other over the same waves: kFast
// builds the planes whose rows are all fast rows (plane_mixed 0), the other
// every other plane, so each is compiled for its own loop alone.
template <bool kFast>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8))) void abstraction_kernel(AbstractionArgs A) {
    const DevCase& c = A.c;
    const uint32_t lane = lane_id();
    const uint32_t task = blockIdx.x * 4u + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    if (task >= c.n_tasks) return;
    uint32_t plane, q0;
    wave_task(c, task, &plane, &q0);
    if ((A.plane_mixed[plane] == 0) != kFast) return;
    const uint32_t iv = plane / c.shape[2];
    const uint32_t q = q0 + lane;
    const bool live = q < c.PQ;
    const uint32_t iy = live ? q / c.shape[0] : 0;
    const uint32_t ix = live ? q - iy * c.shape[0] : 0;
    const double x0 = cell_center(c.lb[0], ix, c.eta[0]);
    const double y0 = cell_center(c.lb[1], iy, c.eta[1]);
    const bool state_ok = live && state_constraint_ok(c, A.fail2d, ix, iy);
    const double rad0 = A.rad[2 * iv];
    const double rad1 = A.rad[2 * iv + 1];
    const int wd = c.wrap_dim;
    uint32_t* const rec_plane = A.rec + static_cast<uint64_t>(plane) * c.I_pad * c.PQ;
    const bool store = A.rec != nullptr;  // false: state hulls only
    // Running hull of the state's valid boxes (see gsc_device.hpp).
    LaneHull H{kHullMaxOffset + 1, -kHullMaxOffset - 1, kHullMaxOffset + 1, -kHullMaxOffset - 1, false, false, 0};
    // Per-input row data is wave-uniform: scalar loads through the constant
    // address space (a plain load could alias the record stores below, and
    // would be a vector load waited for on every input).
    static_assert(kRecBlock == kAbsGroup && kInputBlock % kAbsGroup == 0, "a group is one record block");
    const bool wave_states_ok = __ballot(state_ok) != 0;
    const size_t row0 = static_cast<size_t>(plane) * c.I;
    const ConstPtr<int32_t> fast_rows =
        to_const(reinterpret_cast<const int32_t*>(A.fast_offs + static_cast<size_t>(plane) * c.I_pad));
    const FastLane L = fast_lane(c, A.fail2d, ix, iy, state_ok);
    const bool put = live && store;
    // Inputs a record block (8) at a time: the pairs' scalar loads and
    // arithmetic interleave, and their records leave as two 16-byte stores per
    // lane, padding inputs (I to I_pad) included as kRecInvalid.
    if (!wave_states_ok) {
        // The wave's states all fail the constraint (or lie past the plane):
        // no valid pair.
        uint32_t rr[kAbsGroup];
#pragma unroll
        for (uint32_t j = 0; j < kAbsGroup; ++j) rr[j] = kRecInvalid;
        for (uint32_t i0 = 0; i0 < c.I; i0 += kAbsGroup)
            if (put) store_block(rec_plane, c.PQ, q, i0, rr);
    } else if (kFast) {
        if (c.constraint_type != GSC_CONSTRAINT_NONE)
            fast_plane<true>(fast_rows, c.I, put, rec_plane, c.PQ, q, L, H);
        else
            fast_plane<false>(fast_rows, c.I, put, rec_plane, c.PQ, q, L, H);
    } else {
        const ConstPtr<uint32_t> kinds = to_const(A.row_kind + static_cast<size_t>(plane) * c.I_pad);
        const ConstPtr<int32_t> offs_rows = to_const(A.cell_offs + row0 * 8);
        const ConstPtr<uint32_t> box_rows = to_const(reinterpret_cast<const uint32_t*>(A.row_box + row0));
        for (uint32_t i0 = 0; i0 < c.I; i0 += kAbsGroup) {
            uint32_t rr[kAbsGroup];
#pragma unroll
            for (uint32_t j = 0; j < kAbsGroup; ++j) {
                const uint32_t i = i0 + j;
                const uint32_t kind = kinds[i];
                rr[j] = kRecInvalid;
                if (kind == kRowInvalid) continue;
                if (kind != kRowGeneral) {
                    rr[j] = fast_record(fast_rows + 8 * i, kind == kRowFastChecked, L, H);
                    continue;
                }
                const uint2 b2 = make_uint2(box_rows[4 * i], box_rows[4 * i + 1]);
                const uint2 b3 = make_uint2(box_rows[4 * i + 2], box_rows[4 * i + 3]);
                uint32_t lo[4], hi[4];
                const double2* const inc = A.inc + (row0 + i) * c.K;
                const double* const sums = A.inc_sums + (row0 + i) * 4;
                // The (plane, input)'s cell offsets in one scalar load, ahead of the branches that use them.
                int32_t offs[8];
#pragma unroll
                for (int k = 0; k < 8; ++k) offs[k] = offs_rows[8 * i + k];
                const bool ok = pair_box(c, &kernarg_block<AbstractionArgs>()->c, inc, sums, offs, b2, b3, ix, iy, x0,
                                         y0, rad0, rad1, A.fail2d, lo, hi) &
                                state_ok;
                // Per-lane outcome as selects: inverted ranges on a non-wrap axis or
                // wrapped x/y ranges take the escape path.
                const bool regular = (lo[0] <= hi[0]) & (lo[1] <= hi[1]) & ((b2.x <= b2.y) | (wd == 2)) &
                                     ((b3.x <= b3.y) | (wd == 3));
                const int off0 = static_cast<int>(lo[0]) - static_cast<int>(ix);
                const int off1 = static_cast<int>(lo[1]) - static_cast<int>(iy);
                const uint32_t ext0 = hi[0] - lo[0], ext1 = hi[1] - lo[1];
                const int64_t rel = rec_rel_of(off0, off1, ext1, c.shape[0]);
                const bool fits = regular & (rel >= kRelMin) & (rel <= kRelMax) & (ext0 <= kExtMax) &
                                  (ext1 <= kExtMax);
                const bool packed = ok & fits;
                H.any_valid |= ok;
                H.unrepresentable |= ok & !fits;
                H.x0 = packed ? min(H.x0, off0) : H.x0;
                H.x1 = packed ? max(H.x1, off0 + static_cast<int>(ext0)) : H.x1;
                H.y0 = packed ? min(H.y0, off1) : H.y0;
                H.y1 = packed ? max(H.y1, off1 + static_cast<int>(ext1)) : H.y1;
                H.e1 = packed ? max(H.e1, static_cast<int>(ext1)) : H.e1;
                uint32_t r = packed ? rec_pack(static_cast<int>(rel), ext0, ext1, false) : kRecInvalid;
                if (store && ok && !fits)
                    r = escape_record(kernarg_block<AbstractionArgs>(), lo[0], lo[1], lo[2], lo[3], hi[0], hi[1],
                                      hi[2], hi[3]);
                rr[j] = r;
            }
            if (live && store) store_block(rec_plane, c.PQ, q, i0, rr);
        }
    }
    const int hx0 = H.x0, hx1 = H.x1, hy0 = H.y0, hy1 = H.y1;
    const bool any_valid = H.any_valid, unrepresentable = H.unrepresentable;
    if (!kFast && A.ext1_max) {
        // The tallest compact box of a general row (pair_rows_kernel bounds the
        // fast rows'), for the satisfaction's row scans.
        uint32_t e = live ? static_cast<uint32_t>(H.e1) : 0u;
        for (int off = 32; off > 0; off >>= 1) e = max(e, static_cast<uint32_t>(__shfl_xor(e, off, 64)));
        if (lane == 0 && e != 0) atomicMax(A.ext1_max, e);
    }
    if (live && A.hull) {
        uint32_t h;
        if (!any_valid)
            h = kHullNone;
        else if (unrepresentable || hx0 < -kHullMaxOffset || hx1 > kHullMaxOffset || hy0 < -kHullMaxOffset ||
                 hy1 > kHullMaxOffset)
            h = kHullAll;
        else
            h = hull_pack(hx0, hx1, hy0, hy1);
        A.hull[static_cast<size_t>(plane) * c.PQ + q] = h;
    }
    if (A.chunk_hull) {
        // Union over the wave's 64 cells (absolute coordinates).
        uint32_t h = kHullNone;
        if (live) {
            if (!any_valid)
                h = kHullNone;
            else if (unrepresentable || hx0 < -kHullMaxOffset || hx1 > kHullMaxOffset || hy0 < -kHullMaxOffset ||
                     hy1 > kHullMaxOffset)
                h = kHullAll;
            else
                h = hull_pack(hx0, hx1, hy0, hy1);
        }
        const bool all = __ballot(h == kHullAll) != 0;
        const bool none = __ballot(h != kHullNone) == 0;
        uint32_t x0 = 0xFFFFFFFFu, x1 = 0, y0 = 0xFFFFFFFFu, y1 = 0;
        if (h != kHullNone && h != kHullAll) {
            x0 = ix + hull_byte(h, 0);
            x1 = ix + hull_byte(h, 1);
            y0 = iy + hull_byte(h, 2);
            y1 = iy + hull_byte(h, 3);
        }
        for (int off = 32; off > 0; off >>= 1) {
            x0 = min(x0, static_cast<uint32_t>(__shfl_xor(x0, off, 64)));
            x1 = max(x1, static_cast<uint32_t>(__shfl_xor(x1, off, 64)));
            y0 = min(y0, static_cast<uint32_t>(__shfl_xor(y0, off, 64)));
            y1 = max(y1, static_cast<uint32_t>(__shfl_xor(y1, off, 64)));
        }
        if (lane == 0)
            A.chunk_hull[task] = all ? uint4{kChunkAll, 0, 0, 0} : none ? uint4{kChunkNone, 0, 0, 0} : uint4{x0, x1, y0, y1};
    }
}

// Reference abstraction triplet (min_flat, max_flat, valid) in the reference's
// pair order p = state * I + input; zeros for invalid pairs (kernels.cu:189-216).
__global__ __launch_bounds__(256) void export_kernel(DevCase c, const uint32_t* __restrict__ rec,
                                                      const uint2* __restrict__ box2, const uint2* __restrict__ box3,
                                                      const OverflowEntry* __restrict__ overflow, uint64_t pair_begin,
                                                      uint64_t pair_end, uint32_t* __restrict__ out_min,
                                                      uint32_t* __restrict__ out_max, uint8_t* __restrict__ out_valid) {
    const uint64_t p = pair_begin + static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (p >= pair_end) return;
    const uint64_t o = p - pair_begin;
    const uint32_t s = static_cast<uint32_t>(p / c.I);
    const uint32_t i = static_cast<uint32_t>(p - static_cast<uint64_t>(s) * c.I);
    const uint32_t plane = s / c.PQ;
    const uint32_t q = s - plane * c.PQ;
    const uint32_t r = rec[static_cast<uint64_t>(plane) * c.I_pad * c.PQ + rec_offset(c.PQ, q, i)];
    if (r == kRecInvalid) {
        out_min[o] = 0;
        out_max[o] = 0;
        out_valid[o] = 0;
        return;
    }
    if (r & kRecEscape) {
        const OverflowEntry e = overflow[r & kRecIndexMask];
        out_min[o] = e.min_flat;
        out_max[o] = e.max_flat;
        out_valid[o] = 1;
        return;
    }
    const uint32_t iz = plane % c.shape[2];
    const uint32_t iv = plane / c.shape[2];
    const uint2 b2 = box2[iz * c.n_w + i / c.n_a];
    const uint2 b3 = box3[iv * c.n_a + i % c.n_a];
    uint32_t lo0, lo1;
    rec_lo(r, q, c.shape[0], &lo0, &lo1);
    out_min[o] = lo0 * c.stride[0] + lo1 * c.stride[1] + b2.x * c.stride[2] + b3.x * c.stride[3];
    out_max[o] = (lo0 + rec_ext0(r)) * c.stride[0] + (lo1 + rec_ext1(r)) * c.stride[1] + b2.y * c.stride[2] +
                 b3.y * c.stride[3];
    out_valid[o] = 1;
}

// ---------------------------------------------------------------------------
// Pair satisfaction by prefix counts, for the boxes the row scans do not
// cover: an escaped pair, or a compact box of more than sixteen rows or more
// than two bitmask slices along θ or v (only where the unreached prefix
// table is built).
//
// sat = !dead && count(unreached valid cells in the box) == 0, which equals the
// reference's `valid && cntV > 0 && cntR == cntV` because reachable ⊆ valid
// makes cntV - cntR the count of valid, unreached cells. An overflow entry
// settled `dead` when it was written; a compact record learns it here.
__device__ inline bool count_test(const SatisfyArgs& A, const uint32_t* PU, uint32_t r, uint32_t q, uint2 b2, uint2 b3,
                                  uint32_t* rec_slot) {
    const DevCase& c = A.c;
    const int wd = c.wrap_dim;
    if (r & kRecEscape) {
        const OverflowEntry o = A.overflow[r & kRecIndexMask];
        if (o.dead) return false;
        uint32_t lo[4], hi[4];
        unflatten(c, o.min_flat, lo);
        unflatten(c, o.max_flat, hi);
        return box_count_i64(PU, c.pstride, wd, c.shape[wd], lo, hi) == 0;
    }
    uint32_t lo[4], hi[4];
    rec_lo(r, q, c.shape[0], &lo[0], &lo[1]);
    hi[0] = lo[0] + rec_ext0(r);
    hi[1] = lo[1] + rec_ext1(r);
    lo[2] = b2.x;
    hi[2] = b2.y;
    lo[3] = b3.x;
    hi[3] = b3.y;
    if (box_count_wrap_u32(PU, c.pstride, wd, c.shape[wd], lo, hi) != 0) return false;
    // No unreached valid cell: satisfied unless the box has no valid cell at all;
    // that answer never changes, so it is memoised in the record.
    if (box_count_wrap_u32(A.prefix_valid, c.pstride, wd, c.shape[wd], lo, hi) == 0) {
        *rec_slot = r | kRecDead;
        return false;
    }
    return true;
}

// The box's θ (or v) range as a cell count, wrapping past the axis end (a
// compact record's range wraps only on the wrap axis).
__device__ __forceinline__ uint32_t range_cells(uint2 b, uint32_t n) {
    return (b.y >= b.x ? b.y - b.x : b.y + n - b.x) + 1u;
}

// Whether any of bits [row, row + 31 - sh] is set in the rows a + y * st1,
// y = 0 .. e1 (e1 < 8): the eight row loads are issued together (a short box
// re-reads its last row).
__device__ __forceinline__ bool rows_any_8(const uint32_t* __restrict__ bm, uint32_t a, uint32_t st1, uint32_t e1,
                                           uint32_t sh) {
    uint32_t w0[8], w1[8];
#pragma unroll
    for (uint32_t j = 0; j < 8; ++j) {
        const uint32_t rw = a + min(j, e1) * st1;
        w0[j] = bm[rw >> 5];
        w1[j] = bm[(rw >> 5) + 1];
    }
    bool any = false;
#pragma unroll
    for (uint32_t j = 0; j < 8; ++j) any |= bits_any_sh(w0[j], w1[j], a + min(j, e1) * st1, sh);
    return any;
}

// The same for a box of up to sixteen rows (e1 < 16): a second round of eight
// for the rows past the eighth.
__device__ __forceinline__ bool rows_any_16(const uint32_t* __restrict__ bm, uint32_t a, uint32_t st1, uint32_t e1,
                                            uint32_t sh) {
    if (rows_any_8(bm, a, st1, min(e1, 7u), sh)) return true;
    return e1 >= 8u && rows_any_8(bm, a + 8u * st1, st1, e1 - 8u, sh);
}

// The full test of one pair that the row probe could not rule out (or of an
// escaped pair). Almost every such pair turns out satisfied, and a satisfied
// pair must be shown free of unreached valid cells everywhere in its box, so
// the test scans the whole box where that is a few rows: the record, the
// probe base and the θ / v ranges are loaded together, then every x-row of
// the box (up to two slices of the probe bitmask along θ and two along v, up
// to sixteen rows each, a round of eight row loads at a time) from the
// cache-resident bitmask, and finally one row of
// the valid bitmask, which shows a valid cell in the box in almost every case.
// Other boxes take two more rows of the probe bitmask (the first and last y
// rows, where the unreached cells left in a box tend to sit) and the
// 16-corner counts (count_test). With kBits (frontier skipping on bitmasks,
// where no unreached prefix table is built) the solver guarantees that every
// box is compact and fits the scan. `counted` counts the pairs that reach a
// prefix-table count. kScan false (the sweep schedule, where most tested
// pairs are unsatisfied and the first and last rows usually show it) skips
// the whole-box scan in favour of those two rows and the count.
template <bool kBits, bool kScan = true>
__device__ __forceinline__ bool pair_satisfied(const SatisfyArgs& A, const uint32_t* PU, uint32_t* rec_plane,
                                               const uint32_t* pb_plane, uint32_t plane, uint32_t qs, uint32_t i,
                                               uint32_t& counted) {
    const DevCase& c = A.c;
    const uint32_t st1 = c.stride[1];
    uint32_t* const slot = rec_plane + rec_offset(c.PQ, qs, i);
    const uint32_t iw = i / c.n_a;
    const uint32_t rj = *slot;
    const uint32_t pbi = pb_plane[i];
    const uint2 b2 = A.box2[(plane % c.shape[2]) * c.n_w + iw];
    const uint2 b3 = A.box3[(plane / c.shape[2]) * c.n_a + (i - iw * c.n_a)];
    if (!(rj & kRecEscape)) {
        const uint32_t* const bm = reinterpret_cast<const uint32_t*>(A.unreached_bits);
        const uint32_t e1 = rec_ext1(rj), sh = rec_shift0(rj);
        const uint32_t mid = qs + static_cast<uint32_t>(rec_rel(rj));  // the middle row's first cell
        const uint32_t first = mid - static_cast<uint32_t>(__mul24(static_cast<int>(e1 >> 1), static_cast<int>(st1)));
        // The slices of the probe bitmask the box needs: with the OR bitmask
        // (pbi past its start) a slice at θ cell z, v cell w covers z, z + 1,
        // w, w + 1, so a range of two cells takes one slice and of three two
        // (the second starting one cell up, wrapping like the OR bitmask's
        // neighbours); the plain bitmask takes one slice per cell.
        const uint32_t n2 = c.shape[2], n3 = c.shape[3];
        const uint32_t nz = range_cells(b2, n2), nw = range_cells(b3, n3);
        const bool orb = pbi >= probe_or_bit0(c);
        const uint32_t sz = orb ? nz - 1u : nz, sw = orb ? nw - 1u : nw;
        if (kBits || (kScan && e1 < 16u && sz <= 2u && sw <= 2u)) {
            const uint32_t pitch = orb ? probe_or_pitch(c) : c.PQ;
            // The other slices start one θ cell up (sz = 2), one v cell up
            // (sw = 2), or both.
            const uint32_t z1 = b2.x + 1u == n2 ? 0u : b2.x + 1u, w1 = b3.x + 1u == n3 ? 0u : b3.x + 1u;
            const uint32_t p0 = b2.x + b3.x * n2;
            const uint32_t a = first + pbi;
            if (rows_any_16(bm, a, st1, e1, sh)) return false;
            if (sz == 2u && rows_any_16(bm, a + (z1 + b3.x * n2 - p0) * pitch, st1, e1, sh)) return false;
            if (sw == 2u && rows_any_16(bm, a + (b2.x + w1 * n2 - p0) * pitch, st1, e1, sh)) return false;
            if (sz == 2u && sw == 2u && rows_any_16(bm, a + (z1 + w1 * n2 - p0) * pitch, st1, e1, sh)) return false;
            // No unreached valid cell in the box: satisfied unless it holds no
            // valid cell at all. One row of the valid bitmask (the middle row at
            // the lower θ and v) almost always shows one; otherwise the count
            // decides, and a box with none is memoised dead.
            const uint32_t* const vb = reinterpret_cast<const uint32_t*>(A.valid_bits);
            const uint32_t vr = p0 * c.PQ + mid;
            if (bits_any_sh(vb[vr >> 5], vb[(vr >> 5) + 1], vr, sh)) return true;
            uint32_t lo[4], hi[4];
            rec_lo(rj, qs, c.shape[0], &lo[0], &lo[1]);
            hi[0] = lo[0] + rec_ext0(rj);
            hi[1] = lo[1] + e1;
            lo[2] = b2.x;
            hi[2] = b2.y;
            lo[3] = b3.x;
            hi[3] = b3.y;
            ++counted;
            if (box_count_wrap_u32(A.prefix_valid, c.pstride, c.wrap_dim, c.shape[c.wrap_dim], lo, hi) == 0) {
                *slot = rj | kRecDead;
                return false;
            }
            return true;
        }
        // The first and last y rows at the x-range and lower θ, v.
        const uint32_t ra = first + pbi;
        const uint32_t rb = ra + static_cast<uint32_t>(__mul24(static_cast<int>(e1), static_cast<int>(st1)));
        if (bits_any_sh(bm[ra >> 5], bm[(ra >> 5) + 1], ra, sh) || bits_any_sh(bm[rb >> 5], bm[(rb >> 5) + 1], rb, sh))
            return false;
    }
    ++counted;
    return count_test(A, PU, rj, qs, b2, b3, slot);
}

// Totals and the stop test of one iteration (kernels.cu:722-766), by one wave.
// The iteration number lives on the device (ctl->iterations + 1), so a captured
// sequence of launches replays correctly.
__device__ __forceinline__ void finish_body(IterCtl* ctl, IterRecord* records, uint32_t max_iterations) {
    const uint32_t iter = ctl->iterations + 1;
    IterCounters& slot = ctl->ring[iter % kRing];
    const uint32_t l = threadIdx.x;
    unsigned long long nr = 0, nc = 0, pt = 0, bc = 0, ts = 0;
    if (l < kCounterSlots) {
        nr = slot.new_reachable[l];
        nc = slot.new_certified[l];
        pt = slot.pairs_tested[l];
        bc = slot.boxes_counted[l];
        ts = slot.test_slots[l];
        slot.new_reachable[l] = 0;
        slot.new_certified[l] = 0;
        slot.pairs_tested[l] = 0;
        slot.boxes_counted[l] = 0;
        slot.test_slots[l] = 0;
    }
    nr = wave_sum_u64(nr);
    nc = wave_sum_u64(nc);
    pt = wave_sum_u64(pt);
    bc = wave_sum_u64(bc);
    ts = wave_sum_u64(ts);
    if (l != 0) return;
    ctl->total_reachable += nr;
    ctl->total_certified += nc;
    ctl->iterations = iter;
    IterRecord& rec = records[iter % kRing];
    rec.iteration = iter;
    rec.newly_reachable = nr;
    rec.newly_certified = nc;
    rec.total_reachable = ctl->total_reachable;
    rec.total_certified = ctl->total_certified;
    rec.pairs_tested = pt;
    rec.boxes_counted = bc;
    rec.test_slots = ts;
    uint32_t done = 0;
    if (ctl->total_certified == ctl->candidate_count)
        done = kDoneConverged;
    else if (nr == 0 && nc == 0)
        done = kDoneStalled;
    else if (iter >= max_iterations)
        done = kDoneMaxIterations;
    __atomic_store_n(&ctl->done, done, __ATOMIC_RELAXED);
}


// Whether a cell inside a box turned from unreached to reached between the
// tables of iterations k - 1 and k. The unreached set only shrinks, so the two
// counts differ exactly then. coarse_box_changed tests the box snapped outward
// to kCoarse-cell x / y tiles on the small coarse tables (cache-resident): a
// superset, so "unchanged" there is exact and "changed" may be a false alarm.
__device__ __forceinline__ bool coarse_box_changed(const SatisfyArgs& A, const uint32_t lo[4], const uint32_t hi[4],
                                                   const uint32_t* Ccur, const uint32_t* Cprev) {
    const DevCase& c = A.c;
    const int wd = c.wrap_dim;
    uint32_t clo[4] = {lo[0] / kCoarse, lo[1] / kCoarse, lo[2], lo[3]};
    uint32_t chi[4] = {(hi[0] + kCoarse) / kCoarse - 1, (hi[1] + kCoarse) / kCoarse - 1, hi[2], hi[3]};
    return box_count_wrap_u32(Cprev, c.cstride, wd, c.shape[wd], clo, chi) !=
           box_count_wrap_u32(Ccur, c.cstride, wd, c.shape[wd], clo, chi);
}

// The state's hull box: x and y from its hull bytes around (ix, iy), θ and v
// from its plane.
__device__ __forceinline__ bool hull_changed(const SatisfyArgs& A, uint32_t h, uint32_t ix, uint32_t iy,
                                             uint32_t plane, const uint32_t* Pcur, const uint32_t* Pprev) {
    if (h == kHullAll) return true;
    if (h == kHullNone) return false;
    const DevCase& c = A.c;
    uint32_t lo[4], hi[4];
    lo[0] = ix + hull_byte(h, 0);
    hi[0] = ix + hull_byte(h, 1);
    lo[1] = iy + hull_byte(h, 2);
    hi[1] = iy + hull_byte(h, 3);
    const uint2 hz = A.hull2[plane % c.shape[2]];
    const uint2 hv = A.hull3[plane / c.shape[2]];
    lo[2] = hz.x;
    hi[2] = hz.y;
    lo[3] = hv.x;
    hi[3] = hv.y;
    // The exact test straight away: a coarse test first would add a dependent
    // round of loads for every state in a changed chunk, and most states there
    // pass it anyway (measured 4-6% slower satisfaction with it).
    const int wd = c.wrap_dim;
    return box_count_wrap_u32(Pprev, c.pstride, wd, c.shape[wd], lo, hi) !=
           box_count_wrap_u32(Pcur, c.pstride, wd, c.shape[wd], lo, hi);
}

// The same test for a 64-cell chunk of a plane (uniform across the wave).
__device__ __forceinline__ bool chunk_changed(const SatisfyArgs& A, uint4 ch, uint32_t plane, const uint32_t* Ccur,
                                              const uint32_t* Cprev) {
    if (ch.x == kChunkAll) return true;
    if (ch.x == kChunkNone) return false;
    const DevCase& c = A.c;
    uint32_t lo[4], hi[4];
    lo[0] = ch.x;
    hi[0] = ch.y;
    lo[1] = ch.z;
    hi[1] = ch.w;
    const uint2 hz = A.hull2[plane % c.shape[2]];
    const uint2 hv = A.hull3[plane / c.shape[2]];
    lo[2] = hz.x;
    hi[2] = hz.y;
    lo[3] = hv.x;
    hi[3] = hv.y;
    return coarse_box_changed(A, lo, hi, Ccur, Cprev);
}

// ---------------------------------------------------------------------------
// One instance per launch.
//
// Pair satisfaction fused with the first-input reduction, for one group of WG
// cells of one (θ, v) plane, `block` = plane * groups_per_plane + group.
//
// The workgroup first compacts the states of its group that need testing into
// LDS, so each wave runs 64 of them (no lanes spent on settled or invalid
// states). Each wave then walks the inputs in increasing order in blocks of IB:
// the block's record loads are issued together, then a bitmask probe per pair
// reads the whole x-row of the successor box at its middle y (θ and v at their
// lower ends) — any valid unreached cell there proves the pair unsatisfied —
// and only the pairs the probe cannot rule out get the full prefix count. A
// lane stops at its first satisfied input and the wave when no lane is left
// (the reference breaks at the first satisfied input, reachability.hpp:106 /
// kernels.cu:369-374). With sweep_all every pair of every state is tested, the
// reference kernel's schedule (kernels.cu:668).
// States are updated in place: the prefix snapshot was built before this
// launch and only this workgroup touches these states (Jacobi, as kernels.cu:742).
// The body is satisfy_body.inc.
// Seven waves per SIMD (72 VGPRs, no spill) measured faster than eight with
// spills and than the unbounded 82 VGPRs: the pass waits on memory, so
// occupancy pays until spill traffic eats it.
// The host passes the iteration it enqueues, which is the loop state's next
// iteration unless the loop has stopped: the body loads the state and the
// group's chunk flags together and returns on a stopped loop before touching
// any state.
// kSweep: the sweep_all_pairs schedule, where each lane counts its own pairs
// in input order instead of through the count queue.
template <uint32_t WG, uint32_t kSpan, uint32_t IB, bool kSweep, bool kBits>
__global__ __launch_bounds__(WG) __attribute__((amdgpu_waves_per_eu(7))) void satisfy_kernel(SatisfyArgs A,
                                                                                              uint32_t iter) {
    const uint32_t block = blockIdx.x;
#include "satisfy_body.inc"
}

// Chunk flags of iteration `iter`: one lane per 64-cell chunk, 1 where the
// chunk's hull box (x, y from its cells' hulls, θ and v from its plane) holds a
// state whose unreached count changed since the previous iteration. The
// satisfaction workgroups read their chunks' flags first and end at once when
// none is set, without reading their states. Padding bytes keep their zeros.
// 64 bits of a flat bitmask from bit `pos` on. pos is a state index, and the
// bitmask holds one word past the last state's, so both reads are in bounds.
__device__ __forceinline__ uint64_t bits64_at(const unsigned long long* __restrict__ b, uint32_t pos) {
    const uint32_t w = pos >> 6, sh = pos & 63u;
    const uint64_t lo = b[w];
    return sh ? (lo >> sh) | (static_cast<uint64_t>(b[w + 1]) << (64u - sh)) : lo;
}

// Word j of plane `plane` of the OR bitmask: the chunk's 64 cells of the flat
// bitmask U ORed with the same cells one θ plane up, one v plane up, and both
// (the next θ (v) plane wrapping to the first when θ (v) is the wrap axis,
// absent past the last otherwise); cells past the plane's end read 0.
__device__ __forceinline__ uint64_t or_probe_word(const DevCase& c, const unsigned long long* __restrict__ U,
                                                  uint32_t plane, uint32_t j) {
    const uint32_t q0 = j * 64u;
    const uint32_t n2 = c.shape[2];
    const uint32_t z = plane % n2, w = plane / n2;
    const bool zl = z + 1u < n2, wl = w + 1u < c.shape[3];
    const bool has_z = zl || c.wrap_dim == 2, has_w = wl || c.wrap_dim == 3;
    const uint32_t pz = zl ? plane + 1u : plane - z;
    const uint32_t pw = wl ? plane + n2 : plane - w * n2;
    const uint32_t pzw = wl ? pz + n2 : pz - w * n2;
    const uint32_t PQ = c.PQ;
    uint64_t m = bits64_at(U, plane * PQ + q0);
    if (has_z) m |= bits64_at(U, pz * PQ + q0);
    if (has_w) m |= bits64_at(U, pw * PQ + q0);
    if (has_z && has_w) m |= bits64_at(U, pzw * PQ + q0);
    const uint32_t left = PQ - q0;  // cells of the plane from q0 on
    if (left < 64u) m &= (1ull << left) - 1ull;
    return m;
}

// Chunk k's flag, and (where the OR bitmask is used) its word of the OR
// bitmask: the chunk's 64 cells of valid, unreached states ORed with the same
// cells one θ plane up, one v plane up, and both — the next θ (v) plane wrapping
// to the first when θ (v) is the wrap axis, absent past the last otherwise.
// Each plane's words start on a word boundary, so a chunk is exactly a word.
__device__ __forceinline__ void chunk_flags_body(const SatisfyArgs& A, uint8_t* __restrict__ flags, uint32_t iter,
                                                 uint32_t k) {
    const DevCase& c = A.c;
    if (k >= c.n_tasks) return;
    const uint32_t plane = k / c.chunks_per_plane;
    const uint32_t j = k - plane * c.chunks_per_plane;
    const uint32_t* const CU = A.coarse_unreached[iter & 1u];
    const uint32_t* const Cprev = iter == 1 ? A.coarse_valid : A.coarse_unreached[(iter - 1u) & 1u];
    flags[plane * chunk_flag_pitch(c) + j] = chunk_changed(A, A.chunk_hull[k], plane, CU, Cprev) ? 1u : 0u;
    if (A.probe_or_bits) const_cast<unsigned long long*>(A.probe_or_bits)[k] = or_probe_word(c, A.unreached_bits, plane, j);
}

__global__ __launch_bounds__(256) void chunk_flags_kernel(SatisfyArgs A, uint8_t* __restrict__ flags,
                                                          uint32_t iter) {
    if (loop_done(A.ctl)) return;
    chunk_flags_body(A, flags, iter, blockIdx.x * 256u + threadIdx.x);
}

// ---------------------------------------------------------------------------
// Frontier skipping from bitmasks. A state tested and not satisfied at
// iteration k - 1 gets the same answer at k unless a cell of its hull turned
// from unreached to reached in between (see satisfy_kernel). Here that test is
// made for all states of a (θ, v) plane at once, on bitmasks:
//   N(p')  = prev & ~cur over plane p': the cells reached since the previous
//            iteration (before the first: every valid cell that is not
//            unreached, i.e. the seeds);
//   D(p)   = OR of N(p') over the planes p' of p's θ and v hull ranges
//            (hull2, hull3, the unions over the inputs of the boxes' ranges);
//   E(p)   = D(p) dilated by the plane's hull rectangle: bit (x, y) is set when
//            D(p) has a bit at (x + dx, y + dy) for some offsets dx, dy inside
//            the union of the plane's state hulls.
// Every box of a state lies inside its hull, so a state whose E bit is clear
// saw no change in any of its boxes and is skipped; E is a superset of the
// per-state hull test, exact for every state whose hull is the plane's union
// (all interior cells of a plane whose rows are fast rows). The satisfaction
// reads E as one word per 64-cell chunk: a chunk whose word is 0 has nothing
// to test. Set bits are kept only for states with a valid pair (keep), and
// states whose hull is not representable are always tested (always).

// Once per abstraction, one workgroup per plane: the keep and always words of
// the plane's chunks, and the union of its representable state hulls
// (x0 > x1 when it has none).
__global__ __launch_bounds__(256) void frontier_setup_kernel(DevCase c, const uint32_t* __restrict__ hull,
                                                             unsigned long long* __restrict__ keep,
                                                             unsigned long long* __restrict__ always,
                                                             int4* __restrict__ plane_rect) {
    __shared__ int s_rect[4][4];
    const uint32_t plane = blockIdx.x;
    const uint32_t lane = lane_id(), wave = threadIdx.x >> 6;
    const uint32_t cpp = c.chunks_per_plane;
    int x0 = kHullMaxOffset + 1, x1 = -kHullMaxOffset - 1, y0 = kHullMaxOffset + 1, y1 = -kHullMaxOffset - 1;
    for (uint32_t j = wave; j < cpp; j += 4u) {
        const uint32_t q = j * 64u + lane;
        const uint32_t h = q < c.PQ ? hull[static_cast<size_t>(plane) * c.PQ + q] : kHullNone;
        const uint64_t kb = __ballot(h != kHullNone), ab = __ballot(h == kHullAll);
        if (lane == 0) {
            keep[static_cast<size_t>(plane) * cpp + j] = kb;
            always[static_cast<size_t>(plane) * cpp + j] = ab;
        }
        if (h != kHullNone && h != kHullAll) {
            x0 = min(x0, hull_byte(h, 0));
            x1 = max(x1, hull_byte(h, 1));
            y0 = min(y0, hull_byte(h, 2));
            y1 = max(y1, hull_byte(h, 3));
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        x0 = min(x0, __shfl_xor(x0, off, 64));
        x1 = max(x1, __shfl_xor(x1, off, 64));
        y0 = min(y0, __shfl_xor(y0, off, 64));
        y1 = max(y1, __shfl_xor(y1, off, 64));
    }
    if (lane == 0) {
        s_rect[wave][0] = x0;
        s_rect[wave][1] = x1;
        s_rect[wave][2] = y0;
        s_rect[wave][3] = y1;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < 4; ++w) {
            s_rect[0][0] = min(s_rect[0][0], s_rect[w][0]);
            s_rect[0][1] = max(s_rect[0][1], s_rect[w][1]);
            s_rect[0][2] = min(s_rect[0][2], s_rect[w][2]);
            s_rect[0][3] = max(s_rect[0][3], s_rect[w][3]);
        }
        plane_rect[plane] = make_int4(s_rect[0][0], s_rect[0][1], s_rect[0][2], s_rect[0][3]);
    }
}

// Eight states per thread: their valid and reach bytes as one 8-byte load
// each, the previous unreached byte beside them, and one byte of each bitmask
// written (bit s of a bitmask is bit s % 8 of its byte s / 8).
__device__ __forceinline__ void frontier_bits_body(const FrontierArgs& F, uint32_t iter, uint32_t t) {
    const uint32_t S = F.c.S;
    const uint32_t s0 = t * 8u;
    if (s0 >= S) return;
    uint8_t* const cb = reinterpret_cast<uint8_t*>(F.cur);
    const uint8_t old = iter == 1 ? reinterpret_cast<const uint8_t*>(F.valid_bits)[t] : cb[t];
    uint32_t b = 0;
    if (s0 + 8u <= S) {
        const uint2 v = *reinterpret_cast<const uint2*>(F.valid + s0);
        const uint2 r = *reinterpret_cast<const uint2*>(F.reach + s0);
#pragma unroll
        for (uint32_t j = 0; j < 4; ++j) {
            b |= static_cast<uint32_t>(((v.x >> (8 * j)) & 0xFFu) != 0 && ((r.x >> (8 * j)) & 0xFFu) == 0) << j;
            b |= static_cast<uint32_t>(((v.y >> (8 * j)) & 0xFFu) != 0 && ((r.y >> (8 * j)) & 0xFFu) == 0) << (j + 4);
        }
    } else {
        for (uint32_t j = 0; s0 + j < S; ++j) b |= static_cast<uint32_t>(F.valid[s0 + j] != 0 && F.reach[s0 + j] == 0) << j;
    }
    reinterpret_cast<uint8_t*>(F.fresh)[t] = static_cast<uint8_t>(old & ~b);
    cb[t] = static_cast<uint8_t>(b);
}

__global__ __launch_bounds__(256) void frontier_bits_kernel(FrontierArgs F, uint32_t iter) {
    if (loop_done(F.ctl)) return;
    frontier_bits_body(F, iter, blockIdx.x * 256u + threadIdx.x);
}

// 32 bits of a flat bitmask from bit `pos` on (the bitmask holds one word past
// its last bit, so both reads are in bounds for pos below the bit count).
__device__ __forceinline__ uint32_t bits32_at(const unsigned long long* __restrict__ b, uint32_t pos) {
    const uint32_t* const w = reinterpret_cast<const uint32_t*>(b);
    return __builtin_amdgcn_alignbit(w[(pos >> 5) + 1], w[pos >> 5], pos & 31u);
}

// The v pass of D: per plane (θ cell z, v cell w), row y, word k of the
// row-aligned layout, the states reached since the previous iteration ORed
// over the planes (z, v') of w's v hull range. One thread per word.
__device__ __forceinline__ void frontier_v_body(const FrontierArgs& F, uint32_t t) {
    const DevCase& c = F.c;
    const uint32_t n0 = c.shape[0], n1 = c.shape[1], n2 = c.shape[2], n3 = c.shape[3];
    const uint32_t rw = F.row_words, nwords = n1 * rw;
    if (t >= nwords * n2 * n3) return;
    const uint32_t plane = t / nwords, r = t - plane * nwords;
    const uint32_t y = r / rw, k = r - y * rw;
    const uint32_t z = plane % n2, w = plane / n2;
    const uint2 hv = F.hull3[w];
    const uint32_t nv = hv.y >= hv.x ? hv.y - hv.x + 1u : hv.y + n3 - hv.x + 1u;
    const uint32_t x0 = 32u * k, len = min(32u, n0 - x0);
    const uint32_t mask = len == 32u ? 0xFFFFFFFFu : (1u << len) - 1u;
    const uint32_t cell = y * n0 + x0;
    uint32_t acc = 0;
    for (uint32_t tv = 0; tv < nv; tv += 4u) {
        uint32_t v[4];
#pragma unroll
        for (uint32_t u = 0; u < 4; ++u) {
            const uint32_t pv = hv.x + tv + u >= n3 ? hv.x + tv + u - n3 : hv.x + tv + u;
            v[u] = tv + u < nv ? bits32_at(F.fresh, (z + pv * n2) * c.PQ + cell) : 0u;
        }
        acc |= v[0] | v[1] | v[2] | v[3];
    }
    F.fresh_v[t] = acc & mask;
}

__global__ __launch_bounds__(256) void frontier_v_kernel(FrontierArgs F) {
    if (loop_done(F.ctl)) return;
    frontier_v_body(F, blockIdx.x * 256u + threadIdx.x);
}

// 32 bits of LDS row `row` (rw words, bits past the row's end 0) from the
// signed bit offset s on; bits outside the row read 0.
__device__ __forceinline__ uint32_t row_bits_at(const uint32_t* row, int32_t rw, int32_t s) {
    const int32_t a = s >> 5;  // floor(s / 32)
    const uint32_t lo = (a >= 0 && a < rw) ? row[a] : 0u;
    const uint32_t hi = (a + 1 >= 0 && a + 1 < rw) ? row[a + 1] : 0u;
    return __builtin_amdgcn_alignbit(hi, lo, static_cast<uint32_t>(s) & 31u);
}

// One workgroup per plane: D (the v pass's rows ORed over the θ hull range),
// then E (in LDS, row-aligned: y * rw + x / 32), then the plane's frontier
// words and its words of the OR probe bitmask.
__device__ __forceinline__ void frontier_plane_body(const FrontierArgs& F, uint32_t plane, uint32_t* s_rows) {
    const DevCase& c = F.c;
    const uint32_t n0 = c.shape[0], n1 = c.shape[1], n2 = c.shape[2];
    const uint32_t rw = F.row_words, nwords = n1 * rw;
    uint32_t* const A = s_rows;
    uint32_t* const B = s_rows + nwords;
    const int4 rect = F.plane_rect[plane];
    const bool any_rect = rect.x <= rect.y;  // uniform
    if (any_rect) {
        const uint32_t z = plane % n2, w = plane / n2;
        const uint2 hz = F.hull2[z];
        const uint32_t nz = hz.y >= hz.x ? hz.y - hz.x + 1u : hz.y + n2 - hz.x + 1u;
        for (uint32_t t = threadIdx.x; t < nwords; t += blockDim.x) {
            uint32_t acc = 0;
            for (uint32_t tz = 0; tz < nz; tz += 4u) {
                uint32_t v[4];
#pragma unroll
                for (uint32_t u = 0; u < 4; ++u) {
                    const uint32_t pz = hz.x + tz + u >= n2 ? hz.x + tz + u - n2 : hz.x + tz + u;
                    v[u] = tz + u < nz ? F.fresh_v[static_cast<size_t>(pz + w * n2) * nwords + t] : 0u;
                }
                acc |= v[0] | v[1] | v[2] | v[3];
            }
            A[t] = acc;
        }
        __syncthreads();
        // x: bit x of a row = OR of the row's bits x + dx, dx in [rect.x, rect.y]
        // (four offsets per step, their LDS reads issued together).
        for (uint32_t t = threadIdx.x; t < nwords; t += blockDim.x) {
            const uint32_t y = t / rw, k = t - y * rw;
            const uint32_t len = min(32u, n0 - 32u * k);
            const uint32_t mask = len == 32u ? 0xFFFFFFFFu : (1u << len) - 1u;
            const uint32_t* const row = A + y * rw;
            const int32_t s0 = static_cast<int32_t>(32u * k);
            uint32_t acc = 0;
            for (int32_t dx = rect.x; dx <= rect.y; dx += 4) {
                uint32_t v[4];
#pragma unroll
                for (int32_t u = 0; u < 4; ++u)
                    v[u] = dx + u <= rect.y ? row_bits_at(row, static_cast<int32_t>(rw), s0 + dx + u) : 0u;
                acc |= v[0] | v[1] | v[2] | v[3];
            }
            B[t] = acc & mask;
        }
        __syncthreads();
        // y: row y = OR of rows y + dy, dy in [rect.z, rect.w] (back into A).
        for (uint32_t t = threadIdx.x; t < nwords; t += blockDim.x) {
            const uint32_t y = t / rw, k = t - y * rw;
            uint32_t acc = 0;
            for (int32_t dy = rect.z; dy <= rect.w; dy += 4) {
                uint32_t v[4];
#pragma unroll
                for (int32_t u = 0; u < 4; ++u) {
                    const int32_t yy = static_cast<int32_t>(y) + dy + u;
                    v[u] = (dy + u <= rect.w && yy >= 0 && yy < static_cast<int32_t>(n1)) ? B[yy * rw + k] : 0u;
                }
                acc |= v[0] | v[1] | v[2] | v[3];
            }
            A[t] = acc;
        }
        __syncthreads();
    }
    // The plane's chunk words, a lane per chunk: E's bits in flat cell order
    // (row segments of at most 32 bits), kept for the states with a valid
    // pair, with the always-tested states added; and the OR probe word.
    const uint32_t cpp = c.chunks_per_plane;
    for (uint32_t j = threadIdx.x; j < cpp; j += blockDim.x) {
        const size_t k = static_cast<size_t>(plane) * cpp + j;
        const uint64_t keep = F.keep[k], always = F.always[k];
        const uint64_t orw = F.or_bits ? or_probe_word(c, F.cur, plane, j) : 0ull;
        uint64_t e = 0;
        if (any_rect) {
            uint32_t q = j * 64u, filled = 0;
            const uint32_t qe = min(q + 64u, c.PQ);
            uint32_t y = q / n0, x = q - y * n0;
            while (q < qe) {
                const uint32_t take = min(min(32u, n0 - x), qe - q);
                const uint32_t bits = row_bits_at(A + y * rw, static_cast<int32_t>(rw), static_cast<int32_t>(x));
                const uint32_t m = take == 32u ? bits : bits & ((1u << take) - 1u);
                e |= static_cast<uint64_t>(m) << filled;
                filled += take;
                q += take;
                x += take;
                if (x == n0) {
                    x = 0;
                    ++y;
                }
            }
        }
        F.frontier[k] = (e & keep) | always;
        if (F.or_bits) F.or_bits[k] = orw;
    }
}

__global__ __launch_bounds__(kFrontierThreads) void frontier_plane_kernel(FrontierArgs F) {
    if (loop_done(F.ctl)) return;
    extern __shared__ uint32_t s_rows[];
    frontier_plane_body(F, blockIdx.x, s_rows);
}

hipError_t launch_frontier_plane(const FrontierArgs& F, hipStream_t stream) {
    const uint64_t words = static_cast<uint64_t>(F.c.shape[1]) * F.row_words * F.c.shape[2] * F.c.shape[3];
    hipLaunchKernelGGL(frontier_v_kernel, dim3(static_cast<uint32_t>((words + 255u) / 256u)), dim3(256), 0, stream, F);
    hipLaunchKernelGGL(frontier_plane_kernel, dim3(F.c.shape[2] * F.c.shape[3]), dim3(kFrontierThreads),
                       frontier_lds_bytes(F.c), stream, F);
    return hipGetLastError();
}

// Pair satisfaction + first-input reduction with the successor boxes recomputed
// in FP64 (pair_box, the abstraction kernel's arithmetic) instead of read from
// stored records: no per-pair HBM at all, at the cost of the K trajectory steps
// per tested pair. Same schedule and results as satisfy_kernel; one input at a
// time per lane, lanes leave at their first satisfied input.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8))) void satisfy_recompute_kernel(RecomputeArgs R) {
    const SatisfyArgs& A = R.sat;
    const DevCase& c = A.c;
    constexpr uint32_t kUse = 1, kCand = 2, kReached = 4, kCertified = 8;
    __shared__ uint32_t s_q[256];
    __shared__ uint8_t s_flags[256];
    __shared__ uint32_t s_wave_n[4];
    const uint32_t lane = lane_id();
    const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const uint32_t PQ = c.PQ;
    const uint32_t gpp = (PQ + 255u) / 256u;
    const uint32_t plane = blockIdx.x / gpp;
    const uint32_t g0 = (blockIdx.x - plane * gpp) * 256u;
    const bool skip_unchanged = !A.sweep_all && c.hull_enabled;
    // The loop state and the group's four chunk flags (chunk_flags_kernel, one
    // aligned word) together; a group with no changed chunk ends here.
    const uint64_t st = loop_state(A.ctl);
    uint32_t gf = 0xFFFFFFFFu;
    if (skip_unchanged) gf = *reinterpret_cast<const uint32_t*>(A.chunk_flags + plane * chunk_flag_pitch(c) + (g0 >> 6));
    if (state_done(st) || gf == 0) return;
    const uint32_t iter = state_next_iteration(st);
    const uint32_t* const PU = A.prefix_unreached[iter & 1u];
    // The first iteration compares with "nothing reached yet" (the valid table).
    const uint32_t* const Pprev = iter == 1 ? A.prefix_valid : A.prefix_unreached[(iter - 1u) & 1u];
    const bool chunk_live = ((gf >> (8u * wave)) & 0xFFu) != 0;
    {
        const uint32_t q = g0 + threadIdx.x;
        const uint32_t s = plane * PQ + q;
        uint32_t f = 0;
        bool eval = false;
        if (q < PQ) {
            if (A.valid[s]) {
                const bool is_cand = A.cand[s] != 0;
                const bool reached = A.reach[s] != 0;
                const bool certified = A.ctrl[s] != GSC_INVALID_INPUT;
                const bool use = !((!is_cand && reached) || (is_cand && certified));
                f = (use ? kUse : 0) | (is_cand ? kCand : 0) | (reached ? kReached : 0) | (certified ? kCertified : 0);
            }
            eval = A.sweep_all ? true : (f & kUse) != 0;
            if (eval && skip_unchanged) {
                const uint32_t iy0 = q / c.shape[0];
                eval = chunk_live && hull_changed(A, A.hull[s], q - iy0 * c.shape[0], iy0, plane, PU, Pprev);
            }
        }
        const uint64_t m = __ballot(eval);
        if (lane == 0) s_wave_n[wave] = static_cast<uint32_t>(__popcll(m));
        __syncthreads();
        uint32_t base = 0;
        for (uint32_t w = 0; w < wave; ++w) base += s_wave_n[w];
        if (eval) {
            const uint32_t pos = base + lanes_below(m);
            s_q[pos] = q;
            s_flags[pos] = static_cast<uint8_t>(f);
        }
        __syncthreads();
    }
    const uint32_t n_eval = s_wave_n[0] + s_wave_n[1] + s_wave_n[2] + s_wave_n[3];
    uint32_t nr = 0, nc = 0, tested = 0, counted = 0, slots = 0;
    const uint32_t e = wave * 64u + lane;
    const bool has = e < n_eval;
    if (wave * 64u < n_eval) {
        const uint32_t q = has ? s_q[e] : g0;
        const uint32_t flags = has ? s_flags[e] : 0u;
        const uint32_t s = plane * PQ + q;
        const uint32_t iy = q / c.shape[0];
        const uint32_t ix = q - iy * c.shape[0];
        const uint32_t iz = plane % c.shape[2];
        const uint32_t iv = plane / c.shape[2];
        const double x0 = cell_center(c.lb[0], ix, c.eta[0]);
        const double y0 = cell_center(c.lb[1], iy, c.eta[1]);
        const double rad0 = R.rad[2 * iv];
        const double rad1 = R.rad[2 * iv + 1];
        const bool state_ok = state_constraint_ok(c, R.fail2d, ix, iy);
        const uint32_t* const ubits = reinterpret_cast<const uint32_t*>(A.unreached_bits);
        const int wd = c.wrap_dim;
        uint32_t found = GSC_INVALID_INPUT;
        bool pending = has;
        uint32_t i = 0;
        // Wave-uniform per input: scalar loads.
        const ConstPtr<uint32_t> box2 = to_const(reinterpret_cast<const uint32_t*>(A.box2 + iz * c.n_w));
        const ConstPtr<uint32_t> box3 = to_const(reinterpret_cast<const uint32_t*>(A.box3 + iv * c.n_a));
        for (uint32_t iw = 0; iw < c.n_w; ++iw) {
            const uint2 b2 = make_uint2(box2[2 * iw], box2[2 * iw + 1]);
            for (uint32_t ia = 0; ia < c.n_a; ++ia, ++i) {
                if (__ballot(pending) == 0) break;
                if (lane == 0) slots += 64u;
                if (!pending) continue;
                ++tested;
                const uint2 b3 = make_uint2(box3[2 * ia], box3[2 * ia + 1]);
                const double2* const inc = R.inc + (static_cast<size_t>(plane) * c.I + i) * c.K;
                const double* const sums = R.inc_sums + (static_cast<size_t>(plane) * c.I + i) * 4;
                const ConstPtr<int32_t> offs_p = to_const(R.cell_offs + (static_cast<size_t>(plane) * c.I + i) * 8);
                int32_t offs[8];
#pragma unroll
                for (int k = 0; k < 8; ++k) offs[k] = offs_p[k];
                uint32_t lo[4], hi[4];
                if (!state_ok || !pair_box(c, &kernarg_block<RecomputeArgs>()->sat.c, inc, sums, offs, b2, b3, ix, iy,
                                           x0, y0, rad0, rad1, R.fail2d, lo, hi))
                    continue;
                const bool regular = lo[0] <= hi[0] && lo[1] <= hi[1] && (lo[2] <= hi[2] || wd == 2) &&
                                     (lo[3] <= hi[3] || wd == 3);
                bool sat;
                if (regular) {
                    // Row probe as in satisfy_kernel (x-range at middle y, lower θ and v).
                    if (hi[0] - lo[0] <= kExtMax) {
                        const uint32_t row = lo[0] + (lo[1] + ((hi[1] - lo[1]) >> 1)) * c.stride[1] +
                                             lo[2] * c.stride[2] + lo[3] * c.stride[3];
                        if (bits_any(ubits[row >> 5], ubits[(row >> 5) + 1], row, hi[0] - lo[0])) continue;
                    }
                    ++counted;
                    sat = box_count_wrap_u32(PU, c.pstride, wd, c.shape[wd], lo, hi) == 0 &&
                          box_count_wrap_u32(A.prefix_valid, c.pstride, wd, c.shape[wd], lo, hi) != 0;
                } else {
                    ++counted;
                    sat = box_count_i64(A.prefix_valid, c.pstride, wd, c.shape[wd], lo, hi) != 0 &&
                          box_count_i64(PU, c.pstride, wd, c.shape[wd], lo, hi) == 0;
                }
                if (sat) {
                    if (found == GSC_INVALID_INPUT) found = i;
                    if (!A.sweep_all) pending = false;
                }
            }
        }
        if ((flags & kUse) && found != GSC_INVALID_INPUT) {
            if (!(flags & kReached)) {
                A.reach[s] = 1;
                nr = 1;
            }
            if (!(flags & kCertified)) {
                A.ctrl[s] = found;
                if (flags & kCand) {
                    nc = 1;
                } else if (A.rstep[s] == GSC_UNREACHABLE_STEP) {
                    A.rstep[s] = iter;
                }
            }
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        nr += __shfl_xor(nr, off, 64);
        nc += __shfl_xor(nc, off, 64);
        tested += __shfl_xor(tested, off, 64);
        counted += __shfl_xor(counted, off, 64);
    }
    // Each wave adds its own counts: no closing barrier, so waves left without
    // states after the compaction end at once (as in satisfy_kernel).
    if (lane == 0 && (nr | nc | tested | counted | slots) != 0) {
        IterCounters& ring = A.ctl->ring[iter % kRing];
        const uint32_t k = (blockIdx.x * 4u + wave) % kCounterSlots;
        if (nr) atomicAdd(&ring.new_reachable[k], static_cast<unsigned long long>(nr));
        if (nc) atomicAdd(&ring.new_certified[k], static_cast<unsigned long long>(nc));
        if (tested) atomicAdd(&ring.pairs_tested[k], static_cast<unsigned long long>(tested));
        if (counted) atomicAdd(&ring.boxes_counted[k], static_cast<unsigned long long>(counted));
        if (slots) atomicAdd(&ring.test_slots[k], static_cast<unsigned long long>(slots));
    }
}

__global__ void finish_kernel(IterCtl* ctl, IterRecord* records, uint32_t max_iterations) {
    if (blockIdx.x != 0 || threadIdx.x >= 64) return;
    if (loop_done(ctl)) return;
    finish_body(ctl, records, max_iterations);
}

// ---------------------------------------------------------------------------
// Many instances per launch. Every stage of one iteration is a single launch
// over the concatenated work of all instances: starts[k] .. starts[k + 1] - 1
// are instance k's blocks (starts[n] = grid size), so the block -> instance
// lookup is a binary search over n + 1 words, uniform per workgroup.
__device__ __forceinline__ uint32_t find_instance(const uint32_t* __restrict__ starts, uint32_t n, uint32_t b) {
    uint32_t lo = 0, hi = n;  // starts[lo] <= b < starts[hi]
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (__builtin_amdgcn_readfirstlane(starts[mid]) <= b)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

__global__ __launch_bounds__(256) void prefix_rows_batch_kernel(const BatchInstance* __restrict__ B,
                                                                 const uint32_t* __restrict__ starts, uint32_t n) {
    const uint32_t k = find_instance(starts, n, blockIdx.x);
    const BatchInstance& I = B[k];
    const uint64_t st = loop_state(I.sat.ctl);
    if (state_done(st)) return;
    prefix_rows_body(I.sat.c, I.sat.valid, I.sat.reach, I.prefix_unreached[state_next_iteration(st) & 1u],
                     blockIdx.x - starts[k]);
}

__global__ __launch_bounds__(256) void prefix_axis_batch_kernel(const BatchInstance* __restrict__ B,
                                                                 const uint32_t* __restrict__ starts, uint32_t n,
                                                                 int axis) {
    const uint32_t k = find_instance(starts, n, blockIdx.x);
    const BatchInstance& I = B[k];
    const uint64_t st = loop_state(I.sat.ctl);
    if (state_done(st)) return;
    prefix_axis_body(I.sat.c, axis, I.prefix_unreached[state_next_iteration(st) & 1u], blockIdx.x - starts[k]);
}

__global__ __launch_bounds__(256) void unreached_bits_batch_kernel(const BatchInstance* __restrict__ B,
                                                                    const uint32_t* __restrict__ starts, uint32_t n) {
    const uint32_t k = find_instance(starts, n, blockIdx.x);
    const BatchInstance& I = B[k];
    if (loop_done(I.sat.ctl)) return;
    unreached_bits_body(I.sat.c.S, I.sat.valid, I.sat.reach, I.unreached_bits, blockIdx.x - starts[k]);
}

__global__ __launch_bounds__(256) void coarse_batch_kernel(const BatchInstance* __restrict__ B,
                                                            const uint32_t* __restrict__ starts, uint32_t n) {
    const uint32_t k = find_instance(starts, n, blockIdx.x);
    const BatchInstance& I = B[k];
    const uint64_t st = loop_state(I.sat.ctl);
    if (state_done(st)) return;
    const uint32_t par = state_next_iteration(st) & 1u;
    coarse_body(I.sat.c, I.prefix_unreached[par], I.coarse_unreached[par],
                (blockIdx.x - starts[k]) * 256u + threadIdx.x);
}

// The fused prefix build of every instance whose planes fit in LDS (the
// batch uses it only when all do): a workgroup per plane, the case's pitch in
// BatchInstance::plane_pitch, dynamic LDS sized for the largest plane.
__global__ __launch_bounds__(kPlaneThreads) void prefix_plane_batch_kernel(const BatchInstance* __restrict__ B,
                                                                            const uint32_t* __restrict__ starts,
                                                                            uint32_t n) {
    const uint32_t k = find_instance(starts, n, blockIdx.x);
    const BatchInstance& I = B[k];
    const uint64_t st = loop_state(I.sat.ctl);
    if (state_done(st)) return;
    prefix_plane_body(I.sat.c, I.sat.valid, I.sat.reach, I.prefix_unreached[state_next_iteration(st) & 1u],
                      I.unreached_bits, I.plane_pitch, blockIdx.x - starts[k]);
}

__global__ __launch_bounds__(256) void prefix_last_axis_batch_kernel(const BatchInstance* __restrict__ B,
                                                                      const uint32_t* __restrict__ starts, uint32_t n) {
    const uint32_t k = find_instance(starts, n, blockIdx.x);
    const BatchInstance& I = B[k];
    const uint64_t st = loop_state(I.sat.ctl);
    if (state_done(st)) return;
    const uint32_t par = state_next_iteration(st) & 1u;
    prefix_last_axis_body(I.sat.c, I.prefix_unreached[par], I.coarse_unreached[par],
                          (blockIdx.x - starts[k]) * 256u + threadIdx.x);
}

// Chunk flags of every instance (chunk_flags_kernel), into each case's own table.
__global__ __launch_bounds__(256) void chunk_flags_batch_kernel(const BatchInstance* __restrict__ B,
                                                                const uint32_t* __restrict__ starts, uint32_t n) {
    const uint32_t k = find_instance(starts, n, blockIdx.x);
    const BatchInstance& I = B[k];
    const uint64_t st = loop_state(I.sat.ctl);
    if (state_done(st)) return;
    if (I.sat.sweep_all || !I.sat.c.hull_enabled) return;  // no frontier skipping: the flags are not read
    chunk_flags_body(I.sat, const_cast<uint8_t*>(I.sat.chunk_flags), state_next_iteration(st),
                     (blockIdx.x - starts[k]) * 256u + threadIdx.x);
}

// The bitmask frontier of every instance (frontier_bits_kernel,
// frontier_v_kernel, frontier_plane_kernel), each case's arguments in
// BatchInstance::fr; the plane stage's LDS is sized for the largest case.
__global__ __launch_bounds__(256) void frontier_bits_batch_kernel(const BatchInstance* __restrict__ B,
                                                                  const uint32_t* __restrict__ starts, uint32_t n) {
    const uint32_t k = find_instance(starts, n, blockIdx.x);
    const BatchInstance& I = B[k];
    const uint64_t st = loop_state(I.sat.ctl);
    if (state_done(st)) return;
    frontier_bits_body(I.fr, state_next_iteration(st), (blockIdx.x - starts[k]) * 256u + threadIdx.x);
}

__global__ __launch_bounds__(256) void frontier_v_batch_kernel(const BatchInstance* __restrict__ B,
                                                               const uint32_t* __restrict__ starts, uint32_t n) {
    const uint32_t k = find_instance(starts, n, blockIdx.x);
    const BatchInstance& I = B[k];
    if (loop_done(I.sat.ctl)) return;
    frontier_v_body(I.fr, (blockIdx.x - starts[k]) * 256u + threadIdx.x);
}

__global__ __launch_bounds__(kFrontierThreads) void frontier_plane_batch_kernel(const BatchInstance* __restrict__ B,
                                                                                const uint32_t* __restrict__ starts,
                                                                                uint32_t n) {
    const uint32_t k = find_instance(starts, n, blockIdx.x);
    const BatchInstance& I = B[k];
    if (loop_done(I.sat.ctl)) return;
    extern __shared__ uint32_t s_rows[];
    frontier_plane_body(I.fr, blockIdx.x - starts[k], s_rows);
}

// The batched kernel reads its case's arguments from global memory; there the
// function boundary gives the better register allocation (74 vs 97 VGPRs).
// kBits: the batch runs the bitmask frontier (every case qualifies), else the
// prefix-table stages and chunk flags.
template <uint32_t WG, uint32_t kSpan, uint32_t IB, bool kBits>
__device__ __forceinline__ void satisfy_in_batch(const SatisfyArgs& A, uint32_t block, uint32_t iter) {
    constexpr bool kSweep = false;  // the count queue in both schedules
#include "satisfy_body.inc"
}

template <uint32_t WG, uint32_t kSpan, uint32_t IB, bool kBits>
__global__ __launch_bounds__(WG) __attribute__((amdgpu_waves_per_eu(7))) void satisfy_batch_kernel(const BatchInstance* __restrict__ B,
                                                           const uint32_t* __restrict__ starts, uint32_t n) {
    const uint32_t inst = find_instance(starts, n, blockIdx.x);
    const SatisfyArgs& A = B[inst].sat;
    const uint64_t st = loop_state(A.ctl);
    if (state_done(st)) return;
    satisfy_in_batch<WG, kSpan, IB, kBits>(A, blockIdx.x - starts[inst], state_next_iteration(st));
}

__global__ void finish_batch_kernel(const BatchInstance* __restrict__ B, uint32_t n, BatchStatus* status) {
    const uint32_t k = blockIdx.x;
    if (k >= n || threadIdx.x >= 64) return;
    IterCtl* ctl = B[k].sat.ctl;
    if (!loop_done(ctl)) finish_body(ctl, B[k].records, B[k].max_iterations);
    if (threadIdx.x == 0) {
        BatchStatus st;
        st.done = __atomic_load_n(&ctl->done, __ATOMIC_RELAXED);
        st.iterations = ctl->iterations;
        st.total_reachable = ctl->total_reachable;
        st.total_certified = ctl->total_certified;
        status[k] = st;
    }
}

// Seeds and control reset of every instance (seed_kernel + reset_ctl_kernel).
__global__ __launch_bounds__(256) void seed_batch_kernel(const BatchInstance* __restrict__ B,
                                                          const uint32_t* __restrict__ starts, uint32_t n) {
    const uint32_t k = find_instance(starts, n, blockIdx.x);
    const BatchInstance& I = B[k];
    const uint32_t local = blockIdx.x - starts[k];
    if (local == 0) {
        reset_ctl_body(I.sat.ctl, I.candidate_count);
        // max_iterations = 0 runs no iteration at all (reachability.hpp:53).
        if (threadIdx.x == 0 && I.max_iterations == 0) I.sat.ctl->done = kDoneMaxIterations;
    }
    const uint32_t s = local * 256u + threadIdx.x;
    if (s >= I.sat.c.S) return;
    const uint8_t cd = I.sat.cand[s];
    I.sat.reach[s] = cd;
    I.sat.ctrl[s] = GSC_INVALID_INPUT;
    I.sat.rstep[s] = cd ? 0u : GSC_UNREACHABLE_STEP;
}

// ---------------------------------------------------------------------------
// Host launchers for the tile variants.
namespace {
using SatSingle = void (*)(SatisfyArgs, uint32_t);
using SatBatch = void (*)(const BatchInstance*, const uint32_t*, uint32_t);

template <uint32_t WG, uint32_t kSpan>
SatSingle pick_single(uint32_t ib, bool sweep, bool bits) {
    if (sweep) return ib == 4 ? satisfy_kernel<WG, kSpan, 4, true, false> : satisfy_kernel<WG, kSpan, 8, true, false>;
    if (bits) return ib == 4 ? satisfy_kernel<WG, kSpan, 4, false, true> : satisfy_kernel<WG, kSpan, 8, false, true>;
    return ib == 4 ? satisfy_kernel<WG, kSpan, 4, false, false> : satisfy_kernel<WG, kSpan, 8, false, false>;
}
template <uint32_t WG, uint32_t kSpan>
SatBatch pick_batch(uint32_t ib, bool bits) {
    if (bits) return ib == 4 ? satisfy_batch_kernel<WG, kSpan, 4, true> : satisfy_batch_kernel<WG, kSpan, 8, true>;
    return ib == 4 ? satisfy_batch_kernel<WG, kSpan, 4, false> : satisfy_batch_kernel<WG, kSpan, 8, false>;
}
}  // namespace

bool satisfy_variant_supported(uint32_t cells, uint32_t ib) {
    return (cells == 128 || cells == 256 || cells == 512 || cells == 1024) && (ib == 4 || ib == 8);
}

uint32_t satisfy_threads(uint32_t cells) { return cells < 256u ? cells : 256u; }

uint32_t satisfy_blocks(const DevCase& c, uint32_t wg) {
    return (c.PQ + wg - 1) / wg * c.shape[2] * c.shape[3];
}

hipError_t launch_satisfy(const SatisfyArgs& A, uint32_t iter, uint32_t cells, uint32_t ib, hipStream_t stream) {
    const bool sw = A.sweep_all != 0, bits = A.frontier_bits != nullptr;
    const SatSingle f = cells == 128    ? pick_single<128, 1>(ib, sw, bits)
                        : cells == 512  ? pick_single<256, 2>(ib, sw, bits)
                        : cells == 1024 ? pick_single<256, 4>(ib, sw, bits)
                                        : pick_single<256, 1>(ib, sw, bits);
    hipLaunchKernelGGL(f, dim3(satisfy_blocks(A.c, cells)), dim3(satisfy_threads(cells)), 0, stream, A, iter);
    return hipGetLastError();
}

hipError_t launch_abstraction_kernels(const AbstractionArgs& A, hipStream_t stream) {
    const dim3 grid((A.c.n_tasks + 3u) / 4u);
    hipLaunchKernelGGL(abstraction_kernel<true>, grid, dim3(256), 0, stream, A);
    hipLaunchKernelGGL(abstraction_kernel<false>, grid, dim3(256), 0, stream, A);
    return hipGetLastError();
}

hipError_t launch_chunk_flags(const SatisfyArgs& A, uint8_t* flags, uint32_t iter, hipStream_t stream) {
    hipLaunchKernelGGL(chunk_flags_kernel, dim3((A.c.n_tasks + 255u) / 256u), dim3(256), 0, stream, A, flags, iter);
    return hipGetLastError();
}

hipError_t launch_satisfy_batch(const BatchInstance* B, const uint32_t* starts, uint32_t n, uint32_t blocks,
                                uint32_t cells, uint32_t ib, bool bits, hipStream_t stream) {
    const SatBatch f = cells == 128    ? pick_batch<128, 1>(ib, bits)
                       : cells == 512  ? pick_batch<256, 2>(ib, bits)
                       : cells == 1024 ? pick_batch<256, 4>(ib, bits)
                                       : pick_batch<256, 1>(ib, bits);
    hipLaunchKernelGGL(f, dim3(blocks), dim3(satisfy_threads(cells)), 0, stream, B, starts, n);
    return hipGetLastError();
}

}  // namespace gsc_amd
