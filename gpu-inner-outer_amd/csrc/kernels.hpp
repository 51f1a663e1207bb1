// kernels.hpp — kernel entry points and the device-side control block.
#pragma once

#include <cstddef>

#include "gsc_device.hpp"

namespace gsc_amd {

// Per-iteration counters live in a ring so the fixpoint loop never clears
// memory with hipMemset between iterations: finish_kernel consumes slot
// iter % kRing and zeroes it for reuse.
constexpr uint32_t kRing = 64;
constexpr uint32_t kDoneConverged = 1;      // "all candidate states obtained a controller"
constexpr uint32_t kDoneStalled = 2;        // "iteration stalled before candidate closure"
constexpr uint32_t kDoneMaxIterations = 3;  // "maximum iterations reached"

// Workgroups add into one of kCounterSlots slots (by block index) so that tens
// of thousands of workgroups do not serialise on a single address.
constexpr uint32_t kCounterSlots = 32;
// Input padding: the record array holds I_pad = I rounded up to kInputBlock
// inputs per (plane), so every satisfaction tile variant (4 or 8 inputs per
// block of loads) reads whole blocks.
constexpr uint32_t kInputBlock = 8;

struct IterCounters {
    unsigned long long new_reachable[kCounterSlots];
    unsigned long long new_certified[kCounterSlots];
    unsigned long long pairs_tested[kCounterSlots];
    unsigned long long boxes_counted[kCounterSlots];
    unsigned long long test_slots[kCounterSlots];  // lane slots of executed test blocks
};

struct IterCtl {
    uint32_t done;        // 0 while iterating, else kDone*
    uint32_t iterations;  // last finished iteration
    unsigned long long candidate_count;
    unsigned long long total_reachable;
    unsigned long long total_certified;
    IterCounters ring[kRing];
};

struct IterRecord {
    uint32_t iteration;
    uint32_t pad;
    unsigned long long newly_reachable;
    unsigned long long newly_certified;
    unsigned long long total_reachable;
    unsigned long long total_certified;
    unsigned long long pairs_tested;
    unsigned long long boxes_counted;
    unsigned long long test_slots;
};

// How the abstraction kernel builds the records of one (plane, input) row.
constexpr uint32_t kRowGeneral = 0;      // per pair: cell offsets, filter or exact sum; escapes possible
constexpr uint32_t kRowInvalid = 1;      // every pair invalid (θ or v range rejected; padding input)
constexpr uint32_t kRowFast = 2;         // every box from the row's cell offsets, every box fits a record
constexpr uint32_t kRowFastChecked = 3;  // kRowFast plus the constraint rectangle test

// A fast row's cell offsets, x then y, and the cells whose box stays in the
// domain: cell i qualifies on axis d when i - start[d] <= span[d] as unsigned
// 32-bit values. One 32-byte scalar load.
struct FastRow {
    int32_t off_lo[2], off_hi[2];
    int32_t start[2];
    uint32_t span[2];
};
constexpr int32_t kFastNoCell = 0x40000000;  // start of a row no cell qualifies for (span 0)

struct AbstractionArgs {
    DevCase c;
    const double2* inc;          // [plane][input][K] nominal-trajectory increments
    const double* inc_sums;      // [plane][input][4]: (S -/+ rad) / eta for x, then y (increment_sums_kernel)
    const int32_t* cell_offs;    // [plane][input][8]: CellOffsets for x, then y, or kOffsetsUndecided
    const uint32_t* row_kind;    // [plane][I_pad]: kRow* of each (plane, input) (pair_rows_kernel)
    const uint4* row_box;        // [plane][input]: θ range, then v range (box2 / box3 entries)
    const FastRow* fast_offs;    // [plane][I_pad] (pair_rows_kernel)
    const uint32_t* plane_mixed;  // [plane]: 0 when every row builds as a fast row (pair_rows_kernel)
    const double* rad;           // [n_v][2] x / y box radius
    const uint32_t* fail2d;      // 2D prefix of constraint failures ((n1+1) rows of 1 << fail_shift), or null
    const uint32_t* prefix_valid;
    uint32_t* rec;                // null: compute the state hulls only (recompute mode)
    OverflowEntry* overflow;
    uint32_t* overflow_count;
    uint32_t overflow_capacity;
    uint32_t* ext1_max;           // atomic max of the compact records' y extents, or null
    uint32_t* hull;               // [S] state hulls
    uint4* chunk_hull;            // [n_tasks] union of the hulls of each 64-cell chunk
};

struct SatisfyArgs {
    DevCase c;
    // Prefix tables of valid-and-unreached states, double-buffered by iteration
    // parity: iteration k reads table k & 1 and compares it with table (k-1) & 1,
    // built one iteration earlier, to find states whose hull saw no change.
    const uint32_t* prefix_unreached[2];
    const uint32_t* prefix_valid;
    const uint32_t* coarse_unreached[2];  // coarse samples of prefix_unreached[0/1]
    const uint32_t* coarse_valid;         // coarse sample of prefix_valid
    const uint32_t* hull;   // [S] state hulls
    const uint4* chunk_hull;  // [planes * chunks_per_plane] absolute x0, x1, y0, y1 of 64 cells
    // [planes * chunk_flag_pitch] 1 where the chunk's hull saw a change since the
    // previous iteration (chunk_flags_kernel); the single-case solver only.
    const uint8_t* chunk_flags;
    const uint2* hull2;     // [n_θ] union of the θ ranges over ω (lo > hi: wraps)
    const uint2* hull3;     // [n_v] union of the v ranges over a
    const unsigned long long* unreached_bits;
    // [plane][I_pad]: where the row probes of the pair's box start in the
    // probe bitmasks, for the box's lower θ and v cells (pair_rows_kernel).
    const uint32_t* probe_base;
    // Plane-aligned bitmask of the valid, unreached cells ORed with their θ + 1,
    // v + 1 and θ + 1 / v + 1 neighbours, written each iteration by the chunk
    // flag pass (chunk_flags_body); null where that pass does not run, and then
    // no probe base points into it.
    const unsigned long long* probe_or_bits;
    // Bitmask of the valid states (flat state order, one word past the last
    // state's), written once at set-up.
    const unsigned long long* valid_bits;
    // Plane-aligned bitmask (one word per 64-cell chunk) of the states whose
    // hull may have seen a newly reached cell since the previous iteration,
    // written each iteration by frontier_plane_kernel; null where frontier
    // skipping runs on the prefix tables instead (chunk_flags, hull tests).
    const unsigned long long* frontier_bits;
    uint32_t* rec;
    const OverflowEntry* overflow;
    const uint2* box2;
    const uint2* box3;
    const uint8_t* valid;
    const uint8_t* cand;
    uint8_t* reach;
    uint32_t* ctrl;
    uint32_t* rstep;
    IterCtl* ctl;
    uint32_t sweep_all;
};

// Satisfaction without stored pair records: every tested pair's successor box
// is recomputed from the trajectory tables in the same launch.
struct RecomputeArgs {
    SatisfyArgs sat;       // rec is unused
    const double2* inc;    // as AbstractionArgs
    const double* inc_sums;
    const int32_t* cell_offs;
    const double* rad;
    const uint32_t* fail2d;
};

// Frontier skipping from bitmasks (single case, see frontier_plane_kernel).
struct FrontierArgs {
    DevCase c;
    const uint8_t* valid;
    const uint8_t* reach;
    const unsigned long long* valid_bits;
    unsigned long long* cur;           // valid, unreached states of this iteration (flat)
    unsigned long long* fresh;         // states reached since the previous iteration (flat)
    uint32_t* fresh_v;                 // per plane, row-aligned: `fresh` ORed over the v hull range
    const unsigned long long* keep;    // plane-aligned: states with a valid pair
    const unsigned long long* always;  // plane-aligned: states whose hull is not representable
    const int4* plane_rect;            // per plane: union of its states' hull offsets (x0, x1, y0, y1)
    const uint2* hull2;
    const uint2* hull3;
    unsigned long long* frontier;      // out, plane-aligned
    unsigned long long* or_bits;       // out: the OR probe bitmask, or null
    const IterCtl* ctl;
    uint32_t row_words;                // LDS words per y row: ceil(n_x / 32)
};
// LDS of frontier_plane_kernel: two row-aligned copies of one plane.
__host__ __device__ inline uint32_t frontier_lds_bytes(const DevCase& c) {
    return 2u * c.shape[1] * ((c.shape[0] + 31u) / 32u) * 4u;
}
constexpr uint32_t kFrontierLdsMax = 65536;
constexpr uint32_t kFrontierThreads = 512;


// One instance of a multi-instance (batched) iteration: its satisfaction
// arguments plus what the prefix, bitmask, seed and finish stages write.
struct BatchInstance {
    SatisfyArgs sat;
    FrontierArgs fr;  // the bitmask frontier's arguments (batches that run it)
    uint32_t* prefix_unreached[2];
    uint32_t* coarse_unreached[2];
    unsigned long long* unreached_bits;
    IterRecord* records;
    unsigned long long candidate_count;
    uint32_t max_iterations;
    uint32_t plane_pitch;  // LDS row pitch of the fused prefix build (0: the plane does not fit)
};

// Per-instance state published by finish_batch_kernel after every iteration.
struct BatchStatus {
    uint32_t done;
    uint32_t iterations;
    unsigned long long total_reachable;
    unsigned long long total_certified;
};

__global__ void masks_kernel(DevCase c, const gsc_rect4* obstacles, const gsc_rect4* targets, const uint8_t* allow,
                             uint8_t* valid, uint8_t* cand, unsigned long long* counts);
__global__ void seed_kernel(uint32_t S, const uint8_t* cand, uint8_t* reach, uint32_t* ctrl, uint32_t* rstep);
__global__ void reset_ctl_kernel(IterCtl* ctl, unsigned long long candidate_count);
__global__ void fill_u32_kernel(uint32_t* p, uint32_t n, uint32_t value);
// With ctl: the table of the coming iteration k = ctl->iterations + 1 is P[k & 1];
// without: P0.
__global__ void prefix_rows_kernel(DevCase c, const uint8_t* a, const uint8_t* b, uint32_t* P0, uint32_t* P1,
                                   const IterCtl* ctl);
__global__ void prefix_axis_kernel(DevCase c, int axis, uint32_t* P0, uint32_t* P1, const IterCtl* ctl);
__global__ void increments_kernel(DevCase c, const double2* cs, const double* vt, double2* inc);
// Fused prefix build for planes that fit in LDS (see kernels.hip).
constexpr uint32_t kPlaneThreads = 1024;
constexpr uint32_t kPlaneLdsMax = 65536;  // bytes of LDS for one plane's 2D prefix
// The plane kernel's LDS: the 2D prefix (pitch words per row), then one bit per
// staged state, for the plane's PQ states plus up to 63 on either side.
__host__ __device__ inline uint32_t plane_stage_words(uint32_t PQ) { return (PQ + 127u) / 32u + 3u; }
__host__ __device__ inline uint32_t plane_lds_bytes(uint32_t pitch, uint32_t n1, uint32_t PQ) {
    return (pitch * n1 + plane_stage_words(PQ)) * 4u;
}
__global__ void prefix_plane_kernel(DevCase c, const uint8_t* a, const uint8_t* b, uint32_t* P0, uint32_t* P1,
                                    unsigned long long* bits, uint32_t pitch, const IterCtl* ctl);
__global__ void prefix_last_axis_kernel(DevCase c, uint32_t* P0, uint32_t* P1, uint32_t* C0, uint32_t* C1,
                                        const IterCtl* ctl);
__global__ void increment_sums_kernel(DevCase c, const double2* inc, const double* rad, double* sums, int32_t* offs);
// Coarse sample of a prefix table (with ctl: of table k & 1 into coarse k & 1).
__global__ void coarse_kernel(DevCase c, const uint32_t* P0, const uint32_t* P1, uint32_t* C0, uint32_t* C1,
                              const IterCtl* ctl);
__global__ void coarse_batch_kernel(const BatchInstance* B, const uint32_t* starts, uint32_t n);
__global__ void unreached_bits_kernel(uint32_t S, const uint8_t* valid, const uint8_t* reach, unsigned long long* bits,
                                      const IterCtl* ctl);
// Bitmask of the valid states (bit s of word s / 64), once per set-up.
__global__ void valid_bits_kernel(uint32_t S, const uint8_t* valid, unsigned long long* bits);
// Row probes (see satisfy_kernel): a row of x cells at one y, θ and v, as bits
// [row, row + extent] of a bitmask. unreached_bits (flat state order, one word
// past the last state's) is followed in the same allocation by probe_or_bits
// (n_tasks + 1 words), so both are read from one base: a row starts at bit
// x + y*n_x + plane*PQ of the first, or at bit probe_or_bit0(c) +
// plane*probe_or_pitch(c) + x + y*n_x, where each plane starts on a word.
// The valid bitmask (probe_bits_words words, flat order like the first)
// follows the OR bitmask in the same allocation.
__host__ __device__ inline uint32_t probe_or_pitch(const DevCase& c) { return c.chunks_per_plane * 64u; }
__host__ __device__ inline uint32_t probe_bits_words(const DevCase& c) { return (c.S + 63u) / 64u + 1u; }
__host__ __device__ inline uint32_t probe_or_bit0(const DevCase& c) { return probe_bits_words(c) * 64u; }
// Per (plane, input): the probe base, the row kind and the θ / v ranges
// (use_or: the OR bitmask is built, the chunk flag pass runs, for this case).
__global__ void pair_rows_kernel(DevCase c, const uint2* box2, const uint2* box3, const int32_t* offs,
                                 uint32_t use_or, uint32_t* probe_base, uint32_t* row_kind, uint4* row_box,
                                 FastRow* fast_offs, uint32_t* plane_mixed, uint32_t* ext1_max);
// The abstraction kernels (fast planes, then the others) over every wave task.
hipError_t launch_abstraction_kernels(const AbstractionArgs& A, hipStream_t stream);
__global__ void export_kernel(DevCase c, const uint32_t* rec, const uint2* box2, const uint2* box3,
                              const OverflowEntry* overflow, uint64_t pair_begin, uint64_t pair_end,
                              uint32_t* out_min, uint32_t* out_max, uint8_t* out_valid);
__global__ void finish_kernel(IterCtl* ctl, IterRecord* records, uint32_t max_iterations);

// Satisfaction tile variants: WG cells (and threads) per workgroup, IB inputs per
// block of record loads.
bool satisfy_variant_supported(uint32_t cells, uint32_t ib);
// Satisfaction tile: `cells` per workgroup (128, 256, 512 or 1024; 256 threads
// with 2 or 4 64-cell chunks per wave above 256) and `ib` inputs per block.
uint32_t satisfy_blocks(const DevCase& c, uint32_t cells);
hipError_t launch_satisfy(const SatisfyArgs& A, uint32_t iter, uint32_t cells, uint32_t ib, hipStream_t stream);
// Chunk flags: row pitch per plane (a multiple of 16, so that a workgroup of up
// to 16 waves reads its flags as one aligned word) and the launch that fills
// them for iteration `iter` from the chunk hulls and the coarse tables.
__host__ __device__ inline uint32_t chunk_flag_pitch(const DevCase& c) { return (c.chunks_per_plane + 15u) & ~15u; }
hipError_t launch_chunk_flags(const SatisfyArgs& A, uint8_t* flags, uint32_t iter, hipStream_t stream);
// Once per abstraction: the keep / always masks and plane hull rectangles.
__global__ void frontier_setup_kernel(DevCase c, const uint32_t* hull, unsigned long long* keep,
                                      unsigned long long* always, int4* plane_rect);
// Per iteration: this iteration's unreached bitmask into `cur`, and into
// `fresh` the states unreached before (valid, before the first iteration) and
// reached now; eight states per thread.
__global__ void frontier_bits_kernel(FrontierArgs F, uint32_t iter);
// Per iteration, after frontier_bits_kernel: the v pass, then one workgroup
// per plane (frontier words and the OR probe bitmask).
hipError_t launch_frontier_plane(const FrontierArgs& F, hipStream_t stream);
__global__ void satisfy_recompute_kernel(RecomputeArgs R);

// Multi-instance stages: `starts` holds n + 1 block offsets per stage.
__global__ void prefix_rows_batch_kernel(const BatchInstance* B, const uint32_t* starts, uint32_t n);
__global__ void prefix_axis_batch_kernel(const BatchInstance* B, const uint32_t* starts, uint32_t n, int axis);
__global__ void unreached_bits_batch_kernel(const BatchInstance* B, const uint32_t* starts, uint32_t n);
__global__ void frontier_bits_batch_kernel(const BatchInstance* B, const uint32_t* starts, uint32_t n);
__global__ void frontier_v_batch_kernel(const BatchInstance* B, const uint32_t* starts, uint32_t n);
__global__ void frontier_plane_batch_kernel(const BatchInstance* B, const uint32_t* starts, uint32_t n);
__global__ void finish_batch_kernel(const BatchInstance* B, uint32_t n, BatchStatus* status);
__global__ void seed_batch_kernel(const BatchInstance* B, const uint32_t* starts, uint32_t n);
__global__ void chunk_flags_batch_kernel(const BatchInstance* B, const uint32_t* starts, uint32_t n);
__global__ void prefix_plane_batch_kernel(const BatchInstance* B, const uint32_t* starts, uint32_t n);
__global__ void prefix_last_axis_batch_kernel(const BatchInstance* B, const uint32_t* starts, uint32_t n);
hipError_t launch_satisfy_batch(const BatchInstance* B, const uint32_t* starts, uint32_t n, uint32_t blocks,
                                uint32_t wg, uint32_t ib, bool bits, hipStream_t stream);

}  // namespace gsc_amd
