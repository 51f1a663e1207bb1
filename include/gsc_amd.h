/*
 * gsc_amd.h — C ABI of the MI355X (gfx950) grid symbolic-control backend.
 *
 * The reference exposes its GPU backend as two C++ symbols,
 *   bool gsc::cuda_backend_compiled();                       (include/gsc/cuda_api.hpp:87)
 *   GpuRunReport gsc::run_case_cuda_u32(const CaseConfig&);  (include/gsc/cuda_api.hpp:88)
 * whose arguments carry std::string / std::vector / std::function members
 * (include/gsc/config.hpp:41-67). This header re-states that boundary as plain
 * C: POD descriptors, caller-owned buffers, sizes and a status code, so that
 * any FFI (ctypes, cgo, JNI) can bind it. The C++ mirror of the reference
 * interface (gsc::run_case_cuda_u32 returning gsc::GpuRunReport) is layered on
 * top of these entry points in gpu-inner-outer_amd/host/.
 *
 * Threading: every entry point is synchronous on the calling thread. Separate
 * gsc_solver objects may be used from separate threads; one object must not be
 * used from two threads at once.
 */
#ifndef GSC_AMD_H
#define GSC_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GSC_STATE_DIM 4
#define GSC_INPUT_DIM 2
/* Sentinels, include/gsc/types.hpp:22-23. */
#define GSC_INVALID_INPUT 0xFFFFFFFFu
#define GSC_UNREACHABLE_STEP 0xFFFFFFFFu
#define GSC_MESSAGE_LEN 256

typedef enum gsc_status {
    GSC_OK = 0,
    GSC_ERR_INVALID_ARGUMENT = 1, /* null pointer, bad size or buffer too small */
    GSC_ERR_CONFIG = 2,           /* preparation failed; the reference throws these out of
                                     run_case_cuda_u32 (kernels.cu:505, abstraction.hpp:261-265,
                                     grid.hpp:114-137, 150-166) */
    GSC_ERR_NO_DEVICE = 3,        /* kernels.cu:496-501 "no CUDA device found" */
    GSC_ERR_DEVICE = 4,           /* HIP runtime error (kernels.cu:21-27, caught at :856-858) */
    GSC_ERR_OUT_OF_MEMORY = 5,    /* device allocation failed or HBM budget exceeded */
    GSC_ERR_PARSE = 6,            /* .cfg parse error (config.cpp:243-347) */
    GSC_ERR_BUSY = 7              /* serving queue full */
} gsc_status;

typedef enum gsc_constraint_type {
    GSC_CONSTRAINT_NONE = 0,       /* config.hpp:19 */
    GSC_CONSTRAINT_HYPERBOLIC = 1, /* config.hpp:20, config.cpp:64-68 */
    GSC_CONSTRAINT_ELLIPTIC = 2    /* config.hpp:21, config.cpp:71-75 */
} gsc_constraint_type;

/* HyperRect4D, config.hpp:12-15 (closed intervals, config.cpp:79-86). */
typedef struct gsc_rect4 {
    double lb[GSC_STATE_DIM];
    double ub[GSC_STATE_DIM];
} gsc_rect4;

/*
 * CaseConfig, config.hpp:41-67, as a POD. Arrays are borrowed for the duration
 * of a call. The trailing extension fields have no reference counterpart;
 * left zero they reproduce the reference exactly:
 *   extra_targets   : further target boxes; a state is a candidate if its
 *                     centre lies in `candidate` OR in any extra target box.
 *   target_band_*   : when enabled, a centre in `candidate` only counts if
 *                     fabs(x*y - c) <= w as well (a hyperbolic band in (x, y)).
 *   state_allow     : NULL, or state_count bytes (x-fastest flat order); a 0
 *                     removes that state from the valid set. It carries a
 *                     caller-side pointwise constraint (CaseConfig's kCustom
 *                     std::function, config.hpp:38, 63) that cannot cross a C
 *                     ABI: the reference evaluates it in its host mask build
 *                     (abstraction.hpp:97-101) and its GPU box scan treats it
 *                     as no constraint (kernels.cu:55-73, 79), as this backend
 *                     does with constraint_type NONE plus this mask.
 */
typedef struct gsc_case_desc {
    const char* name;
    double state_lb[GSC_STATE_DIM];
    double state_ub[GSC_STATE_DIM];
    double state_eta[GSC_STATE_DIM];
    int32_t wrap_dim;
    double input_lb[GSC_INPUT_DIM];
    double input_ub[GSC_INPUT_DIM];
    double input_eta[GSC_INPUT_DIM];
    gsc_rect4 candidate;
    gsc_rect4 map;
    const gsc_rect4* obstacles;
    uint32_t obstacle_count;
    int32_t constraint_type; /* gsc_constraint_type */
    double hyperbolic_a, hyperbolic_b, hyperbolic_c;
    double elliptic_a, elliptic_b;
    double disturbance_half_width[GSC_STATE_DIM];
    double sample_time;
    uint32_t integration_substeps;
    uint32_t max_iterations;
    int32_t verbose;
    /* extensions */
    const gsc_rect4* extra_targets;
    uint32_t extra_target_count;
    int32_t target_band_enabled;
    double target_band_c;
    double target_band_w;
    const uint8_t* state_allow;
} gsc_case_desc;

/* Grid4D / InputGrid2D / PrefixLayout4D sizes (grid.hpp:12-101, 103-168, 253-269). */
typedef struct gsc_case_shape {
    uint32_t state_shape[GSC_STATE_DIM];
    uint64_t state_strides[GSC_STATE_DIM];
    uint64_t state_count;
    uint32_t input_shape[GSC_INPUT_DIM];
    uint64_t input_strides[GSC_INPUT_DIM];
    uint64_t input_count;
    uint64_t pair_count;
    uint32_t prefix_shape[GSC_STATE_DIM];
    uint64_t prefix_count;
} gsc_case_shape;

/* IterationStats, cuda_api.hpp:11-21. */
typedef struct gsc_iteration_stats {
    uint32_t iteration;
    uint32_t reserved;
    uint64_t newly_reachable;
    uint64_t newly_certified;
    uint64_t total_reachable;
    uint64_t total_certified;
    double iteration_ms;
    double prefix_build_ms;
    double satisfaction_check_ms;
    double reduction_ms;
} gsc_iteration_stats;

/* KernelTimings, cuda_api.hpp:24-49 (same field names and meaning). */
typedef struct gsc_kernel_timings {
    double abstraction_kernel_ms;
    double abstraction_memcpy_h2d_ms;
    double scatter_mask_ms;
    double scan_axis0_ms;
    double scan_axis1_ms;
    double scan_axis2_ms;
    double scan_axis3_ms;
    double prefix_build_total_ms;
    double pair_satisfaction_ms;
    double reduce_inputs_ms;
    double iteration_memcpy_ms;
    double result_memcpy_d2h_ms;
    double total_kernel_ms;
    double total_memcpy_ms;
    double total_compute_ms;
} gsc_kernel_timings;

/* GpuMemoryStats, cuda_api.hpp:52-59. */
typedef struct gsc_memory_stats {
    uint64_t total_allocated_bytes;
    uint64_t peak_usage_bytes;
    uint64_t abstraction_bytes;
    uint64_t prefix_bytes;
    uint64_t iteration_bytes;
    uint64_t result_bytes;
} gsc_memory_stats;

/* MemoryBudget, abstraction.hpp:16-30 (the reference's estimate, u64 prefix tables). */
typedef struct gsc_memory_budget {
    uint64_t pair_min_bytes;
    uint64_t pair_max_bytes;
    uint64_t pair_valid_bytes;
    uint64_t pair_satisfied_bytes;
    uint64_t reachable_prev_bytes;
    uint64_t reachable_next_bytes;
    uint64_t valid_mask_bytes;
    uint64_t candidate_mask_bytes;
    uint64_t controller_bytes;
    uint64_t reach_step_bytes;
    uint64_t prefix_reachable_bytes;
    uint64_t prefix_valid_bytes;
    uint64_t total_bytes;
} gsc_memory_budget;

/* The scalar part of GpuRunReport, cuda_api.hpp:61-85. */
typedef struct gsc_run_report {
    int32_t compiled_with_gpu;
    int32_t runtime_available;
    int32_t executed;
    int32_t converged;
    char message[GSC_MESSAGE_LEN];        /* report.message (errors) */
    char result_message[GSC_MESSAGE_LEN]; /* report.result.message (stop reason) */
    double abstraction_ms;
    double solve_ms;
    double total_ms;
    uint32_t iterations;
    uint32_t iteration_stats_written;
    uint64_t reachable_states;
    uint64_t certified_candidate_states;
    uint64_t candidate_states;
    uint64_t pair_count;
    gsc_case_shape shape;
    gsc_memory_budget budget;
    gsc_kernel_timings kernel_timings;
    gsc_memory_stats memory_stats;
} gsc_run_report;

/*
 * Host buffers filled by a run (any pointer may be NULL to skip that output).
 * State arrays hold state_count entries, abstraction arrays pair_count
 * entries, iteration_stats iteration_stats_capacity entries. The abstraction
 * triplet is the reference layout (kernels.cu:806-824): min_flat/max_flat of
 * the successor cell box and valid, all zero for an invalid pair.
 */
typedef struct gsc_result_buffers {
    uint8_t* valid_mask;
    uint8_t* candidate_mask;
    uint8_t* reachable_mask;
    uint32_t* controller;
    uint32_t* reach_step;
    uint32_t* abstraction_min_flat;
    uint32_t* abstraction_max_flat;
    uint8_t* abstraction_valid;
    gsc_iteration_stats* iteration_stats;
    uint32_t iteration_stats_capacity;
} gsc_result_buffers;

/* Replaces gsc::cuda_backend_compiled (cuda_api.hpp:87, kernels.cu:488-490). Returns 1. */
int gsc_backend_compiled(void);

/* Number of visible HIP devices (kernels.cu:496-497); 0 when none or on error. */
int gsc_device_count(void);

/*
 * Grid shapes of a case (build_state_grid / build_input_grid / build_prefix_layout,
 * grid.hpp:103-168, 253-269). Fails with GSC_ERR_CONFIG and the reference's
 * message for non-positive eta, ub <= lb or a state count above 2^32-1.
 */
int gsc_case_shape_of(const gsc_case_desc* cfg, gsc_case_shape* out, char* err, size_t err_len);

/*
 * Replaces gsc::run_case_cuda_u32 (cuda_api.hpp:88, kernels.cu:492-875): build the
 * masks, the pair abstraction, then iterate the backward fixpoint to convergence,
 * stall or max_iterations, and copy the requested results to `out`.
 * Returns GSC_OK with report->executed = 1 on success. Preparation failures the
 * reference lets escape as exceptions return GSC_ERR_CONFIG; device errors the
 * reference catches into report.message return GSC_ERR_DEVICE/OUT_OF_MEMORY with
 * report->executed = 0. Every device buffer is released before returning.
 */
int gsc_run_case_u32(const gsc_case_desc* cfg, const gsc_result_buffers* out, gsc_run_report* report);

/*
 * gsc_run_case_u32 that allocates nothing. No counterpart in the reference: it
 * is for callers that cannot allocate device memory (serving loops and
 * harnesses that forbid hipMalloc, code that must not allocate while other
 * work is captured or in flight). Every device buffer of the run comes from one
 * file-scope device array of GSC_STATIC_SCRATCH_BYTES per device, held from
 * library load until the process ends; no hipMalloc, hipFree, pinned-host
 * allocation, stream or event creation happens in the call. It runs on the
 * calling thread's current device, on the null stream, and returns the same
 * results as gsc_run_case_u32. A case whose buffers (plus 12 MiB of export
 * staging when the abstraction triplet is requested) exceed the array fails
 * with GSC_ERR_OUT_OF_MEMORY and a message giving both sizes. Calls on one
 * device share its array (each call re-initialises what it reads) and are
 * serialised by a lock inside the library: without it they would race.
 */
#define GSC_STATIC_SCRATCH_BYTES ((size_t)128 << 20)
int gsc_run_case_u32_static(const gsc_case_desc* cfg, const gsc_result_buffers* out, gsc_run_report* report);

/* ---------------------------------------------------------------------------
 * Device-resident solver. No reference counterpart: the reference allocates and
 * frees everything per call (kernels.cu:513-560, 860-873). These entry points
 * keep masks, abstraction and iteration state in HBM so a caller can re-solve,
 * time stages separately, or read back only what it needs.
 * ------------------------------------------------------------------------- */
typedef struct gsc_solver gsc_solver;

typedef struct gsc_solver_options {
    int32_t device;          /* HIP device ordinal */
    int32_t keep_export;     /* 1: also keep the reference abstraction triplet on the device */
    int32_t time_kernels;    /* 1: record per-kernel HIP events (per-iteration stats) */
    int32_t sweep_all_pairs; /* 1: test every pair every iteration (the reference's work
                                schedule); 0: test pairs of unsettled states only */
    int32_t verbose;         /* print per-iteration progress like kernels.cu:746-756 */
    int32_t recompute_abstraction; /* 1: store no pair records; every iteration recomputes the
                                      successor boxes it tests in FP64 (same results) */
    int32_t tile_wg;         /* cells per satisfaction workgroup: 128, 256, 512 or 1024
                                (at most 256 threads; 64-cell chunks per wave above 256);
                                0: 512, or 256 with sweep_all_pairs */
    int32_t tile_ib;         /* inputs per block of record loads: 4 or 8; 0: 8 */
} gsc_solver_options;

typedef struct gsc_solve_summary {
    uint32_t iterations;
    int32_t converged;
    uint64_t reachable_states;
    uint64_t certified_candidate_states;
    uint64_t candidate_states;
    double abstraction_ms;       /* abstraction kernel(s), HIP events */
    double solve_ms;             /* host wall time of the fixpoint loop */
    double prefix_ms;            /* sum of prefix-build kernel time (HIP events) */
    double satisfaction_ms;      /* sum of pair-satisfaction kernel time (HIP events) */
    double reduction_ms;         /* sum of the per-iteration bookkeeping kernel time */
    uint64_t pairs_tested;       /* pairs whose record was read and tested, all iterations */
    uint64_t pair_bytes_read;    /* algorithmic bytes of pair records streamed, all iterations */
    uint64_t boxes_counted;      /* pair tests that needed a 16-corner prefix count */
    uint64_t test_slots;         /* lane slots spent on tests (pairs_tested / test_slots =
                                    lane utilisation of the satisfaction pass) */
    char message[GSC_MESSAGE_LEN];
    /* How unchanged states were skipped: GSC_FRONTIER_NONE (every unsettled
       state tested every iteration: the sweep schedule, or wrap on x or y),
       GSC_FRONTIER_PREFIX (hull counts on prefix tables) or GSC_FRONTIER_BITS
       (per-plane bitmasks; every box settled by bitmask row scans). */
    uint32_t frontier;
    uint32_t reserved;
} gsc_solve_summary;
#define GSC_FRONTIER_NONE 0
#define GSC_FRONTIER_PREFIX 1
#define GSC_FRONTIER_BITS 2

void gsc_solver_default_options(gsc_solver_options* opts);
int gsc_solver_create(const gsc_case_desc* cfg, const gsc_solver_options* opts,
                      gsc_solver** out, char* err, size_t err_len);
int gsc_solver_shape(const gsc_solver* s, gsc_case_shape* out);
/* Compute the pair abstraction on the device (kernels.cu:164-217 equivalent). */
int gsc_solver_build_abstraction(gsc_solver* s, double* kernel_ms);
/* Reset the iteration state to the seeds (kernels.cu:576-598) and run the fixpoint. */
int gsc_solver_solve(gsc_solver* s, gsc_solve_summary* out);
/* Copy results (and, with keep_export, the abstraction triplet) to host buffers. */
int gsc_solver_download(gsc_solver* s, const gsc_result_buffers* out, gsc_run_report* report);
/* Last error text of this solver. */
const char* gsc_solver_last_error(const gsc_solver* s);
void gsc_solver_destroy(gsc_solver* s);

/* ---------------------------------------------------------------------------
 * .cfg parsing (config.cpp:243-347), kept so the CLI and FFI callers read the
 * same files as the reference. Extension keys: `target` (8 numbers, repeatable)
 * and `target_band` (c, w). Returns a handle owning the descriptor and its arrays.
 * ------------------------------------------------------------------------- */
typedef struct gsc_config gsc_config;
int gsc_config_load(const char* path, gsc_config** out, char* err, size_t err_len);
int gsc_config_parse(const char* text, gsc_config** out, char* err, size_t err_len);
const gsc_case_desc* gsc_config_desc(const gsc_config* c);
void gsc_config_free(gsc_config* c);
/* make_default_case / make_wraparound_case / make_smoke_case / make_hyperbolic_case
 * (config.cpp:129-241); which = 0..3. */
int gsc_config_builtin(int which, gsc_config** out);

/* ---------------------------------------------------------------------------
 * Batched solves. No reference counterpart: the reference solves one case per
 * run_case_cuda_u32 call (kernels.cu:492-875). A batch holds many independent
 * cases of uneven grid and input sizes resident on one device and advances all
 * of them together: every stage of an iteration (prefix build, coarse sample,
 * bitmask, satisfaction + reduction, stop test) is one launch over the concatenated
 * work of all unfinished cases, and `iterations_per_launch` such iterations
 * are captured once in a HIP graph that is replayed until every case has
 * stopped. Each case stops on its own rule and keeps its own results, which are
 * identical to a single-case solve of it.
 * ------------------------------------------------------------------------- */
typedef struct gsc_batch gsc_batch;

typedef struct gsc_batch_options {
    int32_t device;
    int32_t tile_wg;               /* cells per satisfaction workgroup: 128, 256, 512, 1024 */
    int32_t tile_ib;               /* inputs per block of record loads: 4 or 8 */
    int32_t iterations_per_launch; /* iterations per graph replay, 1..64 */
    int32_t order;                 /* 0: cases in the given order, 1: largest pair count first */
    int32_t use_graph;             /* 1: replay a captured HIP graph, 0: launch directly */
    int32_t reserved[2];
} gsc_batch_options;

typedef struct gsc_batch_summary {
    uint32_t cases;
    uint32_t replays;          /* graph replays (or direct launch rounds) of the last solve */
    uint32_t max_iterations;   /* largest iteration count over the cases */
    uint32_t stopped;          /* cases that reached a stop rule */
    uint64_t pairs;            /* total state-input pairs */
    double abstraction_ms;     /* wall time of the last gsc_batch_build_abstraction */
    double solve_ms;           /* wall time of the last gsc_batch_solve (seeds to last stop) */
    double gpu_ms;             /* device time of the replays (HIP events) */
} gsc_batch_summary;

typedef struct gsc_batch_tune_entry {
    int32_t tile_wg;
    int32_t tile_ib;
    int32_t iterations_per_launch;
    int32_t order;
    double solve_ms; /* best of the timed solves of this configuration */
} gsc_batch_tune_entry;

void gsc_batch_default_options(gsc_batch_options* opts);
/* Prepare every case (masks, tables, device buffers). Fails with the first
 * case's configuration error, naming its index. */
int gsc_batch_create(const gsc_case_desc* const* cases, uint32_t count, const gsc_batch_options* opts,
                     gsc_batch** out, char* err, size_t err_len);
int gsc_batch_build_abstraction(gsc_batch* b, double* wall_ms);
/* Seeds every case and iterates all of them to their stop rules. */
int gsc_batch_solve(gsc_batch* b, gsc_batch_summary* out);
/* Times each tile / schedule configuration on this batch (full solves, results
 * unchanged), keeps the fastest and reports every one tried. */
int gsc_batch_tune(gsc_batch* b, gsc_batch_tune_entry* entries, uint32_t capacity, uint32_t* written);
int gsc_batch_get_options(const gsc_batch* b, gsc_batch_options* out);
int gsc_batch_set_options(gsc_batch* b, const gsc_batch_options* opts);
/* Per-case results of the last solve, as gsc_solver_solve / gsc_solver_download. */
int gsc_batch_case_summary(gsc_batch* b, uint32_t index, gsc_solve_summary* out);
int gsc_batch_download(gsc_batch* b, uint32_t index, const gsc_result_buffers* out, gsc_run_report* report);
const char* gsc_batch_last_error(const gsc_batch* b);
void gsc_batch_destroy(gsc_batch* b);

/* ---------------------------------------------------------------------------
 * Serving. No reference counterpart (the reference has no request loop). A
 * server owns one device and an HBM budget carved out once at creation; a
 * worker thread admits submitted cases in arrival order while the budget and
 * `max_active` allow, prepares each one (masks, tables, abstraction) while the
 * others keep iterating, and advances every admitted case together: each step
 * runs `iterations_per_step` fused iterations over all of them (continuous
 * batching). A case leaves the device as soon as it stops; its masks,
 * controller, reach steps and iteration history are copied to host memory and
 * its HBM returns to the budget. Results are identical to a single-case solve.
 * ------------------------------------------------------------------------- */
typedef struct gsc_server gsc_server;

typedef struct gsc_server_options {
    int32_t device;
    int32_t max_active;           /* cases iterating together, 1..1024 */
    int32_t iterations_per_step;  /* fused iterations between admission checks, 1..64 */
    int32_t tile_wg;              /* satisfaction tile, as gsc_batch_options */
    int32_t tile_ib;
    int32_t result_pool_mib;      /* pinned host memory holding finished results until fetched
                                     (11 B per state); results that do not fit go to pageable memory */
    int32_t reserved[2];
    uint64_t hbm_budget_bytes;    /* device memory reserved for admitted cases */
} gsc_server_options;

typedef enum gsc_request_state {
    GSC_REQUEST_QUEUED = 0,
    GSC_REQUEST_RUNNING = 1,
    GSC_REQUEST_DONE = 2,
    GSC_REQUEST_FAILED = 3
} gsc_request_state;

typedef struct gsc_request_info {
    uint64_t id;
    int32_t state;               /* gsc_request_state */
    int32_t status;              /* GSC_OK, or the error of a failed request */
    uint64_t state_count;
    uint64_t pair_count;
    uint64_t device_bytes;       /* HBM admitted against the budget */
    double queued_ms;            /* submit -> admitted */
    double latency_ms;           /* submit -> results on the host */
    gsc_solve_summary summary;   /* as gsc_solver_solve (solve_ms = admitted -> stopped) */
    char error[GSC_MESSAGE_LEN];
} gsc_request_info;

typedef struct gsc_server_stats {
    uint64_t submitted;
    uint64_t completed;
    uint64_t failed;
    uint64_t steps;              /* fused iteration steps run */
    uint32_t active;             /* cases on the device now */
    uint32_t queued;
    uint32_t max_active_seen;    /* most cases iterating together */
    uint32_t reserved;
    uint64_t budget_bytes;
    uint64_t budget_in_use;
    uint64_t budget_peak;
} gsc_server_stats;

void gsc_server_default_options(gsc_server_options* opts);
int gsc_server_create(const gsc_server_options* opts, gsc_server** out, char* err, size_t err_len);
/* Thread-safe; returns at once with the request id. */
int gsc_server_submit(gsc_server* s, const gsc_case_desc* cfg, uint64_t* id);
/* Waits up to timeout_ms (< 0: forever) for the request to finish; GSC_ERR_BUSY
 * on timeout. `info` may be NULL. */
int gsc_server_wait(gsc_server* s, uint64_t id, double timeout_ms, gsc_request_info* info);
/* Copies a finished request's state arrays and iteration history (any pointer
 * of `out` may be NULL) and forgets the request. */
int gsc_server_fetch(gsc_server* s, uint64_t id, const gsc_result_buffers* out, gsc_run_report* report);
int gsc_server_get_stats(gsc_server* s, gsc_server_stats* out);
/* Stops admitting, lets running cases finish, joins the worker and frees the device. */
void gsc_server_destroy(gsc_server* s);

#ifdef __cplusplus
}
#endif

#endif /* GSC_AMD_H */
