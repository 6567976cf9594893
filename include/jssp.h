/*
 * jssp.h -- C ABI of the MI355X beam-search engine for job-shop scheduling.
 *
 * Plain C types only (no HIP or torch types), so cgo / JNI / ctypes / N-API
 * can bind it directly.  Every function returns a jssp_status (0 = OK) unless
 * stated otherwise; no function calls exit().  The message of the last error
 * on an engine is jssp_last_error(engine).
 *
 * Reference interfaces replaced (the reference has no FFI layer; drivers link
 * the solver sources directly, reference Makefile:55-93):
 *   beam_search_ich         cpu/v0_ich/solver.h:6      -> jssp_solve, mode ICH_EXACT
 *   beam_search_pms         cpu/v1_pms/solver.h:6      -> jssp_solve, mode PMS_EXACT
 *   run_gpu_beam_search_v1  gpu/v1/solver.cuh:6-12     -> jssp_solve, mode PMS_GPU, phases 1
 *   run_gpu_beam_search_v2  gpu/v2_twophases/solver.cuh:6-12 -> jssp_solve, mode PMS_GPU, phases 2
 *   parse_instance          cpu/common/instance_parser.h:9   -> jssp_parse_instance_file
 *   resolve_instance_file   cpu/common/instance_parser.h:6   -> jssp_resolve_instance_file
 *   drivers' multi-pass loop + MPI_Allreduce(MIN)  gpu/v1/main.cu:55-71,
 *                           cpu/v1_pms/main.cpp:55-73  -> jssp_run_passes (also
 *                           jssp_solve_seeded, jssp_solve with n_seeds/n_passes)
 *   run_hybrid_iteration_v1/_v2  hybrid/v1/solver.cuh, hybrid/v2_twophases/solver.cuh
 *                           -> jssp_hybrid_iteration; hybrid drivers -> jssp_run_hybrid
 * C++ callers that want the reference's exact function signatures include
 * jssp_compat.hpp instead.
 */
#ifndef JSSP_H_
#define JSSP_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* "No makespan found" sentinel of the reference (cpu/common/jssp_types.h:33,
 * gpu/v1/solver.cu:277). */
#define JSSP_NO_MAKESPAN 500000

/* Limits of this engine (the reference's are n <= 100, m <= 31,
 * n*m + 2 <= 2100, gpu/common/gpu_common.cuh:20-22). */
#define JSSP_MAX_JOBS 128
#define JSSP_MAX_MACHINES 255
#define JSSP_MAX_BEAM 65536

typedef enum jssp_status {
  JSSP_OK = 0,
  JSSP_ERR_INVALID = 1,   /* bad argument or malformed instance */
  JSSP_ERR_LIMIT = 2,     /* beyond JSSP_MAX_* */
  JSSP_ERR_HIP = 3,       /* HIP runtime error (no device, launch failure ...) */
  JSSP_ERR_NOMEM = 4,     /* device or host allocation failed */
  JSSP_ERR_IO = 5,        /* file not found / unreadable */
  JSSP_ERR_BUDGET = 6,    /* request larger than the server's HBM budget */
  JSSP_ERR_PENDING = 7,   /* serving: result not ready yet */
  JSSP_ERR_NOTFOUND = 8,  /* serving: unknown ticket */
  JSSP_ERR_CANCELLED = 9  /* serving: request cancelled or timed out */
} jssp_status;

typedef enum jssp_mode {
  /* cpu/v0_ich/solver.cpp:8-54 bit-exact: operation-level branching, score
   * cmax + (rand()%100)*0.01 from the process-global rand(), bounded
   * max-heap selection.  The GPU expands every child; the host replays the
   * rand() stream and the heap. */
  JSSP_MODE_ICH_EXACT = 0,
  /* Operation-level branching with the GPU reference's FP32 score and hash
   * noise, ready list in erase-then-append order; fully on the device. */
  JSSP_MODE_ICH_DET = 1,
  /* gpu/v1/solver.cu:9-157 semantics: hash-chosen pivot machine, ready list
   * in swap-remove order, FP32 score, survivors sorted by (score, child). */
  JSSP_MODE_PMS_GPU = 2,
  /* cpu/v1_pms/solver.cpp:8-70 bit-exact (rand() pivot and noise, FP64
   * normalised score, prune vs best_known); host replays rand() and heap. */
  JSSP_MODE_PMS_EXACT = 3
} jssp_mode;

/* Flags for jssp_params.flags */
#define JSSP_FLAG_GLOBAL_RAND 0x1u  /* exact modes: do not srand(seed); continue the process-global rand() stream */
#define JSSP_FLAG_RADIX_CUT 0x2u    /* diagnostics: cut the selection threshold bin with the generic radix select only */
#define JSSP_FLAG_GENERIC_DIGIT 0x4u /* diagnostics: use the generic score-bit selection digit only */
#define JSSP_FLAG_UNFUSED 0x8u       /* diagnostics: score each level in its own kernel, never inside the
                                       previous level's rank kernel (same results) */

typedef struct jssp_instance {
  int32_t n_jobs;
  int32_t n_machines;
  const int32_t* machine;   /* [n_jobs * n_machines], 0-based machine of op k of job j at j*m + k */
  const int32_t* duration;  /* [n_jobs * n_machines] */
} jssp_instance;

typedef struct jssp_params {
  int32_t beam_width;   /* K >= 1 */
  int32_t mode;         /* jssp_mode */
  int32_t phases;       /* 1: materialise every child, then select (gpu/v1);
                           2: score only, select, materialise the K survivors (gpu/v2) */
  int32_t best_known;   /* prune bound; JSSP_NO_MAKESPAN for none */
  uint32_t seed;        /* GPU modes: random_seed of the reference (level seed = seed + level);
                           exact modes: srand() seed unless JSSP_FLAG_GLOBAL_RAND */
  int32_t n_seeds;      /* PMS_GPU / ICH_DET: independent beams with seeds seed, seed+1, ...
                           sharing one device best bound (0 or 1 = single beam) */
  int32_t n_passes;     /* multi-seed: passes; pass e prunes with the bound after pass e-1 (0 = 1) */
  uint32_t flags;
} jssp_params;

typedef struct jssp_result {
  int32_t makespan;         /* PMS_GPU / ICH_DET: min(best_known, best found) as gpu/v1/solver.cu:277-297;
                               ICH_EXACT: best cmax of the final beam; PMS_EXACT: best complete cmax */
  int32_t found;            /* 1 when `sequence` was filled */
  int32_t* sequence;        /* caller-owned [n*m] or NULL: 1-based reference op ids (j*m + k + 1) */
  int64_t expanded_states;  /* children evaluated, summed over levels */
  int32_t levels;           /* levels run */
  int32_t best_seed_index;  /* multi-seed: which beam produced the result */
  double device_ms;         /* wall time of the solve on the engine's stream (HIP events) */
} jssp_result;

typedef struct jssp_engine jssp_engine;

/* Engine: owns one HIP stream, a growable device arena and the tuning table.
 * One engine per host thread (engines are independent). */
int jssp_engine_create(int device, jssp_engine** out);
void jssp_engine_destroy(jssp_engine* engine);
const char* jssp_last_error(const jssp_engine* engine);

/* One solve.  Inputs are host pointers; they are uploaded inside the call. */
int jssp_solve(jssp_engine* engine, const jssp_instance* inst, const jssp_params* params,
               jssp_result* result);

/* As jssp_solve, and also returns the per-level survivor trail
 * (trail[l*K + q] = (parent << 8) | job for survivor q after level l,
 * (n*m - 1) * K entries) and pcount[l] (n*m + 1 entries).  Either may be NULL. */
int jssp_solve_trace(jssp_engine* engine, const jssp_instance* inst, const jssp_params* params,
                     jssp_result* result, int32_t* trail, int32_t* pcount);

/* Multi-seed / multi-pass solve with explicit level seeds: beam i of pass e
 * uses seeds[e * n_seeds + i] (n_seeds * n_passes entries) instead of
 * seed + e * n_seeds + i.  This is the reference drivers' loop
 * (gpu/v1/main.cu:60-71, gpu/v2_twophases/main.cu:60-71: seed = rand() per
 * iteration on each rank, best_known = the all-reduced minimum of the
 * previous iteration) with the ranks as the beams of a pass.  Modes ICH_DET
 * and PMS_GPU. */
int jssp_solve_seeded(jssp_engine* engine, const jssp_instance* inst, const jssp_params* params,
                      const uint32_t* seeds, jssp_result* result);

/* The reference's multi-rank drivers on one device: gpu/v1/main.cu:55-71,
 * gpu/v2_twophases/main.cu:55-71 (PMS_GPU) and cpu/v1_pms/main.cpp:55-73
 * (PMS_EXACT), also the per-instance loops of gpu/v1/benchmark.cu:52-66 and
 * cpu/v1_pms/benchmark.cpp:48-66.  Rank r draws from its own rand() stream,
 * seeded srand(time_seed + r); each of `iterations` passes runs one solve per
 * rank with best_known = the minimum over every rank of the passes before it
 * (MPI_Allreduce MIN), starting from JSSP_NO_MAKESPAN.
 *   PMS_GPU:   the rank's solve seed is its next rand() value (seed = rand();
 *              run_gpu_beam_search_v1/_v2(inst, K, global_best, seed)); the
 *              ranks of a pass run as the beams of one batched launch.
 *   PMS_EXACT: beam_search_pms(inst, K, global_best) consuming the rank's
 *              stream (cpu/v1_pms/solver.cpp:32,52).
 * result: makespan = the final global best, sequence = the first (pass, rank)
 * solve that reached it, best_seed_index = pass * ranks + rank,
 * expanded_states summed over all solves.  pass_best[e] (`iterations`
 * entries, may be NULL) = the global best after pass e.  The process-wide
 * rand() state is not touched. */
typedef struct jssp_driver_params {
  int32_t beam_width;  /* K */
  int32_t mode;        /* JSSP_MODE_PMS_GPU or JSSP_MODE_PMS_EXACT */
  int32_t phases;      /* PMS_GPU: 1 = gpu/v1 driver, 2 = gpu/v2_twophases driver */
  int32_t ranks;       /* MPI world size (>= 1) */
  int32_t iterations;  /* passes (>= 1) */
  uint32_t time_seed;  /* the time(NULL) value the reference driver read */
} jssp_driver_params;
int jssp_run_passes(jssp_engine* engine, const jssp_instance* inst, const jssp_driver_params* params,
                    jssp_result* result, int32_t* pass_best);

/* out[i] = the i-th value of rand() after srand(seed) (glibc's generator),
 * computed without touching the process-wide rand() state. */
void jssp_rand_stream(uint32_t seed, int32_t count, int32_t* out);

/* BS-PMS on the host (cpu/v1_pms/solver.cpp:8-70; the CPU side of the hybrid
 * drivers), one pass: srand(seed) then beam_search_pms(inst, K, best_known),
 * from a private rand() stream.  Needs no engine or device.  result as
 * PMS_EXACT (makespan = best complete cmax; sequence filled when found and
 * result->sequence is set; device_ms = 0). */
int jssp_host_solve_pms(const jssp_instance* inst, int32_t beam_width, int32_t best_known, uint32_t seed,
                        jssp_result* result);

/* One hybrid CPU+GPU iteration (hybrid/v1/solver.cu:11-44,
 * hybrid/v2_twophases/solver.cu:11-44): the device pass (PMS_GPU, beam
 * gpu_beam_width, seed gpu_seed; phases 1 = gpu/v1, 2 = gpu/v2) and
 * threads - 1 host BS-PMS passes (beam cpu_beam_width, thread tid drawing
 * srand(cpu_seed_base + tid), tid = 1 .. threads - 1), all pruning with
 * best_known, run at the same time.  result->makespan = min(best_known,
 * every pass); best_seed_index = the thread that reached it first (0 = the
 * device pass, then threads in order); sequence filled when a pass improved
 * on best_known. */
int jssp_hybrid_iteration(jssp_engine* engine, const jssp_instance* inst, int32_t gpu_beam_width,
                          int32_t cpu_beam_width, int32_t best_known, uint32_t gpu_seed, int32_t phases,
                          int32_t threads, uint32_t cpu_seed_base, jssp_result* result);

/* The hybrid drivers (hybrid/v1/main.cu:45-85, hybrid/v2_twophases/main.cu,
 * and the per-instance loops of hybrid/v1/benchmark.cu and
 * hybrid/v2_twophases/benchmark.cu): rank r seeds
 * srand(time_seed + r); each iteration every rank runs jssp_hybrid_iteration
 * with gpu_seed = its next rand(), cpu_beam_width = max(1, gpu_beam_width /
 * accel_factor), cpu_seed_base = time_seed + r*1000 + iteration*100 and
 * best_known = the minimum over all ranks of the iterations before
 * (MPI_Allreduce MIN, from JSSP_NO_MAKESPAN).  The ranks' device passes of an
 * iteration run as one batch, their host passes on a thread pool beside it.
 * result / pass_best as jssp_run_passes; best_seed_index = (iteration *
 * ranks + rank) * threads + tid. */
typedef struct jssp_hybrid_params {
  int32_t gpu_beam_width;
  int32_t accel_factor;  /* >= 1 (the reference's default is 34) */
  int32_t phases;        /* 1 = hybrid/v1, 2 = hybrid/v2_twophases */
  int32_t threads;       /* threads per rank (OpenMP threads): 1 device + threads-1 host passes */
  int32_t ranks;
  int32_t iterations;
  uint32_t time_seed;
} jssp_hybrid_params;
int jssp_run_hybrid(jssp_engine* engine, const jssp_instance* inst, const jssp_hybrid_params* params,
                    jssp_result* result, int32_t* pass_best);

/* As jssp_solve for one ICH_DET or PMS_GPU beam (n_seeds, n_passes <= 1),
 * but allocating nothing: no counterpart in the reference.  For callers that
 * cannot allocate device or pinned host memory per call (tight serving
 * loops, harnesses that forbid allocation; the engine itself is created
 * beforehand).  The workspace is the library's static per-device scratch
 * array of JSSP_NOALLOC_SCRATCH_BYTES (64 MiB, held by every device the
 * library runs on until the process ends); the instance is copied up from
 * the caller's arrays and the results come back into `result` (sequence
 * filled only when found).  JSSP_ERR_LIMIT when the solve's workspace does
 * not fit, e.g. 50x20 K=4096 fits two-phase (~31 MB) but not single-phase
 * (~130 MB).  Calls on one device share its array: the library serialises
 * them (each call is synchronous).  Same results as jssp_solve. */
#define JSSP_NOALLOC_SCRATCH_BYTES (64ll << 20)
int jssp_solve_noalloc(jssp_engine* engine, const jssp_instance* inst, const jssp_params* params,
                       jssp_result* result);

/* Batched solve: advances all `count` instances together, one level per
 * step; a step is three launches (score, filter, rank) whose grids cover every
 * instance, replayed from a captured hipGraph.  Results equal those of
 * jssp_solve on each instance alone.  Modes ICH_DET and PMS_GPU. */
int jssp_solve_batch(jssp_engine* engine, int32_t count, const jssp_instance* insts,
                     const jssp_params* params, jssp_result* results);
/* As jssp_solve_batch; use_graph = 0 enqueues the step kernels directly
 * instead of replaying the captured step graph (same results). */
int jssp_solve_batch_ex(jssp_engine* engine, int32_t count, const jssp_instance* insts,
                        const jssp_params* params, jssp_result* results, int32_t use_graph);

/* Batch-shape autotuner: per (n, m, K) class of a batch, (1) the score shape:
 * a separate score kernel with 16, 32 or 64 lanes per parent (tile = 16
 * waves), or (lanes = 1) the fused form where the rank kernel scores the next
 * level, one wave per parent (two-phase device modes with the exact selection
 * digit only); (2) the rank block shape: 8-wave blocks ranking 4 survivors
 * per wave in one pass over the staged keys, 8-wave blocks ranking 1, or
 * 16-wave blocks ranking 1; (3) the filter block cap per instance (16, 64 or
 * 256).  When on (the default), a batch first times each class it has not
 * seen, stage by stage, on a batch of the class's instances repeated until
 * about 4 M children per level fill the device, over up to 16 full-width
 * levels, and keeps the fastest; table_path (or NULL) persists the table as
 * text lines "n m log2K lanes rank_shape filter_cap_index" (rank_shape 0, 1, 2
 * as listed; filter_cap_index 0 = default 64, 1 = 16, 2 = 64, 3 = 256; lines
 * with only the first four fields keep the defaults).  Without a tuned entry
 * a batch uses the fused form where it applies, 8-wave blocks ranking 4
 * survivors per wave and 64 filter blocks.  The shape never changes results. */
int jssp_engine_set_autotune(jssp_engine* engine, int32_t on, const char* table_path);
/* Tuned score shape of a class (1 = fused, else lanes per parent), 0 if not tuned yet. */
int jssp_engine_tuned_lanes(const jssp_engine* engine, int32_t n_jobs, int32_t n_machines, int32_t beam_width);
/* Whole tuned shape of a class: returns 1 and fills the non-NULL outputs
 * (lanes as jssp_engine_tuned_lanes; rank waves per block, survivors ranked
 * per wave in one pass, filter blocks per instance), 0 if not tuned yet. */
int jssp_engine_tuned_shape(const jssp_engine* engine, int32_t n_jobs, int32_t n_machines, int32_t beam_width,
                            int32_t* lanes, int32_t* rank_waves, int32_t* rank_batch, int32_t* filter_blocks);

/* Per-kernel timing accumulated over the solves run since profiling was
 * switched on (jssp_engine_set_profiling(engine, 1) also resets it), for
 * roofline reporting.  kind: 0 = expand/score (+ first selection
 * histogram), 1 = selection filter, 2 = rank + materialise/gather.
 * total_ms is the sum of the kernels' durations, from HIP events that each
 * launch records around its own execution (hipExtLaunchKernelGGL) on the
 * engine's stream; algorithmic_bytes the bytes the algorithm must move. */
int jssp_engine_set_profiling(jssp_engine* engine, int32_t on);
int jssp_engine_kernel_time(const jssp_engine* engine, int32_t kind, double* total_ms,
                            int64_t* launches, double* algorithmic_bytes);
/* Per launch of `kind` in the last profiled single-beam device solve, in
 * level order: duration in ms (-1 if not measured) and algorithmic bytes,
 * up to `capacity` entries; returns the number of launches (-1 on bad
 * arguments).  Shows the full-width levels apart from the narrow ones. */
int jssp_engine_kernel_launches(const jssp_engine* engine, int32_t kind, float* ms, double* algorithmic_bytes,
                                int32_t capacity);
/* Selection diagnostics of the last profiled single-beam device solve, per
 * level: out[0..9] = kept children, survivors, score-bit lower bound, digit
 * shift, threshold bin, keys below it, keys in it, survivors taken from it,
 * sure / tie counters. */
int jssp_engine_level_stats(const jssp_engine* engine, int32_t level, uint32_t* out, int32_t count);

/* ---- serving: continuous batching with HBM-budget admission ------------ */
typedef struct jssp_server jssp_server;
typedef struct jssp_server_config {
  int64_t hbm_budget_bytes;  /* admission budget for resident beams */
  int32_t max_slots;         /* concurrently resident requests */
  int32_t reserved;
} jssp_server_config;

int jssp_server_create(jssp_engine* engine, const jssp_server_config* cfg, jssp_server** out);
void jssp_server_destroy(jssp_server* server);
/* Queues a request; inputs are copied.  *ticket identifies it. */
int jssp_serve_submit(jssp_server* server, const jssp_instance* inst, const jssp_params* params,
                      uint64_t* ticket);
/* Advances the device by up to `levels` level steps, admitting queued work as
 * slots and budget free up.  Returns the number of active requests in *active.
 * Steps are pipelined: the call returns with its step still running on the
 * device and collects the requests that the previous step finished, so a
 * request that finishes during a step is reported by the following call (a
 * call with levels == 0 waits for everything enqueued and collects it). */
int jssp_serve_step(jssp_server* server, int32_t levels, int32_t* active);
/* JSSP_OK and the result when done, JSSP_ERR_PENDING otherwise.  The result's
 * sequence pointer is caller-owned as in jssp_solve. */
int jssp_serve_poll(jssp_server* server, uint64_t ticket, jssp_result* result);
int jssp_serve_cancel(jssp_server* server, uint64_t ticket);
/* Runs until every submitted request is finished. */
int jssp_serve_drain(jssp_server* server);
/* Submission / admission / completion times (seconds, steady clock) of a ticket. */
int jssp_serve_times(jssp_server* server, uint64_t ticket, double* submit_s, double* admit_s,
                     double* done_s);

/* ---- host-side instance I/O and checks (no device needed) -------------- */
/* resolve_instance_file (cpu/common/instance_parser.cpp:8-34): a direct path,
 * then ./laXX.txt, ./taXX.txt, {.,..,../..}/lawrence_instances/laXX.txt,
 * {.,..,../..}/taillard_instances/taXX.  Returns 1 if found. */
int jssp_resolve_instance_file(const char* arg, char* out_path, int32_t path_size);
/* parse_instance (cpu/common/instance_parser.cpp:37-93): header "n m", then
 * n rows of m (machine, duration) pairs, 0-based machines.  Fills up to `cap`
 * ops.  A truncated body is an error here (the reference silently reads
 * machine 0 / duration 0). */
int jssp_parse_instance_file(const char* path, int32_t* n_jobs, int32_t* n_machines,
                             int32_t* machine, int32_t* duration, int32_t cap);
/* Seeded synthetic instance: per job a Fisher-Yates machine order, durations
 * uniform in [dmin, dmax] (splitmix64 stream). */
int jssp_generate_instance(uint64_t seed, int32_t n_jobs, int32_t n_machines, int32_t dmin,
                           int32_t dmax, int32_t* machine, int32_t* duration);
/* Re-simulates a sequence of 1-based op ids; *makespan = -1 if infeasible. */
int jssp_validate_sequence(const jssp_instance* inst, const int32_t* sequence, int32_t* makespan);
/* max(max job length, max machine load) */
int32_t jssp_lower_bound(const jssp_instance* inst);

/* Library version string. */
const char* jssp_version(void);

#ifdef __cplusplus
}
#endif
#endif /* JSSP_H_ */
