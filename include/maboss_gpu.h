/*
 * MaBoSS MI355X engine: C-ABI of the simulate/statistics path.
 *
 * The reference (MaBoSS.GPU) has no library API; its host/device cut is the
 * kernel_compiler / kernel_wrapper / stats trio plus the CLI.  Every entry
 * point below names the reference interface it replaces.  Conventions:
 *   - every function returns an int status (MBG_OK = 0) and never exits the
 *     process; mbg_last_error() describes the last failure of this thread;
 *   - handles are opaque; objects created by the library are released by the
 *     matching mbg_*_free;
 *   - device buffers passed in are owned by the caller;
 *   - a handle must not be used from two threads at once;
 *   - streams are hipStream_t passed as void* (NULL = the default stream).
 */
#ifndef MABOSS_GPU_H
#define MABOSS_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MBG_OK 0
#define MBG_ERR_ARG 1      /* invalid argument */
#define MBG_ERR_IO 2       /* file could not be read or written */
#define MBG_ERR_PARSE 3    /* .bnd/.cfg syntax error */
#define MBG_ERR_MODEL 4    /* unknown node, missing rate, undefined $variable, ... */
#define MBG_ERR_LIMIT 5    /* model or run exceeds a supported size */
#define MBG_ERR_COMPILE 6  /* hipRTC compile or link failed */
#define MBG_ERR_HIP 7      /* HIP runtime error (no device, out of memory, ...) */
#define MBG_ERR_OVERFLOW 8 /* a device hash table filled up */
#define MBG_ERR_INTERNAL 9

typedef struct mbg_model mbg_model;
typedef struct mbg_program mbg_program;
typedef struct mbg_result mbg_result;

/* Last error message of the calling thread ("" if none). */
const char* mbg_last_error(void);
/* Library version string. */
const char* mbg_version(void);

/* ---- model (replaces driver::parse, src/parser/driver.cpp:16-31, and the
 *      mask / istate / limit logic of src/main.cpp:15-36, :128-152) ---- */
int mbg_model_load(const char* bnd_path, const char* cfg_path, mbg_model** out);
int mbg_model_load_strings(const char* bnd_text, const char* cfg_text, mbg_model** out);
void mbg_model_free(mbg_model* m);

typedef struct mbg_model_info
{
	int32_t n_nodes;      /* N, in .bnd declaration order = bit order */
	int32_t n_words;      /* ceil(N / 32) 32-bit state words */
	int32_t n_observed;   /* non-internal nodes */
	int32_t windows;      /* ceil(max_time / time_tick) */
	int32_t discrete_time;
	float max_time;
	float time_tick;
	int64_t sample_count;
	uint64_t seed;        /* seed_pseudorandom, exact for integer literals */
	int32_t n_params;     /* $variables used by rate expressions */
	int32_t total_literals;
} mbg_model_info;

int mbg_model_get_info(const mbg_model* m, mbg_model_info* out);
/* Name of node i (valid while m lives). */
const char* mbg_model_node_name(const mbg_model* m, int32_t i);
/* 1 if node i is internal (src/parser/parse_types.cpp:20-22). */
int mbg_model_node_internal(const mbg_model* m, int32_t i);
/* Override a .cfg constant (sample_count, max_time, time_tick, discrete_time,
 * seed_pseudorandom, ...), as if written last in the .cfg. */
int mbg_model_set_constant(mbg_model* m, const char* name, double value);
/* Override a $variable ("$u_A" or "u_A"). */
int mbg_model_set_variable(mbg_model* m, const char* name, double value);
/* Name of rate parameter k (a $variable, with '$'). */
const char* mbg_model_param_name(const mbg_model* m, int32_t k);
/* Value of rate parameter k as the model defines it. */
float mbg_model_param_value(const mbg_model* m, int32_t k);
/* Evaluate node i's rate in a state with the host interpreter; used by tests. */
int mbg_model_eval_rate(const mbg_model* m, int32_t node, const uint32_t* state_words, float* rate);

/* ---- code generation + JIT (replaces generator::generate_code,
 *      src/generator.cpp:11-55, and kernel_compiler::compile_simulation,
 *      src/kernel_compiler.cpp:61-164) ---- */
typedef struct mbg_compile_opts
{
	int32_t entropy_formula; /* 0 = log-sum form (default), 1 = reference form */
	int32_t no_cache;        /* 1: bypass the code-object cache */
	int32_t with_gradient;   /* mbg_generate_code only: append the gradient kernel's
	                            model code and include (MBG_ERR_LIMIT if the model has none) */
	int32_t lane_occupancy;  /* waves per SIMD the lane kernels are compiled for
	                            (register budget 512 / waves per lane); 0 = automatic:
	                            4, 3 or 2 by node count, one fewer while the kernel
	                            would spill more than 4 registers (down to 2) */
	int32_t grad_occupancy;  /* waves per SIMD the gradient kernel is compiled for;
	                            0 = automatic: 2 up to 64 rate parameters, 1 above */
	int32_t reserved[3];
} mbg_compile_opts;

/* Generated HIP source of the model (caller frees with mbg_free_string). */
int mbg_generate_code(const mbg_model* m, const mbg_compile_opts* opts, char** source);
void mbg_free_string(char* s);
/* Generate, compile with hipRTC for gfx950 and link against the precompiled
 * kernels.  Needs a GPU (the link targets the current device). */
int mbg_program_compile(const mbg_model* m, const mbg_compile_opts* opts, mbg_program** out);
void mbg_program_free(mbg_program* p);
/* Seconds spent generating + compiling + linking (0 on a cache hit path is
 * not guaranteed; it includes the cache lookup). */
double mbg_program_compile_seconds(const mbg_program* p);

/* ---- simulation (replaces simulation_runner::run_simulation,
 *      src/simulation_runner.cu:54-180, and the stats_composite pipeline,
 *      src/statistics/stats_composite.cpp:3-30) ---- */
typedef struct mbg_sim_params
{
	int64_t n_trajectories;  /* <0: sample_count of the model */
	uint64_t first_id;       /* global id of the first trajectory (prefix runs) */
	int64_t seed;            /* <0: seed_pseudorandom of the model */
	uint32_t variant;        /* stream tag (batched variant index) */
	uint32_t max_steps;      /* 0: default cap on SSA steps per trajectory */
	int32_t window_mode;     /* -1 auto (dense LDS up to 32 KB of cells, dense HBM up to 8 GB,
	                            else hash), 0 dense HBM, 1 hash, 2 dense LDS, 3 sorted pairs */
	int32_t final_mode;      /* -1 auto, 0 dense HBM, 1 hash, 2 dense LDS */
	uint64_t window_capacity;/* hash slots (power of two) or sorted-mode pairs, 0 = auto */
	uint64_t final_capacity;
	uint64_t fixed_capacity;
	int32_t block_size;      /* 0 = 256 */
	int32_t blocks_per_cu;   /* 0 = occupancy limit */
	double normalization;    /* <=0: n_trajectories (reference divides by sample_count) */
	int32_t flags;           /* MBG_SIM_* bits */
	int32_t kernel;          /* -1 auto: lane up to 256 nodes; above, the wave kernel when it can
	                            run the model (fast-form rates, nodes of up to 6 inputs), else
	                            lane up to 512; 0 lane per trajectory, 1 wave per trajectory */
	const float* rates;      /* [n_params] run-time values of the rate parameters (mbg_model_param_name
	                            order), host memory; NULL: the model's.  Lane kernel only. */
	int32_t moment_states;   /* with MBG_SIM_ERRORS: observed states a lane holds per window in the
	                            second-moment replay, 8 (0) or 1; more states in a window go to a
	                            host-side list, so both give the same sums */
	int32_t refill_lanes;    /* lane kernel: idle lanes a wave gathers before it starts new
	                            trajectories (one round of initial-state draws serves up to
	                            64 / ceil(N/4)); 0: min(that, 8).  Changes no result */
	int32_t reserved[2];
} mbg_sim_params;

/* Keep the statistics in device memory and return only the counters and
 * timings (no device-to-host copy of the histograms); used to time the
 * device-side path. */
#define MBG_SIM_DEVICE_ONLY 1
/* Fill the ErrorTH / ErrorProba columns: after the run its trajectories are
 * replayed to sum, exactly, the square of each trajectory's time in each
 * window cell and of its entropy sum in each window; the columns are the
 * standard errors of the means, sqrt((E[x^2] - E[x]^2) / (n - 1)).  No
 * counterpart in the reference, which prints 0 in both columns
 * (src/statistics/window_average_small.cpp:235, :244-245).  Lane kernel
 * (up to 512 nodes) and windows x 2^n_obs <= 2^24; without the flag the
 * columns are 0, as in the reference. */
#define MBG_SIM_ERRORS 2

/* Fill p with the defaults above. */
void mbg_sim_params_default(mbg_sim_params* p);
/* Simulate and collect window, final-state and fixed-point statistics.
 * Inputs live in HBM for the whole run.  On the device the occupied cells of
 * every table are compacted and sorted by key; they are then copied to
 * pinned host memory, and the result is complete when this returns (the
 * stream is synchronised). */
int mbg_simulate(mbg_program* p, const mbg_sim_params* params, void* stream, mbg_result** out);
/* The same run, returning once it is enqueued: the statistics reach the host
 * on the program's own copy stream while the caller's stream moves on, so a
 * following run's kernels overlap this run's copy.  Counters (transitions,
 * finished, truncated) and the storage modes are final on return; the
 * statistics after mbg_result_wait, which every mbg_result_* accessor calls
 * first.  A program has two output slots: a third run in flight waits on the
 * host for the first one's copy.  No counterpart in the reference, whose
 * statistics copies are synchronous cudaMemcpy calls
 * (src/statistics/window_average_small.cpp:91-107, final_states.cpp:45-51). */
int mbg_simulate_async(mbg_program* p, const mbg_sim_params* params, void* stream, mbg_result** out);

/* The window probabilities of a run as a dense device array, for the
 * differentiable form: probs[w << n_observed | x] (device, doubles,
 * windows << n_observed entries) = P_w(x) normalised as by
 * mbg_result_window_probs; rates: the rate parameters in device memory
 * (float[n_params], mbg_model_param_name order; NULL: the model's values);
 * params->rates must be NULL.  Stream ordered: nothing goes through the host
 * and the host does not wait.  Lane kernel with dense window cells (at most
 * 24 observed nodes), else MBG_ERR_LIMIT.  No counterpart in the reference,
 * whose rates are compiled in (src/generator.cpp:104-106). */
int mbg_window_probs_device(mbg_program* p, const mbg_sim_params* params, const float* rates, double* probs,
							void* stream);

/* ---- batched variants.  No counterpart in the reference: its rates are
 *      baked into the generated code (src/generator.cpp:104-106), one model
 *      runs per process (src/main.cpp:154-164) and mutants are a TODO
 *      (README.md:33).  Variant v draws from the stream tagged v, so its
 *      result equals a plain run with variant = v and the same rates. ---- */
typedef struct mbg_variant
{
	int64_t n_trajectories;
	float max_time;           /* <= 0: the model's max_time (tick and windows follow) */
	int32_t pad;
	const float* rates;       /* [n_params] in mbg_model_param_name order, host memory; NULL: the model's */
	const int8_t* mutations;  /* [n_nodes]: +1 forced on (never switches off, starts on), -1 forced
	                             off, 0 free; host memory; NULL: none */
} mbg_variant;

typedef struct mbg_batch_params
{
	int64_t seed;             /* <0: seed_pseudorandom of the model */
	uint32_t max_steps;       /* 0: default cap */
	int32_t block_size;       /* 0: autotuned (64, 128 or 256) */
	int32_t blocks_per_cu;    /* 0: autotuned / occupancy */
	int32_t traj_per_lane;    /* a work chunk holds 64*traj_per_lane ids of one variant; 0: autotuned */
	int32_t use_graph;        /* 1: capture the launch in a hipGraph and replay it */
	int32_t flags;            /* MBG_SIM_DEVICE_ONLY */
	int32_t lds_cells;        /* window cells each wave caches in LDS (a power of two, 64-4096) before
	                             the HBM atomics, flushed when the wave's work chunk ends; 0: autotuned
	                             with the block size and trajectories per lane (none when both are
	                             given), -1: none */
	int32_t reserved[7];
} mbg_batch_params;

typedef struct mbg_batch_info
{
	double kernel_ms;         /* device time of the batched launch (graph or kernel) */
	double total_ms;          /* host wall time of mbg_simulate_batched */
	uint64_t transitions;     /* all variants */
	int32_t block_size, blocks_per_cu, traj_per_lane, used_graph;
	int32_t lds_cells;        /* the launch's LDS cell cache slots per wave (0: none) */
	int32_t tuned_cached;     /* 1: the launch shape came from the on-disk tuning cache, which keeps
	                             the autotuner's choice per generated model code, batch shape and
	                             device in $MBG_CACHE_DIR (else ~/.cache/maboss_gpu_amd) */
} mbg_batch_info;

void mbg_batch_params_default(mbg_batch_params* p);
/* Simulate n_variants variants in one launch.  results[v] receives variant
 * v's statistics (free each with mbg_result_free); info may be NULL.  The
 * autotuner times every block size (64, 128, 256) x trajectories per lane
 * (1, 2, 4, 6, 8, 12) x LDS cell cache (none, 256, 1024 slots) on the batch and keeps
 * the fastest; results do not depend on the choice (integer sums). */
int mbg_simulate_batched(mbg_program* p, const mbg_batch_params* params, int32_t n_variants,
						 const mbg_variant* variants, void* stream, mbg_result** results, mbg_batch_info* info);

/* ---- gradient form.  No counterpart in the reference (its rates are
 *      compile-time constants, src/generator.cpp:104-106, and nothing is
 *      differentiated).  Likelihood-ratio (score function) estimate of
 *      dL/dtheta for the model's rate parameters (mbg_model_param_name
 *      order), given dL/dP for the window probabilities P[w][x] of the
 *      forward run with the same trajectory ids, seed, stream tag and rates.
 *      Continuous time; every rate must be a bare $parameter in the fast
 *      form (`@logic ? $u : 0`, `@logic ? 0 : $d`) or parameter-free. ---- */
typedef struct mbg_grad_params
{
	int64_t n_trajectories;   /* <0: sample_count of the model */
	uint64_t first_id;        /* global id of the first trajectory */
	int64_t seed;             /* <0: seed_pseudorandom of the model */
	uint32_t variant;         /* stream tag of the forward run (batched variant index) */
	uint32_t max_steps;       /* 0: default cap */
	int32_t block_size;       /* 0: from the LDS budget (64..256) */
	int32_t blocks_per_cu;    /* 0: occupancy limit */
	int32_t fold_lanes;       /* ended trajectories a wave gathers before folding them into the
	                             block's accumulators (1-64); 0: 16.  Changes no result beyond
	                             the order of double additions */
	int32_t reserved[5];
} mbg_grad_params;

typedef struct mbg_grad_info
{
	double kernel_ms;         /* device time of the gradient kernel */
	uint64_t transitions;     /* SSA steps replayed */
	int32_t block_size, grid;
} mbg_grad_info;

void mbg_grad_params_default(mbg_grad_params* p);
/* MBG_OK when the model has a gradient form, else the reason (mbg_last_error). */
int mbg_grad_check(mbg_program* p);
/* grad_params[j] = sum_{w,x} grad_probs[w << n_obs | x] * dP[w][x]/dtheta_j.
 * rates: device [n_params] floats (NULL: the model's values); grad_probs:
 * device [windows << n_observed] doubles; grad_params: device [n_params]
 * doubles, overwritten.  Stream ordered; when info is non-NULL the stream is
 * synchronised and info filled.  Calls on one program share its workspace:
 * issue them on one stream. */
int mbg_backward(mbg_program* p, const mbg_grad_params* params, const float* rates, const double* grad_probs,
				 double* grad_params, void* stream, mbg_grad_info* info);

/* mbg_backward that allocates nothing.  No counterpart in the reference: it
 * is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments and result as mbg_backward without `info`; the kernel's
 * workspace (counters and model tables) comes from a static device array of
 * MBG_NOALLOC_SCRATCH_BYTES in this library, of which each GPU holds one
 * copy for the life of the process.  Call mbg_grad_check on the program
 * first (it loads the gradient kernel); MBG_ERR_ARG otherwise, and
 * MBG_ERR_LIMIT if the workspace would exceed the scratch.  Every call on
 * one GPU shares that GPU's array: calls that overlap on one GPU (on
 * different streams or threads) race. */
#define MBG_NOALLOC_SCRATCH_BYTES 16384
int mbg_backward_noalloc(mbg_program* p, const mbg_grad_params* params, const float* rates,
						 const double* grad_probs, double* grad_params, void* stream);

/* ---- results (replaces stats::finalize / visualize / write_csv,
 *      src/statistics/stats.h:15-29) ---- */
void mbg_result_free(mbg_result* r);
/* Block until the result's statistics are on the host (a no-op for results
 * of mbg_simulate and after the first call). */
int mbg_result_wait(mbg_result* r);

typedef struct mbg_result_info
{
	int32_t windows;
	int32_t n_observed;
	int32_t n_words;
	int32_t discrete_time;
	float time_tick;
	float max_time;
	double normalization;            /* trajectories the probabilities divide by */
	uint64_t n_trajectories;
	uint64_t transitions;            /* SSA iterations (records) */
	uint64_t finished;               /* trajectories that reached max_time or a fixed point */
	uint64_t truncated;              /* trajectories stopped by the step cap */
	uint64_t n_fixed_points;         /* distinct fixed points */
	uint64_t n_final_states;         /* distinct final observed states */
	uint64_t n_window_cells;         /* nonzero (window, state) cells */
	double kernel_ms;                /* device time of the whole device path: simulate kernel,
	                                    tail expansion, compaction and sort of the occupied cells
	                                    (with MBG_SIM_DEVICE_ONLY: simulate and tail expansion) */
	double total_ms;                 /* host wall time from the mbg_simulate(_async) call until the
	                                    statistics were on the host */
	int32_t window_mode, final_mode;
	uint64_t window_capacity;
	int32_t kernel;                  /* 0 lane per trajectory, 1 wave per trajectory */
	int32_t pad;
	double simulate_ms;              /* device time of the simulate kernel alone */
	uint64_t window_pairs;           /* sorted window mode: pair slots filled by the kernels */
	uint64_t moment_spills;          /* MBG_SIM_ERRORS: partial times the second-moment replay
	                                    summed on the host (windows with more states than a lane holds) */
	int32_t has_errors;              /* 1: the error columns come from second moments */
	int32_t pad2;
	uint64_t window_spills;          /* hashed window mode: cells whose probe passed 256 slots, kept in a
	                                    spill list and summed with the table's cells after the run */
	uint32_t window_record_bytes;    /* bytes per window cell copied to the host: 12 when the cells
	                                    travel packed (key and value in one 96-bit record), else 16 */
	int32_t pad3;
	uint64_t window_escapes;         /* packed cells whose value did not fit its record and was copied
	                                    in a side list */
} mbg_result_info;

int mbg_result_get_info(const mbg_result* r, mbg_result_info* out);
/* Number of nonzero observed states in window w. */
int64_t mbg_result_window_size(const mbg_result* r, int32_t w);
/* Window w: observed-state indices (bit j = j-th non-internal node) and raw
 * fixed-point dwell sums (continuous: dwell*2^32/tick, discrete: counts),
 * sorted by index.  Arrays hold mbg_result_window_size(r, w) entries. */
int mbg_result_window(const mbg_result* r, int32_t w, uint64_t* states, uint64_t* raw);
/* Window w: probabilities (as the reference normalises them). */
int mbg_result_window_probs(const mbg_result* r, int32_t w, uint64_t* states, double* probs);
/* Per-window transition entropy TH and state entropy H. */
int mbg_result_entropies(const mbg_result* r, double* th, double* h);
/* Standard errors (runs with MBG_SIM_ERRORS; zeros otherwise): ErrorTH per
 * window, and ErrorProba of window w's cells in mbg_result_window's order. */
int mbg_result_error_th(const mbg_result* r, double* err_th);
int mbg_result_window_errors(const mbg_result* r, int32_t w, double* err);
/* Final observed states and counts, sorted by index. */
int64_t mbg_result_final_size(const mbg_result* r);
int mbg_result_final(const mbg_result* r, uint64_t* states, uint64_t* counts);
/* Fixed points: n_words words per state, in the reference's order (most
 * significant word first, src/state.h:43-49), and their counts. */
int64_t mbg_result_fixed_size(const mbg_result* r);
int mbg_result_fixed(const mbg_result* r, uint32_t* states, uint64_t* counts);

/* <prefix>_probtraj.csv and <prefix>_fp.csv (src/statistics/window_average_small.cpp:181-250,
 * src/statistics/fixed_states.cu:154-187); with_finals adds <prefix>_finalstates.csv. */
int mbg_result_write_csv(const mbg_result* r, const char* prefix, int with_finals);
/* The stdout report of the reference (final points, fixed points, windows)
 * to a file path, or to stdout when path is NULL
 * (src/statistics/final_states.cpp:53-66, fixed_states.cu:140-152,
 * window_average_small.cpp:142-179). */
int mbg_result_print(const mbg_result* r, const char* path);
/* Exact integer dump used by the parity tests (one record per line). */
int mbg_result_write_raw(const mbg_result* r, const char* path);

/* ---- checkpoint / resume at trajectory-id granularity.  The reference's
 *      runner keeps its state only between epochs of one process
 *      (src/simulation_runner.cu:123-161) and saves nothing.  A result's
 *      statistics are exact integers and trajectory id i always draws the
 *      same stream, so a run split into id ranges and added back together
 *      equals the single run bit for bit. ---- */
/* The result's statistics (window cells, entropy sums, final states, fixed
 * points, second moments) and identity (model and rate fingerprint, seed,
 * stream tag, trajectory ids, kernel) to a file, replaced atomically; the
 * file ends with a hash of its contents. */
int mbg_result_save(const mbg_result* r, const char* path);
/* A saved result, for a program of the same model (MBG_E_IO for a damaged or
 * truncated file, MBG_E_ARG for another model's). */
int mbg_result_load(const mbg_program* p, const char* path, mbg_result** out);
/* Add `other` into `into`: the same run identity and adjacent trajectory-id
 * ranges (MBG_E_ARG otherwise); `into` then covers both. */
int mbg_result_merge(mbg_result* into, const mbg_result* other);
/* Simulate ids [first_id, first_id + n_trajectories) in chunks of `chunk`
 * ids (0: one chunk), saving the merged statistics to `path` after each
 * chunk.  When `path` already holds a checkpoint of the same run identity
 * and first id, the run resumes after its last id (a checkpoint that covers
 * every id returns at once).  Honours MBG_SIM_ERRORS; not
 * MBG_SIM_DEVICE_ONLY. */
int mbg_simulate_checkpointed(mbg_program* p, const mbg_sim_params* params, uint64_t chunk, const char* path,
							  void* stream, mbg_result** out);

#ifdef __cplusplus
}
#endif

#endif /* MABOSS_GPU_H */
