// Argument block of the fused simulate kernels; shared by the host runner and
// the device code (plain C layout, passed by value as the only kernel argument).
#pragma once

#include "mbg_abi.h"

// Histogram storage modes.
#define MBG_HIST_DENSE 0  // cell (w << n_obs) | obs, in HBM
#define MBG_HIST_HASH 1   // open-addressing table keyed by (w << n_obs) | obs
#define MBG_HIST_LDS 2    // dense, privatised in LDS and flushed once per block
#define MBG_HIST_SORT 3   // (key, value) pairs appended, then sorted and reduced by key

// Counter slots (u64) written by the kernels.
#define MBG_CTR_TRANSITIONS 0    // SSA iterations = records of the reference
#define MBG_CTR_WIN_OVERFLOW 1   // window hash table full
#define MBG_CTR_FIN_OVERFLOW 2   // final-state hash table full
#define MBG_CTR_FP_OVERFLOW 3    // fixed-point table full
#define MBG_CTR_TRUNCATED 4      // trajectories stopped by the step cap
#define MBG_CTR_FP_TIMEOUT 5     // fixed-point slot never published (bounded wait ended)
#define MBG_CTR_QUEUE 6          // trajectory work queue
#define MBG_CTR_FINISHED 7       // trajectories ended at max_time or a fixed point
#define MBG_CTR_TAIL_OVERFLOW 8  // tail table full
#define MBG_CTR_PAIRS 9          // window pair slots reserved (MBG_HIST_SORT; may exceed the capacity)
#define MBG_CTR_WIN_SPILL 10     // window cells sent to the spill list (MBG_HIST_HASH; may exceed the capacity)
#define MBG_PAIR_PAD_KEY 0xffffffffffffffffull // key of reserved slots left unused
#define MBG_NCOUNTERS 11

#define MBG_EMPTY_KEY 0xffffffffffffffffull

// Node of the wave-per-trajectory kernel: logic as a truth table over at
// most MBG_WAVE_MAXPAR parent nodes, rates of the fast form
// rate = (state == logic) ? 0 : (state ? down : up).
#define MBG_WAVE_MAXPAR 6
#define MBG_WAVE_NOOBS 0xffffu
struct mbg_wnode
{
	unsigned short par[MBG_WAVE_MAXPAR]; // parent node indices (unused: 0)
	unsigned char npar;
	unsigned char pad;
	unsigned short obs_pos; // bit of the observed index, or MBG_WAVE_NOOBS
	unsigned long long tt;  // bit j = logic value for parent bits j
	float up, down;
};

// Tables of the wave kernel, in device memory; the kernel stages them in LDS.
struct mbg_wave_tables
{
	const mbg_wnode* nodes;        // [n_nodes]
	const unsigned int* child_off; // [n_nodes + 1] CSR offsets into child
	const unsigned short* child;   // nodes whose rate depends on a node
	const unsigned short* obs;     // [n_obs] observed nodes in order
	int n_child, npl;              // entries of child; nodes per lane
	int tab_bytes;                 // LDS bytes of the staged tables (16-aligned)
	int obs_tab_off;               // byte offset of the observed nodes' {up, down, ul, dl}
};

struct mbg_sim_args
{
	unsigned long long n_traj;  // trajectories in this launch
	unsigned long long id_base; // global id of the launch's first trajectory
	unsigned int seed_lo, seed_hi, variant, max_steps;
	int n_nodes, n_obs, windows, discrete;
	float max_time, tick;
	float qscale;   // 2^32 / tick (continuous) or 2^32 (discrete entropy)
	float pad0;
	const float* istate; // [n_nodes] probability that node starts on

	int win_mode, fin_mode, ent_lds, lds_bytes; // storage modes, LDS size in bytes
	int win_lds_off, fin_lds_off, ent_lds_off; // u64 offsets into dynamic LDS
	int own_lds_off; // u64 offset of 64 bytes per wave (the lane that owns each item of a
	                 // cooperative pass), then the block's 3 u64 counters (lane kernel)

	unsigned long long* win_vals;   // dense cells (MBG_HIST_DENSE)
	unsigned long long* win_table;  // hash slots {key, value} (MBG_HIST_HASH), 16 B each
	unsigned long long win_mask;    // hash capacity - 1
	unsigned long long* win_ent;    // [windows] entropy sums
	unsigned long long* fin_vals;
	unsigned long long* fin_table;
	unsigned long long fin_mask;
	unsigned int* fp_table;         // slots of (2 + W) u32: tag, count, state words
	unsigned long long fp_mask;
	// Tail records: the final record of a trajectory adds the same quantum
	// to every window after its first one, so those windows are counted per
	// (first full window, observed state) here and expanded after the run.
	unsigned long long* tail_table; // hash slots {key, count}; null: per-window adds
	unsigned long long tail_mask;
	// Block-private LDS caches in front of the tail and fixed-point tables:
	// trajectories of a block that end in the same state (attractors) are
	// counted in LDS and reach the global table once per block.
	int tail_lds_off, tail_lds_slots; // u64 offset, slots {key + 1, count} (power of two; 0: none)
	int fp_lds_off, fp_lds_slots;     // u64 offset, slots of (2 + key words) u32 (power of two; 0: none)
	int fin_lds_cache_off, fin_lds_cache_slots; // hashed final states: u64 offset, slots {key + 1, count}
	unsigned long long* counters;   // [MBG_NCOUNTERS]
	// MBG_HIST_HASH: cells whose probe ran past MBG_WIN_PROBES slots, appended
	// as (key, value) pairs and sum-reduced with the table's cells after the
	// run, so a table that fills up loses nothing and needs no second run
	int refill_lanes;               // lane kernel: idle lanes a wave gathers before it refills them
	int pad_refill;
	unsigned long long* win_spill_keys;
	unsigned long long* win_spill_vals;
	unsigned long long win_spill_cap;
	unsigned long long* pair_keys;  // MBG_HIST_SORT: appended cell keys
	unsigned long long* pair_vals;  //                and their values
	unsigned long long pair_cap;    //                capacity in pairs
	int pair_lds_off;               //                u64 offset of the per-wave chunk cursors in LDS
	int pair_chunk;                 //                pairs a wave reserves at a time
	int pair_stage_off;             //                u64 offset of the per-wave staging buffers in LDS
	int pair_stage;                 //                pairs a wave stages (0: write directly)
	const mbg_param_view* pv;       // per-variant tables (batched/gradient); null for plain
	mbg_wave_tables wt;             // wave-per-trajectory kernel only
	int wave_lds_off;               // byte offset of the per-wave rate arrays in LDS
	int bent_windows;               // batched form: per-wave LDS entropy sums, windows (0: global atomics)

	// batched form only
	const struct mbg_batch_variant* bv;  // [n_variants]
	const struct mbg_batch_chunk* chunks; // work list, each chunk inside one variant
	unsigned int n_chunks;
	unsigned int bcache_slots;           // per-wave LDS window-cell cache slots (power of two; 0: none)
	unsigned long long* vctr;            // per-variant {transitions, finished, truncated}
};

// One variant of the batched form: its own rates, mutant masks, initial
// probabilities, max_time / windows and slices of the statistics storage.
struct mbg_batch_variant
{
	unsigned long long n_traj;
	unsigned long long win_off; // first window cell of [windows << n_obs]
	unsigned long long fin_off; // first final-state cell of [2^n_obs]
	unsigned int ent_off;       // first entropy sum of [windows]
	int windows;
	float max_time;
	unsigned int ctr_off; // first of its 3 counters in vctr
	const MBG_CONST float* istate;  // [n_nodes]
	mbg_param_view pv;
};

// A run of trajectory ids [begin, end) of one variant, taken by one wave.
struct mbg_batch_chunk
{
	unsigned int variant, pad;
	unsigned long long begin, end;
};

// Likelihood-ratio gradient kernel (no counterpart in the reference).  One
// lane per trajectory replays the forward trajectory (same stream, same
// float arithmetic) and accumulates, per rate parameter, the score terms of
// sum_w Gw * S(e_w) in the suffix form described in mbg_grad_kernel.h.
// Second-moment replay (no counterpart in the reference): the run's
// trajectories again, squaring each trajectory's per-window totals.
#define MBG_MOM_CTR_QUEUE 0 // trajectory work queue
#define MBG_MOM_CTR_SPILL 1 // spill records written (may exceed the capacity)
#define MBG_MOM_NCOUNTERS 2

struct mbg_moments_args
{
	unsigned long long n_traj, id_base;
	unsigned int seed_lo, seed_hi, variant, max_steps;
	int n_nodes, n_obs, windows, discrete;
	float max_time, tick, qscale;
	int refill_lanes;               // idle lanes a wave gathers before it starts new trajectories
	const float* istate;
	mbg_param_view pv;              // rates (run-time values) and rate * log2(rate), masks all enabled
	unsigned long long* sq;         // [windows << n_obs][2]: sum of P^2, low and high words
	unsigned long long* esq;        // [windows][2]: sum of T^2
	unsigned long long* spill;      // [spill_cap][4]: id, window, obs, partial P
	unsigned long long spill_cap;
	unsigned long long* counters;   // [MBG_MOM_NCOUNTERS]
};

struct mbg_grad_args
{
	unsigned long long n_traj, id_base;
	unsigned int seed_lo, seed_hi, variant, max_steps;
	int n_nodes, n_obs, windows, n_params;
	float max_time, tick;
	int g_lds;         // 1: the upstream gradient table is staged in LDS
	int lds_bytes;     // dynamic LDS of a block
	int fold_lanes;    // ended trajectories a wave gathers before folding them
	int pad_fold;
	const float* istate;
	mbg_param_view pv; // rates (runtime values), masks all enabled
	const double* gprob;           // [windows << n_obs] dL/dP
	const unsigned short* jparam;  // [2 * n_nodes]: parameter of node i's up (2i) / down (2i+1) jump, 0xffff none
	double* grad;                  // [n_params], accumulated with atomics (zeroed by the host)
	unsigned long long* counters;  // [MBG_NCOUNTERS]
	double scale;                  // 1 / (n_traj_total * tick)
};
