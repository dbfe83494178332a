// Host engine: kernel compiler (hipRTC), simulation runner and results.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <functional>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../device/mbg_kernel_args.h"
#include "generator.h"
#include "model.h"

namespace mbg
{

// Throws mbg::error(MBG_E_HIP) with the HIP error string.
void hip_check(hipError_t e, const char* what);

// Tables of the wave-per-trajectory kernel (built from the model on the host).
struct wave_tables
{
	int npl = 1;         // nodes per lane
	int rows = 1;        // rows of a wave's rate array: npl rounded up to a power of two
	int obs_tab_off = 0; // byte offset of the observed nodes' {up, down, ul, dl} in image
	std::vector<mbg_wnode> nodes;
	std::vector<unsigned> child_off;
	std::vector<unsigned short> child;
	std::vector<unsigned short> obs;
	std::vector<char> image; // nodes, child_off, child, obs, slots, observed rates; 16-byte padded
};

// Throws MBG_E_LIMIT with the reason when the model does not fit the wave kernel.
wave_tables build_wave_tables(const model& m);

// Device allocation that only grows; reused by successive runs of a program.
struct device_buffer
{
	void* ptr = nullptr;
	size_t bytes = 0;
	void* ensure(size_t n);
	void release();
	~device_buffer() { release(); }
};

// Pinned (page-locked) host blocks for statistics copied back from the
// device, kept for reuse: pinning hundreds of MB costs far more than the copy
// itself, so a block released by a result goes back to the pool and the next
// run of the same size takes it.  Thread safe.
class pinned_pool : public std::enable_shared_from_this<pinned_pool>
{
public:
	~pinned_pool();
	// A block of at least `bytes` bytes; the block returns to the pool when
	// the last reference goes.
	std::shared_ptr<char> get(size_t bytes);

private:
	void put(char* p, size_t bytes);
	std::mutex mu_;
	std::vector<std::pair<size_t, char*>> free_;
	size_t retained_ = 0;
};

// A sorted run of (key, value) cells in host memory: window cells keyed by
// (window << n_obs) | observed state, or final states keyed by the observed
// state.  `hold` keeps the memory alive (a pinned block shared by the results
// of one launch, or a vector); `sub` is subtracted from every stored key (the
// first cell of a variant in a batched launch).
struct cell_view
{
	std::shared_ptr<const void> hold;
	// plain form: keys and values in two arrays
	const uint64_t* key = nullptr;
	const uint64_t* val = nullptr;
	// packed form (rec set): 12-byte records of three 32-bit words, the
	// 96-bit integer key | field << kbits.  With esc_min = 0 the field is the
	// value; otherwise the field has 96 - kbits bits, and a field f >= esc_min
	// stands for the value xval[f - esc_min] of the side list
	const uint32_t* rec = nullptr;
	unsigned kbits = 0;
	uint64_t esc_min = 0;
	const uint64_t* xval = nullptr;
	uint64_t nx = 0;
	uint64_t first = 0; // packed form: record index of this view's cell 0
	uint64_t n = 0;
	uint64_t sub = 0;
	uint64_t k(uint64_t i) const { return (rec ? packed_key(first + i) : key[i]) - sub; }
	uint64_t v(uint64_t i) const { return rec ? packed_val(first + i) : val[i]; }
	// index of the first cell whose key is >= kk
	uint64_t lower_bound(uint64_t kk) const;
	// cells [lo, hi) of this view, their keys reduced by s
	cell_view slice(uint64_t lo, uint64_t hi, uint64_t s) const;
	// cells from owned vectors
	static cell_view from(std::vector<uint64_t> keys, std::vector<uint64_t> vals);

private:
	uint64_t packed_lo(uint64_t r) const
	{
		return (uint64_t)rec[3 * r] | ((uint64_t)rec[3 * r + 1] << 32);
	}
	uint64_t packed_key(uint64_t r) const { return packed_lo(r) & ((1ull << kbits) - 1ull); }
	uint64_t packed_val(uint64_t r) const;
};

// Tables of the gradient kernel (enable masks, initial-state probabilities,
// jump parameters, the model's rates) in pinned host memory, uploaded at the
// start of each gradient launch.
struct grad_host_tables
{
	std::shared_ptr<char> host;
	size_t en_bytes = 0, ist_bytes = 0, jp_bytes = 0, bytes = 0;
};

// A buffer a run left cleared (its tables are cleared right after they have
// been read, before the run's copy to the host starts, so the next run finds
// them ready and does not clear them while that copy is in flight).
struct cleared_buf
{
	const void* ptr = nullptr;
	size_t bytes = 0;
	bool holds(const void* p, size_t b) const { return p && ptr == p && bytes >= b; }
};

// Device storage a run's statistics leave the device from.  A program keeps
// two, used in turn, so that one run's copy to the host (on the program's
// copy stream) overlaps the next run's kernels; a slot is reused only after
// the copy of its previous run has completed.
struct out_slot
{
	// window cells: compacted (hashed / dense modes) or appended (sorted
	// mode) pairs, the sort's alternate buffers
	device_buffer pk, pv, pka, pva;
	device_buffer ent;            // [windows] entropy sums
	device_buffer fk, fv, fka, fva; // final states, compacted
	device_buffer fpc;            // fixed points, compacted records {count, words}
	device_buffer cnt;            // [4] counts of the compactions
	device_buffer prec, px;       // window cells packed for the copy: records; escape count + side list
	cleared_buf cl_ent, cl_cnt;
	hipEvent_t copied = nullptr;  // the last copy out of this slot
	bool pending = false;         // copied has been recorded
	~out_slot();
};

// The reference's kernel_compiler (src/kernel_compiler.{h,cpp}) plus the run
// workspace: generated source -> hipRTC -> code object -> module and kernel
// handles, and the device storage reused by successive runs.
struct program
{
	// run workspace (kept between runs so repeated launches do not reallocate)
	device_buffer b_counters, b_win_vals, b_win_keys, b_fin_vals, b_fin_keys, b_fp, b_istate, b_pv, b_wave;
	device_buffer b_sort_tmp;
	cleared_buf cl_counters, cl_win_keys, cl_win_vals, cl_fin_keys, cl_fin_vals, cl_tail, cl_fp, cl_pent;
	const void* istate_dev = nullptr; // b_istate holds the model's initial-state probabilities
	const void* wave_dev = nullptr;   // b_wave holds the wave kernel's tables
	uint64_t pairs_per_traj = 0; // learned from the last sorted run (capacity hint)
	out_slot slots[2];
	int slot_next = 0;
	hipStream_t copy_stream = nullptr; // device-to-host copies of the statistics
	std::shared_ptr<pinned_pool> pool = std::make_shared<pinned_pool>();
	std::shared_ptr<char> mailbox;     // small pinned block: counters and counts read mid-run
	// recorded after the last device work a run enqueued on its stream; the
	// next run (which may come on another stream) waits for it, since both
	// use the shared workspace above
	hipEvent_t run_done = nullptr;
	void order_after_last_run(hipStream_t stream);
	void mark_run_done(hipStream_t stream);

	std::shared_ptr<const model> mdl;
	generated gen;
	gen_options opts;
	int W = 1;
	// lane kernels per flavour (0 plain, 1 batched, 2 plain with run-time
	// rates); the plain flavours also per window and final-state storage mode,
	// which they take as compile-time constants
	hipModule_t mods[3][4][3] = {};
	hipFunction_t fns[3][4][3] = {};
	double compile_seconds = 0;
	bool no_cache = false;
	int device = 0;
	int num_cus = 0;

	hipFunction_t fn_fill = nullptr; // mbg_fill_u64 of the parametric module

	// utility kernels (slot fill, tail expansion, table compaction), precompiled
	hipModule_t mod_util = nullptr;
	hipFunction_t fn_fill_slots = nullptr, fn_expand_tails = nullptr, fn_table_compact = nullptr,
				  fn_fp_compact = nullptr, fn_rate_view = nullptr, fn_cells_to_probs = nullptr, fn_pack_cells = nullptr;
	void load_util();
	// waits for every copy in flight (before buffers are released)
	void drain_copies();
	device_buffer b_tail;
	device_buffer b_win_spill; // hashed window mode: spilled cells, keys then values

	// batched form: launch shape chosen by the autotuner, keyed by the batch
	struct tuned
	{
		uint64_t key = 0;
		int block = 0, per_cu = 0, tpl = 0, slots = 0;
		bool cached = false;
	} batch_tuned;
	// captured batched launch, replayed while the batch layout is unchanged
	hipGraphExec_t batch_graph = nullptr;
	uint64_t batch_graph_key = 0;
	device_buffer b_bwin, b_bent, b_bfin, b_bfp, b_bctr, b_bvar, b_bchunk, b_btab;
	device_buffer b_bck, b_bcv, b_bfk, b_bfv, b_bfpc, b_bcnt; // compacted cells of a batched launch

	// likelihood-ratio gradient kernel (compiled per model on first use)
	std::unique_ptr<grad_code> gcode;
	hipModule_t mod_grad = nullptr;
	hipFunction_t fn_grad = nullptr;
	const grad_code& grad();          // throws MBG_E_LIMIT when the model has no gradient form
	hipFunction_t grad_function();
	device_buffer b_gen;                 // gradient workspace: counters, tables

	// second-moment replay (compiled per model and per states-per-lane on first use)
	hipModule_t mod_mom[2] = {};
	hipFunction_t fn_mom[2] = {};
	hipFunction_t moments_function(int states_per_lane);
	device_buffer b_mom;
	grad_host_tables gtab;                    // the gradient kernel's tables, built on first use
	const grad_host_tables& grad_tables();
	size_t lds_per_block = 0;            // device property, read once
	hipFunction_t fn_fill_u64 = nullptr; // utility module's fill (graph safe)

	// wave-per-trajectory kernel (precompiled, table driven)
	hipModule_t mod_wave = nullptr;
	std::unique_ptr<wave_tables> wtab;
	std::string wave_unsupported; // reason, when the model does not fit the wave kernel

	~program();
	// win_mode / fin_mode: MBG_HIST_* of the run (ignored for flavour 1)
	hipFunction_t function(int flavour, int win_mode, int fin_mode);
	device_buffer b_rate_view; // run-time rates of the plain flavour 2
	device_buffer b_pent;      // entropy sums of a device-probability run
	// initial-state probabilities in pinned memory (their upload is stream ordered)
	std::shared_ptr<char> istate_pinned;
	// Kernel for the wave path at the model's nodes-per-lane (loads on first use).
	hipFunction_t wave_function();
};

std::unique_ptr<program> compile_program(std::shared_ptr<const model> m, const gen_options& opt, bool no_cache);
// The memory budget of a run's tables: a quarter of the device memory free
// when it starts (8 GiB when the runtime cannot tell).
uint64_t device_budget();
// Idle lanes a lane-kernel wave gathers before it starts new trajectories.
int default_refill_lanes(int n_nodes);
// Directory of the on-disk caches (code objects, tuned launch shapes):
// $MBG_CACHE_DIR, else ~/.cache/maboss_gpu_amd.
std::string code_cache_dir();

// Largest state-word count served by the lane-per-trajectory kernels (512
// nodes, the reference's limit, CMakeLists.txt:12-13); above auto_lane_words
// the automatic choice is the wave kernel when it can run the model (its
// table-driven evaluator covers fast-form rates of nodes with up to 6
// inputs), else the lane kernel with the model's generated code.  The
// batched and gradient forms keep the 256-node limit.
constexpr int max_lane_words = 16;
constexpr int auto_lane_words = 8;
constexpr int max_form_words = 8;

struct sim_params
{
	int64_t n_traj = -1;
	uint64_t first_id = 0;
	int64_t seed = -1;
	uint32_t variant = 0;
	uint32_t max_steps = 0;
	int window_mode = -1, final_mode = -1;
	uint64_t window_capacity = 0, final_capacity = 0, fixed_capacity = 0;
	int block_size = 0, blocks_per_cu = 0;
	double normalization = 0;
	int kernel = -1; // -1 auto, 0 lane per trajectory, 1 wave per trajectory
	bool device_only = false; // leave the histograms in HBM
	std::vector<float> rates;  // [n_params] run-time rate values; empty: the model's
	int refill_lanes = 0;      // lane kernel: idle lanes a wave gathers before refilling (0: automatic)
	bool errors = false;       // second moments for ErrorTH / ErrorProba (a replay after the run)
	int moment_states = 0;     // observed states a lane holds per window in that replay (0: 8)
};

struct result
{
	// run description
	int windows = 0, n_obs = 0, n_words = 0, discrete = 0;
	float tick = 0, max_time = 0;
	double norm = 0;
	uint64_t n_traj = 0;
	// identity of the run, for checkpoints: trajectories [first_id, first_id
	// + n_traj) of the stream (seed, variant) of the model and rates whose
	// fingerprint this is; user_norm: the probabilities divide by norm, not n_traj
	uint64_t first_id = 0, seed = 0, fingerprint = 0;
	uint32_t variant = 0;
	bool user_norm = false;
	std::vector<std::string> names;
	std::vector<int> observed;

	// statistics; raw values are fixed point (see simulate_lane.hip).  Valid
	// after wait() (a run may still be copying them to the host).
	cell_view win;                  // window cells sorted by (window << n_obs) | state
	std::vector<uint64_t> woff;     // [windows + 1]: window w's cells are [woff[w], woff[w + 1])
	std::vector<uint64_t> ent;      // per window
	cell_view fin;                  // final states sorted by state
	std::vector<uint32_t> fp_words; // fixed points, n_words words each, reference order
	std::vector<uint64_t> fp_count;

	uint64_t transitions = 0, finished = 0, truncated = 0;
	double kernel_ms = 0, total_ms = 0, simulate_ms = 0;
	uint64_t window_pairs = 0;
	uint64_t window_spills = 0; // hashed window cells that went through the spill list
	uint32_t window_record_bytes = 16; // bytes per window cell copied to the host (12 packed, 16 plain)
	uint64_t window_escapes = 0;       // packed cells whose value went through the side list
	int window_mode = 0, final_mode = 0;
	int kernel = 0; // 0 lane, 1 wave
	uint64_t window_capacity = 0;

	// Second moments (sim_params::errors): per window cell and per window,
	// the exact sum over trajectories of the square of the trajectory's
	// fixed-point time in the cell, and of its entropy sum in the window.
	bool has_moments = false;
	std::vector<unsigned __int128> msq;  // [windows << n_obs], dense
	std::vector<unsigned __int128> mesq; // [windows]
	uint64_t moment_spills = 0;          // partial times the replay sent to the host
	double error_th(int w) const;                       // standard error of TH (0 without moments)
	double error_prob(int w, uint64_t obs, uint64_t raw) const; // standard error of a cell's probability

	double prob(uint64_t raw) const;     // window cell -> probability
	double th(int w) const;              // transition entropy of window w
	double h(int w) const;               // state entropy of window w (float, as the reference)
	std::string state_name(uint64_t obs) const;
	std::string full_state_name(const uint32_t* s) const;
	uint64_t n_cells() const { return woff.empty() ? 0 : woff.back() - woff.front(); }
	uint64_t n_fixed() const { return fp_count.size(); }
	const uint32_t* fixed_state(uint64_t i) const { return fp_words.data() + i * (size_t)n_words; }
	uint64_t obs_mask() const { return n_obs >= 64 ? ~0ull : ((1ull << n_obs) - 1ull); }
	// The cells, window offsets and fixed-point order from what the run left
	// in the views (windows from the cell keys; fixed points sorted most
	// significant word first, src/state.h:43-49).
	void index_cells();
	void sort_fixed();

	void write_csv(const std::string& prefix, bool with_finals) const;
	void print(std::ostream& os) const;
	void write_raw(std::ostream& os) const;

	// Completion of a run whose copy is still in flight: wait() blocks until
	// it has landed and then indexes it.  Safe to call more than once.
	void wait() const;
	mutable std::function<void()> on_ready;
	~result();
};

std::unique_ptr<result> simulate(program& p, const sim_params& sp, hipStream_t stream);

// Checkpoints (trajectory-id granularity): a result's exact statistics and
// identity to a file and back; two results of adjacent id ranges of the same
// run added into one; and a run in chunks of ids that saves the merged
// statistics after each chunk and, started again with the same file,
// resumes after its last id.
uint64_t run_fingerprint(const program& p, const sim_params& sp);
void save_result(const result& r, const std::string& path);
std::unique_ptr<result> load_result(const program& p, const std::string& path);
void merge_into(result& into, const result& other);
std::unique_ptr<result> simulate_checkpointed(program& p, const sim_params& sp, uint64_t chunk,
											  const std::string& path, hipStream_t stream);
// The second moments of a finished run (sim_params::errors): replays its
// trajectories with the moments kernel and fills r.msq / r.mesq.  Lane
// kernel, dense windows (windows << n_obs <= 2^24).
void run_moments(program& p, const sim_params& sp, hipStream_t stream, result& r);
// As simulate(), but returns once the run is enqueued: the statistics are
// copied to pinned host memory on the program's copy stream and the result
// completes in wait().
std::unique_ptr<result> simulate_async(program& p, const sim_params& sp, hipStream_t stream);
// The window probabilities P[w][x] of a run as dense doubles in device memory
// (probs: [windows << n_obs]), with the rate parameters read from device
// memory (rates: [n_params]; null: the model's).  Stream ordered: nothing is
// copied to or from the host and the host does not wait.  Lane kernel,
// dense window storage (at most 24 observed nodes).
void window_probs_device(program& p, const sim_params& sp, const float* rates, double* probs, hipStream_t stream);
// Window and final-state storage modes a default run of the model uses.
void default_storage_modes(const model& m, int& win_mode, int& fin_mode);

// ---- batched form
struct variant_spec
{
	int64_t n_traj = 0;
	float max_time = 0;         // <= 0: the model's
	std::vector<float> rates;   // [n_params] or empty: the model's values
	std::vector<int8_t> mut;    // [n_nodes]: +1 forced on, -1 forced off, 0 free; or empty
};

struct batch_params
{
	int64_t seed = -1;
	uint32_t max_steps = 0;
	int block = 0, per_cu = 0, tpl = 0; // 0: autotuned
	int lds_cells = 0;                  // per-wave LDS window-cell cache slots; 0: autotuned, -1: none
	bool graph = true;                  // capture fills + launch in a hipGraph, replay it
	bool device_only = false;
};

struct batch_timing
{
	double kernel_ms = 0, total_ms = 0;
	uint64_t transitions = 0;
	int block = 0, per_cu = 0, tpl = 0, lds_cells = 0;
	bool graph = false;
	bool tuned_cached = false; // the launch shape came from the on-disk tuning cache
};

std::vector<std::unique_ptr<result>> simulate_batched(program& p, const batch_params& bp,
													  const std::vector<variant_spec>& vs, hipStream_t stream,
													  batch_timing* timing);

// Sorted window mode: radix-sort n (key, value) pairs on their low key_bits
// bits and sum equal keys.  keys/vals hold the input, keys_alt/vals_alt are
// scratch of the same size; *out_keys / *out_vals point at the unique keys
// and sums (one of the two buffer pairs), their count goes to the device word
// unique_count.  Stream ordered.
void sort_reduce_pairs(device_buffer& temp, unsigned long long* keys, unsigned long long* vals,
					   unsigned long long* keys_alt, unsigned long long* vals_alt, uint64_t n, int key_bits,
					   unsigned long long* unique_count, hipStream_t stream, unsigned long long** out_keys,
					   unsigned long long** out_vals);
// The same sort without the reduction (the keys are already distinct: cells
// compacted from a hash table).  *out_keys / *out_vals: the buffer pair that
// holds the sorted pairs.
void sort_pairs(device_buffer& temp, unsigned long long* keys, unsigned long long* vals, unsigned long long* keys_alt,
				unsigned long long* vals_alt, uint64_t n, int key_bits, hipStream_t stream,
				unsigned long long** out_keys, unsigned long long** out_vals);
// Nonzero entries of a dense array, in index order: out_keys receives their
// indices, out_vals their values and *count (device) their number.
void compact_dense(device_buffer& temp, const unsigned long long* dense, uint64_t n, unsigned long long* out_keys,
				   unsigned long long* out_vals, unsigned long long* count, hipStream_t stream);

// ---- gradient form
struct grad_params
{
	int64_t n_traj = -1;
	uint64_t first_id = 0;
	int64_t seed = -1;
	uint32_t variant = 0;
	uint32_t max_steps = 0;
	int block = 0, per_cu = 0;
	int fold_lanes = 0; // 0: 16
};

struct grad_timing
{
	double kernel_ms = 0;
	uint64_t transitions = 0;
	int block = 0, grid = 0;
};

// dL/dtheta (device, [n_params]) from dL/dP (device, [windows << n_obs]) on
// the trajectories of the forward run with the same ids, seed, variant and
// rates (device [n_params], or null for the model's values).  Stream
// ordered; with timing the stream is synchronised and timing filled.
// scratch: when non-null, the device workspace (counters and tables, at most
// k_noalloc_scratch_bytes) is taken from there and nothing is allocated; the
// gradient kernel must already be loaded (mbg_grad_check).
void simulate_grad(program& p, const grad_params& gp, const float* rates, const double* gprob, double* grad,
				   hipStream_t stream, grad_timing* timing, void* scratch = nullptr);

// Static device scratch of the no-allocation entry point (noalloc.hip): its
// size, the per-device set-up (resolves the array's address on the current
// device; called when a program is compiled) and the address on a device
// (null before set-up there).
constexpr size_t k_noalloc_scratch_bytes = 16 * 1024;
void noalloc_setup(int device);
void* noalloc_scratch(int device);

} // namespace mbg
