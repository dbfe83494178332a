// Simulation runner: sizes and allocates the statistics storage, launches the
// fused persistent simulate kernel and gathers the results.
//
// Replaces simulation_runner::run_simulation (src/simulation_runner.cu:54-180)
// and the per-batch stats fan-out (src/statistics/stats_composite.cpp:5-12).
// There are no epochs: the reference's batches of at most 1e6 trajectories x
// 100 records (:50-51), its partitions/refills (:129-158) and its per-epoch
// memset of 400 MB (:161) are replaced by one launch in which lanes pull
// trajectory ids from a queue until all are done.  A consequence worth
// stating: the reference re-draws the dwell time (without flipping) of every
// trajectory still running at an epoch boundary; here the flip always follows
// its own dwell.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>

#include "../device/mbg_kernel_args.h"
#include "engine.h"
#include "timer.h"

namespace mbg
{

void* device_buffer::ensure(size_t n)
{
	if (n <= bytes && ptr)
		return ptr;
	release();
	size_t want = std::max<size_t>(n, 256);
	hip_check(hipMalloc(&ptr, want), "hipMalloc");
	bytes = want;
	return ptr;
}

void device_buffer::release()
{
	if (ptr)
		(void)hipFree(ptr);
	ptr = nullptr;
	bytes = 0;
}

namespace
{

// Empty hash slots {key = all ones, value = 0}, written by a fill kernel.
void hip_fill_slots(program& p, unsigned long long* table, uint64_t slots, hipStream_t stream)
{
	p.load_util();
	void* args[] = { &table, &slots };
	const unsigned grid = (unsigned)std::min<uint64_t>(4096, (slots + 255) / 256);
	hip_check(hipModuleLaunchKernel(p.fn_fill_slots, grid, 1, 1, 256, 1, 1, 0, stream, args, nullptr),
			  "hipModuleLaunchKernel(mbg_fill_slots)");
}

uint64_t next_pow2(uint64_t v)
{
	uint64_t p = 1;
	while (p < v)
		p <<= 1;
	return p;
}

int bits_for(uint64_t v)
{
	int b = 0;
	while ((1ull << b) <= v && b < 64)
		b++;
	return b;
}

constexpr size_t k_lds_budget = 48 * 1024;
constexpr uint64_t k_dense_limit = 8ull << 30;

// LDS budget left after `used` bytes (0 when the budget is spent)
size_t lds_left(size_t used) { return used < k_lds_budget ? k_lds_budget - used : 0; }

struct plan
{
	int win_mode, fin_mode, ent_lds;
	int win_off = 0, fin_off = 0, ent_off = 0, lds_bytes = 0;
	int tail_off = 0, tail_slots = 0, fp_off = 0, fp_slots = 0; // LDS caches
	int fin_cache_off = 0, fin_cache_slots = 0;
	int own_off = 0;                                             // per-wave item owners (lane kernel)
	int pair_off = 0;                                            // per-wave pair chunk cursors
	int stage_off = 0, stage = 0;                                // per-wave pair staging (lane kernel)
	uint64_t win_cells = 0, fin_cells = 0, win_cap = 0, fin_cap = 0, fp_cap = 0;
};

// Largest power of two <= limit of slots of the given size that fits the
// remaining LDS budget, or 0 below min_slots.
int cache_slots(size_t budget_left, size_t slot_bytes, int limit, int min_slots)
{
	int n = limit;
	while (n >= min_slots && (size_t)n * slot_bytes > budget_left)
		n >>= 1;
	return n >= min_slots ? n : 0;
}

plan make_plan(const model& m, const sim_params& sp, uint64_t n_traj, int windows, int n_obs, bool wave_kernel,
			   uint64_t budget)
{
	plan p;
	// tables in HBM are sized from the run's memory budget (a quarter of the
	// device memory free when the run starts): a table slot costs its 16
	// bytes plus the compacted and sorted copies of its cell (32 bytes)
	const uint64_t slot_budget = std::max<uint64_t>(1ull << 16, budget / 48);
	const uint64_t dense_limit = std::min<uint64_t>(k_dense_limit, std::max<uint64_t>(budget, 1ull << 20));
	const uint64_t win_dense_bytes = n_obs <= 40 ? ((uint64_t)windows << n_obs) * 8 : ~0ull;
	const uint64_t fin_dense_bytes = n_obs <= 40 ? (1ull << n_obs) * 8 : ~0ull;
	size_t lds = 0;

	p.ent_lds = (size_t)windows * 8 <= 16 * 1024;
	if (p.ent_lds)
	{
		p.ent_off = 0;
		lds += (size_t)windows * 8;
	}

	p.win_mode = sp.window_mode;
	// the LDS histogram leaves room for the per-wave tables and caches below
	if (p.win_mode < 0)
		p.win_mode = win_dense_bytes <= 32 * 1024 && win_dense_bytes + 2048 <= lds_left(lds) ? MBG_HIST_LDS
					 : win_dense_bytes <= dense_limit ? MBG_HIST_DENSE
													   : MBG_HIST_HASH;
	if (p.win_mode < 0 || p.win_mode > MBG_HIST_SORT)
		throw error(MBG_E_ARG, "unknown window mode");
	if ((p.win_mode == MBG_HIST_DENSE || p.win_mode == MBG_HIST_LDS) && win_dense_bytes > std::max(dense_limit, budget))
		throw error(MBG_E_LIMIT, "dense window histogram too large; use the hashed or sorted mode");
	if (p.win_mode == MBG_HIST_LDS)
	{
		if (win_dense_bytes > lds_left(lds))
			throw error(MBG_E_LIMIT, "window histogram does not fit in LDS");
		p.win_off = (int)(lds / 8);
		lds += win_dense_bytes;
	}

	p.fin_mode = sp.final_mode;
	if (p.fin_mode > MBG_HIST_LDS)
		throw error(MBG_E_ARG, "unknown final-state mode");
	if (p.fin_mode < 0)
		p.fin_mode = fin_dense_bytes <= lds_left(lds) ? MBG_HIST_LDS
					 : fin_dense_bytes <= (1ull << 30)	 ? MBG_HIST_DENSE
														 : MBG_HIST_HASH;
	if (p.fin_mode != MBG_HIST_HASH && fin_dense_bytes > std::max(dense_limit, budget))
		throw error(MBG_E_LIMIT, "dense final-state histogram too large; use the hashed mode");
	if (p.fin_mode == MBG_HIST_LDS)
	{
		if (fin_dense_bytes > lds_left(lds))
			throw error(MBG_E_LIMIT, "final-state histogram does not fit in LDS");
		p.fin_off = (int)(lds / 8);
		lds += fin_dense_bytes;
	}
	if (p.win_mode == MBG_HIST_SORT)
	{
		p.pair_off = (int)(lds / 8);
		lds += 16 * 32; // per-wave chunk cursor {next, left} (32 bytes) for up to 16 waves
	}
	if (!wave_kernel)
	{
		// 64 bytes per wave (blocks of at most 4 waves), then the block's 3 counters
		p.own_off = (int)(lds / 8);
		lds += 4 * 64 + 3 * 8;
	}
	if ((p.win_mode == MBG_HIST_SORT || p.win_mode == MBG_HIST_HASH) && !wave_kernel)
	{
		// 192 staged pairs per wave (64 arrive at a time; in hashed mode a lane
		// whose insert is still in flight takes no new pair, so a flush may
		// leave more than 64), blocks of at most 4 waves (the lane kernel's bound)
		p.stage = 192;
		p.stage_off = (int)(lds / 8);
		lds += 4 * (size_t)p.stage * 16;
	}
	// block-private caches in front of the tail and fixed-point tables
	if (!m.discrete_time() && p.win_mode != MBG_HIST_LDS)
	{
		// the wave kernel ends one trajectory per wave every few thousand
		// steps and needs its LDS for node tables: a smaller cache
		p.tail_slots = cache_slots(lds_left(lds), 16, wave_kernel ? 128 : 512, 64);
		p.tail_off = (int)(lds / 8);
		lds += (size_t)p.tail_slots * 16;
	}
	{
		// the wave kernel keeps a smaller cache: its LDS holds the node tables
		const size_t slot = 4 * (2 + (size_t)m.n_words());
		p.fp_slots = wave_kernel ? cache_slots(lds_left(lds), slot, 16, 8) : cache_slots(lds_left(lds), slot, 128, 16);
		p.fp_off = (int)(lds / 8);
		lds += ((size_t)p.fp_slots * slot + 7) & ~(size_t)7;
	}
	// block-private cache in front of a hashed final-state table: trajectories
	// that end in the same state add to one LDS slot
	if (p.fin_mode == MBG_HIST_HASH)
	{
		p.fin_cache_slots = cache_slots(lds_left(lds), 16, wave_kernel ? 128 : 512, 64);
		p.fin_cache_off = (int)(lds / 8);
		lds += (size_t)p.fin_cache_slots * 16;
	}
	if (lds > k_lds_budget)
		throw error(MBG_E_LIMIT, "the run's LDS statistics (" + std::to_string(lds) + " bytes) exceed the "
									 + std::to_string(k_lds_budget) + "-byte budget");
	p.lds_bytes = (int)lds;

	if (p.win_mode == MBG_HIST_HASH || p.win_mode == MBG_HIST_SORT)
		if (n_obs + bits_for((uint64_t)windows) > 63)
			throw error(MBG_E_LIMIT, "observed nodes plus window bits exceed the 63-bit cell key");
	if (p.win_mode == MBG_HIST_HASH)
		p.win_cap = sp.window_capacity ? next_pow2(sp.window_capacity)
									   : next_pow2(std::min<uint64_t>(std::max<uint64_t>(n_traj * windows, 1 << 16),
																	   slot_budget));
	else if (p.win_mode != MBG_HIST_SORT) // the sorted mode's pair capacity is the runner's
		p.win_cells = (uint64_t)windows << n_obs;
	if (p.fin_mode == MBG_HIST_HASH)
	{
		if (n_obs > 63)
			throw error(MBG_E_LIMIT, "more than 63 observed nodes");
		p.fin_cap = sp.final_capacity ? next_pow2(sp.final_capacity)
									  : next_pow2(std::min<uint64_t>(std::max<uint64_t>(2 * n_traj, 1 << 12), slot_budget));
	}
	else
		p.fin_cells = 1ull << n_obs;
	p.fp_cap = sp.fixed_capacity ? next_pow2(sp.fixed_capacity)
								 : next_pow2(std::min<uint64_t>(std::max<uint64_t>(2 * n_traj, 1 << 12),
																std::max<uint64_t>(1ull << 16, budget / (8 * (2 + (uint64_t)m.n_words())))));
	(void)m;
	return p;
}

} // namespace

namespace
{

void launch_util(hipFunction_t fn, unsigned grid, void** args, hipStream_t stream, const char* what)
{
	hip_check(hipModuleLaunchKernel(fn, grid, 1, 1, 256, 1, 1, 0, stream, args, nullptr), what);
}

// Occupied slots {key, value} of a hash table as (key, value) pairs; their
// number is added to *count, and the number of values >= esc_min (when not 0)
// to *esc_count.
void table_compact(program& p, const unsigned long long* table, uint64_t cap, unsigned long long* keys,
				   unsigned long long* vals, uint64_t out_cap, unsigned long long* count, hipStream_t stream,
				   uint64_t esc_min = 0, unsigned long long* esc_count = nullptr)
{
	void* args[] = { &table, &cap, &keys, &vals, &out_cap, &count, &esc_min, &esc_count };
	const unsigned g = (unsigned)std::max<uint64_t>(1, (cap + 256 * 16 - 1) / (256 * 16)); // 16 slots per thread
	launch_util(p.fn_table_compact, g, args, stream, "hipModuleLaunchKernel(mbg_table_compact)");
}

// Published fixed-point slots as records {count, state words}.
void fp_compact(program& p, const unsigned int* table, uint64_t cap, int stride, unsigned int* out, uint64_t out_cap,
				unsigned long long* count, hipStream_t stream)
{
	void* args[] = { &table, &cap, &stride, &out, &out_cap, &count };
	const unsigned g = (unsigned)std::min<uint64_t>(4096, std::max<uint64_t>(1, (cap + 255) / 256));
	launch_util(p.fn_fp_compact, g, args, stream, "hipModuleLaunchKernel(mbg_fp_compact)");
}

// Key-sorted window cells as 12-byte records (values >= esc_min, when not 0,
// to the side list xval, their number added to *xcount).
void pack_cells(program& p, const unsigned long long* keys, const unsigned long long* vals, uint64_t n,
				unsigned kbits, uint64_t esc_min, unsigned int* rec, unsigned long long* xval, uint64_t xcap,
				unsigned long long* xcount, hipStream_t stream)
{
	void* args[] = { &keys, &vals, &n, &kbits, &esc_min, &rec, &xval, &xcap, &xcount };
	const unsigned g = (unsigned)((n + 1023) / 1024); // 4 cells per thread
	launch_util(p.fn_pack_cells, g, args, stream, "hipModuleLaunchKernel(mbg_pack_cells)");
}

struct run_events
{
	hipEvent_t e0 = nullptr, es = nullptr, e1 = nullptr, copied = nullptr;
	run_events()
	{
		hip_check(hipEventCreate(&e0), "hipEventCreate");
		hip_check(hipEventCreate(&es), "hipEventCreate");
		hip_check(hipEventCreate(&e1), "hipEventCreate");
		hip_check(hipEventCreateWithFlags(&copied, hipEventDisableTiming), "hipEventCreateWithFlags");
	}
	~run_events()
	{
		for (hipEvent_t e : { e0, es, e1, copied })
			if (e)
				(void)hipEventDestroy(e);
	}
};

size_t align256(size_t v) { return (v + 255) & ~(size_t)255; }

} // namespace

int default_refill_lanes(int n_nodes)
{
	// a refill round initialises 64 / ceil(N/4) trajectories; wait for 4 to 8
	// idle lanes: waiting for m costs about (m - 1) / 128 of the wave's work.
	// Measured on MI355X: grad_30's simulate 3.27 -> 2.49 ms at 8, plain_100
	// 10.18 -> 9.76 ms at 4, no gain past 8 and a loss at 16 on plain_100.
	const int blocks = (n_nodes + 3) / 4;
	const int per_round = blocks <= 64 ? 64 / blocks : 1;
	return std::min(std::max(per_round, 4), 8);
}

uint64_t device_budget()
{
	size_t free_b = 0, total_b = 0;
	if (hipMemGetInfo(&free_b, &total_b) != hipSuccess || !free_b)
		return 8ull << 30;
	return (uint64_t)free_b / 4;
}

void default_storage_modes(const model& m, int& win_mode, int& fin_mode)
{
	const plan pl = make_plan(m, sim_params {}, (uint64_t)m.sample_count(), m.windows(), m.n_observed(), false,
							  device_budget());
	win_mode = pl.win_mode;
	fin_mode = pl.fin_mode;
}

namespace
{

// One run.  With probs set (window_probs_device) the run is stream ordered
// end to end: rates from device memory (d_rates, or the model's baked ones
// when null), dense window cells converted to probabilities into probs, no
// host copy and no host wait; the function then returns null.
std::unique_ptr<result> run_impl(program& p, const sim_params& sp, hipStream_t stream, const float* d_rates,
								 double* probs)
{
	scope_timer tm("runner> simulate");
	const double t0 = wall_seconds();
	const model& m = *p.mdl;
	auto res = std::make_unique<result>();
	const int W = m.n_words();
	const int n_obs = m.n_observed();
	const int windows = m.windows();
	const uint64_t n_traj = sp.n_traj < 0 ? (uint64_t)m.sample_count() : (uint64_t)sp.n_traj;
	const uint64_t seed = sp.seed < 0 ? m.seed() : (uint64_t)sp.seed;
	if (windows <= 0 || windows > (1 << 24))
		throw error(MBG_E_LIMIT, "window count out of range");

	res->windows = windows;
	res->n_obs = n_obs;
	res->n_words = W;
	res->discrete = m.discrete_time();
	res->tick = m.time_tick();
	res->max_time = m.max_time();
	res->n_traj = n_traj;
	res->norm = sp.normalization > 0 ? sp.normalization : (double)n_traj;
	res->user_norm = sp.normalization > 0;
	res->first_id = sp.first_id;
	res->seed = seed;
	res->variant = sp.variant;
	res->fingerprint = run_fingerprint(p, sp);
	for (auto& n : m.nodes)
		res->names.push_back(n.name);
	res->observed = m.observed_nodes();
	res->ent.assign(windows, 0);
	res->woff.assign((size_t)windows + 1, 0);

	const bool use_wave =
		sp.kernel == 1 || (sp.kernel < 0 && (W > max_lane_words || (W > auto_lane_words && p.wtab != nullptr)));
	if (!use_wave && W > max_lane_words)
		throw error(MBG_E_LIMIT, "the lane kernel serves up to " + std::to_string(32 * max_lane_words) + " nodes");
	if (use_wave && p.opts.entropy == entropy_formula::reference)
		throw error(MBG_E_ARG, "the wave kernel computes the log-sum entropy only");
	if (!sp.rates.empty() && use_wave)
		throw error(MBG_E_LIMIT, "run-time rates are served by the lane kernel (up to 512 nodes)");
	if (!sp.rates.empty() && sp.rates.size() != p.gen.param_names.size())
		throw error(MBG_E_ARG, "expected " + std::to_string(p.gen.param_names.size()) + " rates");
	res->kernel = use_wave ? 1 : 0;
	plan pl = make_plan(m, sp, n_traj, windows, n_obs, use_wave, device_budget());
	if (probs && (use_wave || (pl.win_mode != MBG_HIST_DENSE && pl.win_mode != MBG_HIST_LDS) || n_obs > 24))
		throw error(MBG_E_LIMIT, "device window probabilities need the lane kernel and dense window cells "
								 "(at most 24 observed nodes)");
	const bool rt_rates = !sp.rates.empty() || (probs && d_rates);
	hipFunction_t fn = use_wave ? p.wave_function() : p.function(rt_rates ? 2 : 0, pl.win_mode, pl.fin_mode);
	p.load_util();
	if (!p.copy_stream)
		hip_check(hipStreamCreateWithFlags(&p.copy_stream, hipStreamNonBlocking), "hipStreamCreateWithFlags");
	if (!p.mailbox)
		p.mailbox = p.pool->get(4096);
	auto* mb = (unsigned long long*)p.mailbox.get(); // [0, 16) counters, [16, 22) compaction counts

	// The run's output slot: its previous copy to the host must be complete
	// before this run writes into it or grows its buffers (a host-side wait:
	// a buffer may be reallocated below).
	out_slot& slot = p.slots[p.slot_next];
	if (!probs)
		p.slot_next ^= 1;
	if (slot.pending && !probs)
	{
		hip_check(hipEventSynchronize(slot.copied), "hipEventSynchronize");
		slot.pending = false;
	}
	p.order_after_last_run(stream);

	// run-time rates: a parameter view {rate, rate*log2(rate), enable masks} in device memory
	const mbg_param_view* d_view = nullptr;
	if (probs && d_rates)
	{
		// the view over the caller's device rates, built on the device
		const int np = (int)p.gen.param_names.size();
		const size_t off_lr = (sizeof(mbg_param_view) + 15) & ~(size_t)15, off_en = off_lr + ((size_t)np * 4 + 15) / 16 * 16;
		char* base = (char*)p.b_rate_view.ensure(off_en + 8 * (size_t)W);
		auto* view = (mbg_param_view*)base;
		auto* lr = (float*)(base + off_lr);
		auto* en = (uint32_t*)(base + off_en);
		int nw = W;
		void* args[] = { &view, &d_rates, &lr, &en, (void*)&np, &nw };
		const unsigned g = (unsigned)std::max(1, (std::max(np, 2 * W) + 255) / 256);
		hip_check(hipModuleLaunchKernel(p.fn_rate_view, g, 1, 1, 256, 1, 1, 0, stream, args, nullptr),
				  "hipModuleLaunchKernel(mbg_rate_view)");
		d_view = view;
	}
	else if (!sp.rates.empty())
	{
		const size_t np = sp.rates.size();
		const size_t off_rate = (sizeof(mbg_param_view) + 15) & ~(size_t)15;
		const size_t off_lr = off_rate + np * 4, off_en = off_lr + np * 4;
		std::vector<char> img(off_en + 8 * (size_t)W, 0);
		char* base = (char*)p.b_rate_view.ensure(img.size());
		mbg_param_view v;
		v.rate = (const float*)(base + off_rate);
		v.rate_lr = (const float*)(base + off_lr);
		v.up_en = (const uint32_t*)(base + off_en);
		v.dn_en = (const uint32_t*)(base + off_en + 4 * (size_t)W);
		std::memcpy(img.data(), &v, sizeof(v));
		for (size_t k = 0; k < np; k++)
		{
			const float lr = rate_log2_rate(sp.rates[k]);
			std::memcpy(img.data() + off_rate + 4 * k, &sp.rates[k], 4);
			std::memcpy(img.data() + off_lr + 4 * k, &lr, 4);
		}
		std::memset(img.data() + off_en, 0xff, 8 * (size_t)W);
		hip_check(hipMemcpyAsync(base, img.data(), img.size(), hipMemcpyHostToDevice, stream), "hipMemcpyAsync");
		// the pageable source must stay valid until the copy has read it
		hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize");
		d_view = (const mbg_param_view*)base;
	}
	// tail records for window storage in HBM (continuous time); not in the
	// stream-ordered probability run, which reads no overflow counter
	const bool tails = !m.discrete_time() && pl.win_mode != MBG_HIST_LDS && !probs;
	uint64_t tail_cap = tails ? next_pow2(std::min<uint64_t>(std::max<uint64_t>(n_traj / 4, 1 << 12), 1 << 20)) : 0;
	// sorted window mode: pair capacity from the caller, the last run, or windows per trajectory
	uint64_t pair_cap = 0;
	if (pl.win_mode == MBG_HIST_SORT)
	{
		const uint64_t per = p.pairs_per_traj ? p.pairs_per_traj + p.pairs_per_traj / 8 + 1
											  : std::max<uint64_t>((uint64_t)windows, 16);
		pair_cap = sp.window_capacity ? sp.window_capacity : std::max<uint64_t>(n_traj * per, 1 << 16);
	}
	// hashed window mode: the spill list of cells past the probe bound
	uint64_t spill_cap = pl.win_mode == MBG_HIST_HASH ? std::max<uint64_t>(1 << 16, pl.win_cap / 8) : 0;
	std::vector<float> istate;
	for (auto& n : m.nodes)
		istate.push_back(n.istate);
	auto ev = std::make_shared<run_events>();
	const size_t fp_stride = 2 + (size_t)W;

	// Clearing: every table starts a run cleared.  A run clears what it used
	// right after it has been read (compaction, or the counters' read), so a
	// following run usually finds its tables ready; `use` clears a buffer the
	// last run did not leave cleared and marks it in use.
	auto use = [](cleared_buf& c, const void* ptr, size_t bytes, auto&& clear) {
		if (!c.holds(ptr, bytes))
			clear();
		c.ptr = nullptr;
	};
	auto memset0 = [&](void* ptr, size_t bytes) {
		hip_check(hipMemsetAsync(ptr, 0, bytes, stream), "hipMemsetAsync");
	};

	// Window cells travel to the host as 12-byte records when their keys take
	// at most 60 bits (the copy bounds a step whose statistics are large):
	// up to 31 key bits leave the record room for any 64-bit value; above
	// that, values of 2^(95 - key bits) or more go to a side list, which the
	// hashed table's compaction counts so every copy size is known at the
	// run's one mid-point synchronisation.
	const unsigned pk_bits = (unsigned)std::max(n_obs + bits_for((uint64_t)windows - 1), 1);
	const uint64_t esc_min = pk_bits <= 31 ? 0ull : 1ull << (95 - pk_bits);
	for (int attempt = 0;; attempt++)
	{
		// ---- storage
		auto* counters = (unsigned long long*)p.b_counters.ensure(MBG_NCOUNTERS * 8);
		use(p.cl_counters, counters, MBG_NCOUNTERS * 8, [&] { memset0(counters, MBG_NCOUNTERS * 8); });
		cleared_buf& cl_ent = probs ? p.cl_pent : slot.cl_ent;
		auto* ent = (unsigned long long*)(probs ? p.b_pent : slot.ent).ensure((size_t)windows * 8);
		use(cl_ent, ent, (size_t)windows * 8, [&] { memset0(ent, (size_t)windows * 8); });
		unsigned long long *win_vals = nullptr, *win_keys = nullptr, *fin_vals = nullptr, *fin_keys = nullptr;
		unsigned long long *pk = nullptr, *pv = nullptr;
		if (pl.win_mode == MBG_HIST_SORT)
		{
			pk = (unsigned long long*)slot.pk.ensure(pair_cap * 8);
			pv = (unsigned long long*)slot.pv.ensure(pair_cap * 8);
		}
		else if (pl.win_mode == MBG_HIST_HASH)
		{
			win_keys = (unsigned long long*)p.b_win_keys.ensure(pl.win_cap * 16);
			use(p.cl_win_keys, win_keys, pl.win_cap * 16, [&] { hip_fill_slots(p, win_keys, pl.win_cap, stream); });
		}
		else
		{
			win_vals = (unsigned long long*)p.b_win_vals.ensure(pl.win_cells * 8);
			use(p.cl_win_vals, win_vals, pl.win_cells * 8, [&] { memset0(win_vals, pl.win_cells * 8); });
		}
		auto* win_spill = spill_cap ? (unsigned long long*)p.b_win_spill.ensure(spill_cap * 16) : nullptr;
		if (pl.fin_mode == MBG_HIST_HASH)
		{
			fin_keys = (unsigned long long*)p.b_fin_keys.ensure(pl.fin_cap * 16);
			use(p.cl_fin_keys, fin_keys, pl.fin_cap * 16, [&] { hip_fill_slots(p, fin_keys, pl.fin_cap, stream); });
		}
		else
		{
			fin_vals = (unsigned long long*)p.b_fin_vals.ensure(pl.fin_cells * 8);
			use(p.cl_fin_vals, fin_vals, pl.fin_cells * 8, [&] { memset0(fin_vals, pl.fin_cells * 8); });
		}
		unsigned long long* tail_tab = nullptr;
		if (tails)
		{
			tail_tab = (unsigned long long*)p.b_tail.ensure(tail_cap * 16);
			use(p.cl_tail, tail_tab, tail_cap * 16, [&] { hip_fill_slots(p, tail_tab, tail_cap, stream); });
		}
		auto* fp = (unsigned int*)p.b_fp.ensure(pl.fp_cap * fp_stride * 4);
		use(p.cl_fp, fp, pl.fp_cap * fp_stride * 4, [&] { memset0(fp, pl.fp_cap * fp_stride * 4); });
		// the run's tables cleared again once they have been read
		auto clear_used = [&] {
			memset0(counters, MBG_NCOUNTERS * 8);
			p.cl_counters = cleared_buf { counters, MBG_NCOUNTERS * 8 };
			if (win_keys)
			{
				hip_fill_slots(p, win_keys, pl.win_cap, stream);
				p.cl_win_keys = cleared_buf { win_keys, pl.win_cap * 16 };
			}
			if (win_vals)
			{
				memset0(win_vals, pl.win_cells * 8);
				p.cl_win_vals = cleared_buf { win_vals, pl.win_cells * 8 };
			}
			if (fin_keys)
			{
				hip_fill_slots(p, fin_keys, pl.fin_cap, stream);
				p.cl_fin_keys = cleared_buf { fin_keys, pl.fin_cap * 16 };
			}
			if (fin_vals)
			{
				memset0(fin_vals, pl.fin_cells * 8);
				p.cl_fin_vals = cleared_buf { fin_vals, pl.fin_cells * 8 };
			}
			if (tail_tab)
			{
				hip_fill_slots(p, tail_tab, tail_cap, stream);
				p.cl_tail = cleared_buf { tail_tab, tail_cap * 16 };
			}
			memset0(fp, pl.fp_cap * fp_stride * 4);
			p.cl_fp = cleared_buf { fp, pl.fp_cap * fp_stride * 4 };
		};

		// the model's initial-state probabilities, uploaded once per buffer
		// (from pinned memory the program owns: the copy is stream ordered)
		auto* d_istate = (float*)p.b_istate.ensure(istate.size() * 4);
		if (p.istate_dev != d_istate)
		{
			p.istate_pinned = p.pool->get(istate.size() * 4);
			std::memcpy(p.istate_pinned.get(), istate.data(), istate.size() * 4);
			hip_check(hipMemcpyAsync(d_istate, p.istate_pinned.get(), istate.size() * 4, hipMemcpyHostToDevice,
									 stream),
					  "hipMemcpyAsync");
			p.istate_dev = d_istate;
		}

		// ---- kernel arguments
		mbg_sim_args a;
		std::memset(&a, 0, sizeof(a));
		a.n_traj = n_traj;
		a.id_base = sp.first_id;
		a.seed_lo = (uint32_t)seed;
		a.seed_hi = (uint32_t)(seed >> 32);
		a.variant = sp.variant;
		a.max_steps = sp.max_steps ? sp.max_steps : (1u << 30);
		a.n_nodes = m.n_nodes();
		a.n_obs = n_obs;
		a.windows = windows;
		a.refill_lanes = sp.refill_lanes > 0 ? sp.refill_lanes : default_refill_lanes(m.n_nodes());
		a.discrete = m.discrete_time();
		a.max_time = m.max_time();
		a.tick = m.time_tick();
		a.qscale = a.discrete ? 0x1p32f : 0x1p32f / a.tick;
		a.istate = d_istate;
		a.win_mode = pl.win_mode;
		a.fin_mode = pl.fin_mode;
		a.ent_lds = pl.ent_lds;
		a.lds_bytes = pl.lds_bytes;
		a.win_lds_off = pl.win_off;
		a.fin_lds_off = pl.fin_off;
		a.ent_lds_off = pl.ent_off;
		a.win_vals = win_vals;
		a.win_table = win_keys;
		a.win_mask = pl.win_cap ? pl.win_cap - 1 : 0;
		a.win_spill_keys = win_spill;
		a.win_spill_vals = win_spill ? win_spill + spill_cap : nullptr;
		a.win_spill_cap = spill_cap;
		a.win_ent = ent;
		a.fin_vals = fin_vals;
		a.fin_table = fin_keys;
		a.fin_mask = pl.fin_cap ? pl.fin_cap - 1 : 0;
		a.fp_table = fp;
		a.fp_mask = pl.fp_cap - 1;
		a.counters = counters;
		a.pv = d_view;
		a.tail_table = tail_tab;
		a.tail_mask = tail_cap ? tail_cap - 1 : 0;
		a.tail_lds_off = pl.tail_off;
		a.tail_lds_slots = tails ? pl.tail_slots : 0;
		a.fp_lds_off = pl.fp_off;
		a.fp_lds_slots = pl.fp_slots;
		a.own_lds_off = pl.own_off;
		a.fin_lds_cache_off = pl.fin_cache_off;
		a.fin_lds_cache_slots = pl.fin_cache_slots;
		a.pair_keys = pk;
		a.pair_vals = pv;
		a.pair_cap = pair_cap;
		a.pair_lds_off = pl.pair_off;
		a.pair_chunk = 1024;
		a.pair_stage_off = pl.stage_off;
		a.pair_stage = pl.stage;

		// ---- launch: a persistent grid filling every CU
		int block = sp.block_size > 0 ? sp.block_size : 256;
		if (use_wave && sp.block_size <= 0)
		{
			// the wave kernel stages the node tables once per block: the block
			// size (1 to 8 waves) that keeps the most waves resident per CU
			// (measured on scale_1000: 6 waves x 2 blocks ran 1.4x faster than
			// 4 x 2 or 8 x 1)
			int best_waves = 0;
			for (int bs = 512; bs >= 64; bs -= 64)
			{
				const int lds = ((pl.lds_bytes + 15) & ~15) + (int)p.wtab->image.size() + (bs / 64) * p.wtab->rows * 64 * 4;
				if (lds > 160 * 1024)
					continue;
				int occ = 0;
				hip_check(hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&occ, fn, bs, lds),
						  "hipModuleOccupancyMaxActiveBlocksPerMultiprocessor");
				if (occ * bs / 64 > best_waves)
				{
					best_waves = occ * bs / 64;
					block = bs;
				}
			}
		}
		{
			int max_threads = 0;
			hip_check(hipFuncGetAttribute(&max_threads, HIP_FUNC_ATTRIBUTE_MAX_THREADS_PER_BLOCK, fn),
					  "hipFuncGetAttribute");
			if (block > max_threads || block % 64 != 0)
				throw error(MBG_E_ARG, "block_size " + std::to_string(block) + " is not a multiple of 64 up to "
										   + std::to_string(max_threads) + " for this kernel");
		}
		int lds_total = pl.lds_bytes;
		uint64_t work_per_block = (uint64_t)block; // trajectories in flight per block
		if (use_wave)
		{
			const wave_tables& wt = *p.wtab;
			auto* img = (char*)p.b_wave.ensure(wt.image.size());
			if (p.wave_dev != img)
			{
				// the tables are the program's (they outlive the stream-ordered copy)
				hip_check(hipMemcpyAsync(img, wt.image.data(), wt.image.size(), hipMemcpyHostToDevice, stream),
						  "hipMemcpyAsync");
				hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize");
				p.wave_dev = img;
			}
			a.wt.nodes = (const mbg_wnode*)img;
			a.wt.child_off = (const unsigned*)(img + wt.nodes.size() * sizeof(mbg_wnode));
			a.wt.child = (const unsigned short*)(a.wt.child_off + wt.child_off.size());
			a.wt.obs = a.wt.child + wt.child.size();
			a.wt.n_child = (int)wt.child.size();
			a.wt.npl = wt.npl;
			a.wt.obs_tab_off = wt.obs_tab_off;
			a.wt.tab_bytes = (int)wt.image.size();
			a.wave_lds_off = (pl.lds_bytes + 15) & ~15;
			lds_total = a.wave_lds_off + a.wt.tab_bytes + (block / 64) * wt.rows * 64 * 4;
			if (lds_total > 160 * 1024)
				throw error(MBG_E_LIMIT, "wave kernel tables do not fit in LDS; use a smaller block size");
			work_per_block = (uint64_t)(block / 64);
		}
		else if (block > 256)
			throw error(MBG_E_ARG, "the lane kernel takes blocks of at most 256 lanes");
		int per_cu = sp.blocks_per_cu;
		if (per_cu <= 0)
		{
			hip_check(hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, block, lds_total),
					  "hipModuleOccupancyMaxActiveBlocksPerMultiprocessor");
			per_cu = std::max(per_cu, 1);
		}
		uint64_t blocks_needed = (n_traj + work_per_block - 1) / work_per_block;
		uint64_t grid = std::min<uint64_t>((uint64_t)p.num_cus * per_cu, std::max<uint64_t>(blocks_needed, 1));

		hip_check(hipEventRecord(ev->e0, stream), "hipEventRecord");
		if (n_traj > 0)
		{
			void* args[] = { &a };
			hip_check(hipModuleLaunchKernel(fn, (unsigned)grid, 1, 1, block, 1, 1, lds_total, stream, args, nullptr),
					  "hipModuleLaunchKernel(simulate)");
		}
		hip_check(hipEventRecord(ev->es, stream), "hipEventRecord");
		if (n_traj > 0 && tails)
		{
			void* args[] = { &a };
			const unsigned eg = (unsigned)std::min<uint64_t>(1024, (tail_cap + 255) / 256);
			hip_check(hipModuleLaunchKernel(p.fn_expand_tails, eg, 1, 1, 256, 1, 1, 0, stream, args, nullptr),
					  "hipModuleLaunchKernel(mbg_expand_tails)");
		}

		if (probs)
		{
			// dense cells to probabilities, stream ordered, and done
			void* cargs[] = { &win_vals, (void*)&pl.win_cells, nullptr, &probs };
			double denom = res->discrete ? res->norm : res->norm * 4294967296.0;
			cargs[2] = &denom;
			const unsigned g = (unsigned)std::min<uint64_t>(4096, std::max<uint64_t>(1, (pl.win_cells + 255) / 256));
			hip_check(hipModuleLaunchKernel(p.fn_cells_to_probs, g, 1, 1, 256, 1, 1, 0, stream, cargs, nullptr),
					  "hipModuleLaunchKernel(mbg_cells_to_probs)");
			clear_used();
			memset0(ent, (size_t)windows * 8);
			p.cl_pent = cleared_buf { ent, (size_t)windows * 8 };
			p.mark_run_done(stream);
			return nullptr;
		}

		// ---- the occupied cells of every table, compacted into the output slot
		// counts: [0] window cells, [1] final states, [2] fixed points, [3] unique
		// sorted cells, [4] window values for the packed copy's side list, [5]
		// the pack kernel's side-list cursor
		auto* cnt = (unsigned long long*)slot.cnt.ensure(6 * 8);
		if (!sp.device_only)
		{
			use(slot.cl_cnt, cnt, 6 * 8, [&] { memset0(cnt, 6 * 8); });
			if (pl.win_mode == MBG_HIST_HASH)
			{
				// room after the table's cells for the spilled ones
				auto* tk = (unsigned long long*)slot.pk.ensure((pl.win_cap + spill_cap) * 8);
				auto* tv = (unsigned long long*)slot.pv.ensure((pl.win_cap + spill_cap) * 8);
				table_compact(p, win_keys, pl.win_cap, tk, tv, pl.win_cap, cnt + 0, stream, esc_min, cnt + 4);
			}
			else if (pl.win_mode != MBG_HIST_SORT)
			{
				auto* tk = (unsigned long long*)slot.pk.ensure(pl.win_cells * 8);
				auto* tv = (unsigned long long*)slot.pv.ensure(pl.win_cells * 8);
				compact_dense(p.b_sort_tmp, win_vals, pl.win_cells, tk, tv, cnt + 0, stream);
			}
			if (pl.fin_mode == MBG_HIST_HASH)
			{
				auto* k = (unsigned long long*)slot.fk.ensure(pl.fin_cap * 8);
				auto* v = (unsigned long long*)slot.fv.ensure(pl.fin_cap * 8);
				table_compact(p, fin_keys, pl.fin_cap, k, v, pl.fin_cap, cnt + 1, stream);
			}
			else
			{
				auto* k = (unsigned long long*)slot.fk.ensure(pl.fin_cells * 8);
				auto* v = (unsigned long long*)slot.fv.ensure(pl.fin_cells * 8);
				compact_dense(p.b_sort_tmp, fin_vals, pl.fin_cells, k, v, cnt + 1, stream);
			}
			auto* fpc = (unsigned int*)slot.fpc.ensure(pl.fp_cap * (fp_stride - 1) * 4);
			fp_compact(p, fp, pl.fp_cap, (int)fp_stride, fpc, pl.fp_cap, cnt + 2, stream);
		}
		hip_check(hipMemcpyAsync(mb, counters, MBG_NCOUNTERS * 8, hipMemcpyDeviceToHost, stream), "hipMemcpyAsync");
		if (!sp.device_only)
		{
			hip_check(hipMemcpyAsync(mb + 16, cnt, 6 * 8, hipMemcpyDeviceToHost, stream), "hipMemcpyAsync");
			// the sorted mode's unique count (cnt[3]) is written and read after the sort
			if (pl.win_mode != MBG_HIST_SORT)
			{
				memset0(cnt, 6 * 8);
				slot.cl_cnt = cleared_buf { cnt, 6 * 8 };
			}
		}
		clear_used();
		if (sp.device_only)
		{
			memset0(ent, (size_t)windows * 8);
			cl_ent = cleared_buf { ent, (size_t)windows * 8 };
			hip_check(hipEventRecord(ev->e1, stream), "hipEventRecord");
		}
		hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize");
		unsigned long long ctr[MBG_NCOUNTERS];
		std::memcpy(ctr, mb, sizeof(ctr));
		const uint64_t n_pairs = ctr[MBG_CTR_PAIRS];
		const bool pair_over = pl.win_mode == MBG_HIST_SORT && n_pairs > pair_cap;
		if (pl.win_mode == MBG_HIST_SORT && n_traj)
			p.pairs_per_traj = (n_pairs + n_traj - 1) / n_traj;

		// the window table's own overflow is the spill list's: cells past the
		// probe bound that found the list full
		const uint64_t n_spill = ctr[MBG_CTR_WIN_SPILL];
		const bool win_over = ctr[MBG_CTR_WIN_OVERFLOW] != 0, fin_over = ctr[MBG_CTR_FIN_OVERFLOW] != 0,
				   fp_over = ctr[MBG_CTR_FP_OVERFLOW] != 0, tail_over = ctr[MBG_CTR_TAIL_OVERFLOW] != 0;
		if (ctr[MBG_CTR_FP_TIMEOUT])
			throw error(MBG_E_INTERNAL, "fixed-point table: a slot was never published");
		if (win_over || fin_over || fp_over || tail_over || pair_over)
		{
			if (attempt >= 6)
				throw error(MBG_E_OVERFLOW, std::string("device hash table overflow after resizing (")
												+ (win_over ? "window " : "") + (fin_over ? "final " : "")
												+ (fp_over ? "fixed-point " : "") + (tail_over ? "tail " : "")
												+ (pair_over ? "window pairs" : "") + ")");
			if (tail_over)
				tail_cap *= 4;
			if (pair_over)
			{
				if (sp.window_capacity)
					throw error(MBG_E_OVERFLOW, "window pair buffer overflow; raise window_capacity");
				pair_cap = n_pairs + n_pairs / 16 + 1;
			}
			if (win_over)
			{
				// the next launch's list takes every spilled cell; the table grows
				// unless the caller fixed its capacity
				spill_cap = next_pow2(n_spill + n_spill / 8 + 1);
				if (!sp.window_capacity)
					pl.win_cap *= 2;
			}
			if (fin_over)
			{
				if (sp.final_capacity)
					throw error(MBG_E_OVERFLOW, "final-state hash table overflow; raise final_capacity");
				pl.fin_cap *= 2;
			}
			if (fp_over)
			{
				if (sp.fixed_capacity)
					throw error(MBG_E_OVERFLOW, "fixed-point table overflow; raise fixed_capacity");
				pl.fp_cap *= 2;
			}
			continue;
		}

		res->window_pairs = pl.win_mode == MBG_HIST_SORT ? n_pairs : 0;
		res->window_spills = n_spill;
		res->transitions = ctr[MBG_CTR_TRANSITIONS];
		res->finished = ctr[MBG_CTR_FINISHED];
		res->truncated = ctr[MBG_CTR_TRUNCATED];
		res->window_mode = pl.win_mode;
		res->final_mode = pl.fin_mode;
		res->window_capacity = pl.win_mode == MBG_HIST_SORT ? pair_cap : pl.win_cap;

		if (sp.device_only)
		{
			p.mark_run_done(stream);
			float ms = 0, sim_ms = 0;
			hip_check(hipEventElapsedTime(&ms, ev->e0, ev->e1), "hipEventElapsedTime");
			hip_check(hipEventElapsedTime(&sim_ms, ev->e0, ev->es), "hipEventElapsedTime");
			res->kernel_ms = ms;
			res->simulate_ms = sim_ms;
			res->total_ms = (wall_seconds() - t0) * 1e3;
			return res;
		}

		// ---- order the cells on the device
		const int key_bits = n_obs + bits_for((uint64_t)windows - 1);
		uint64_t n_cells = mb[16], n_fin = mb[17], n_fp = mb[18];
		unsigned long long *wk = (unsigned long long*)slot.pk.ptr, *wv = (unsigned long long*)slot.pv.ptr;
		if (pl.win_mode == MBG_HIST_HASH && n_spill)
		{
			// the spilled cells after the table's, sorted together and equal keys summed
			hip_check(hipMemcpyAsync(wk + n_cells, win_spill, n_spill * 8, hipMemcpyDeviceToDevice, stream),
					  "hipMemcpyAsync");
			hip_check(hipMemcpyAsync(wv + n_cells, win_spill + spill_cap, n_spill * 8, hipMemcpyDeviceToDevice, stream),
					  "hipMemcpyAsync");
			auto* ka = (unsigned long long*)slot.pka.ensure((n_cells + n_spill) * 8);
			auto* va = (unsigned long long*)slot.pva.ensure((n_cells + n_spill) * 8);
			sort_reduce_pairs(p.b_sort_tmp, wk, wv, ka, va, n_cells + n_spill, key_bits, cnt + 3, stream, &wk, &wv);
			hip_check(hipMemcpyAsync(mb + 19, cnt + 3, 8, hipMemcpyDeviceToHost, stream), "hipMemcpyAsync");
			hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize");
			n_cells = mb[19];
			memset0(cnt + 3, 8); // the slot's counts stay cleared for its next run
		}
		else if (pl.win_mode == MBG_HIST_HASH && n_cells)
		{
			auto* ka = (unsigned long long*)slot.pka.ensure(n_cells * 8);
			auto* va = (unsigned long long*)slot.pva.ensure(n_cells * 8);
			sort_pairs(p.b_sort_tmp, wk, wv, ka, va, n_cells, key_bits, stream, &wk, &wv);
		}
		else if (pl.win_mode == MBG_HIST_SORT)
		{
			n_cells = 0;
			if (n_pairs)
			{
				auto* ka = (unsigned long long*)slot.pka.ensure(pair_cap * 8);
				auto* va = (unsigned long long*)slot.pva.ensure(pair_cap * 8);
				sort_reduce_pairs(p.b_sort_tmp, pk, pv, ka, va, n_pairs, key_bits + 1, // + the padding bit
								  cnt + 3, stream, &wk, &wv);
				hip_check(hipMemcpyAsync(mb + 19, cnt + 3, 8, hipMemcpyDeviceToHost, stream), "hipMemcpyAsync");
				hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize");
				n_cells = mb[19];
			}
		}
		unsigned long long *fk = (unsigned long long*)slot.fk.ptr, *fv = (unsigned long long*)slot.fv.ptr;
		if (pl.fin_mode == MBG_HIST_HASH && n_fin)
		{
			auto* ka = (unsigned long long*)slot.fka.ensure(n_fin * 8);
			auto* va = (unsigned long long*)slot.fva.ensure(n_fin * 8);
			sort_pairs(p.b_sort_tmp, fk, fv, ka, va, n_fin, std::max(n_obs, 1), stream, &fk, &fv);
		}
		// ---- window cells as 12-byte records for the copy: the hashed table's
		// cells (unless spilled cells were merged into them after the count)
		// or other modes' cells whose values always fit
		const bool packed = n_cells && pk_bits <= 60
							&& ((pl.win_mode == MBG_HIST_HASH && !n_spill) || (pl.win_mode != MBG_HIST_SORT && !esc_min));
		const uint64_t n_x = packed && esc_min ? mb[20] : 0;
		unsigned int* rec = nullptr;
		unsigned long long* xv = nullptr;
		if (packed)
		{
			rec = (unsigned int*)slot.prec.ensure(n_cells * 12);
			xv = (unsigned long long*)slot.px.ensure(std::max<uint64_t>(n_x, 1) * 8);
			pack_cells(p, wk, wv, n_cells, pk_bits, esc_min, rec, xv, n_x, cnt + 5, stream);
		}
		res->window_record_bytes = packed ? 12 : 16;
		res->window_escapes = n_x;
		hip_check(hipEventRecord(ev->e1, stream), "hipEventRecord");
		p.mark_run_done(stream);

		// ---- copy out on the copy stream, into one pinned block
		// packed: records, side list, the pack kernel's side-list count
		const size_t o_wv = packed ? align256(n_cells * 12) : align256(n_cells * 8);
		const size_t o_ent = packed ? o_wv + align256(n_x * 8) + 256 : o_wv + align256(n_cells * 8);
		const size_t o_fk = o_ent + align256((size_t)windows * 8), o_fv = o_fk + align256(n_fin * 8);
		const size_t o_fp = o_fv + align256(n_fin * 8), total = o_fp + align256(n_fp * (fp_stride - 1) * 4);
		std::shared_ptr<char> blk = p.pool->get(total);
		char* h = blk.get();
		hipStream_t cs = p.copy_stream;
		hip_check(hipStreamWaitEvent(cs, ev->e1, 0), "hipStreamWaitEvent");
		auto d2h = [&](size_t off, const void* src, size_t bytes) {
			if (bytes)
				hip_check(hipMemcpyAsync(h + off, src, bytes, hipMemcpyDeviceToHost, cs), "hipMemcpyAsync");
		};
		if (packed)
		{
			d2h(0, rec, n_cells * 12);
			d2h(o_wv, xv, n_x * 8);
			d2h(o_wv + align256(n_x * 8), cnt + 5, 8);
			// the cursor cleared for the slot's next run, after its copy
			hip_check(hipMemsetAsync(cnt + 5, 0, 8, cs), "hipMemsetAsync");
		}
		else
		{
			d2h(0, wk, n_cells * 8);
			d2h(o_wv, wv, n_cells * 8);
		}
		d2h(o_ent, ent, (size_t)windows * 8);
		d2h(o_fk, fk, n_fin * 8);
		d2h(o_fv, fv, n_fin * 8);
		d2h(o_fp, slot.fpc.ptr, n_fp * (fp_stride - 1) * 4);
		// the slot's entropy sums cleared for its next run, after their copy
		hip_check(hipMemsetAsync(ent, 0, (size_t)windows * 8, cs), "hipMemsetAsync");
		slot.cl_ent = cleared_buf { ent, (size_t)windows * 8 };
		// the slot keeps an event of its own (the result's goes with the result)
		if (!slot.copied)
			hip_check(hipEventCreateWithFlags(&slot.copied, hipEventDisableTiming), "hipEventCreateWithFlags");
		hip_check(hipEventRecord(slot.copied, cs), "hipEventRecord");
		hip_check(hipEventRecord(ev->copied, cs), "hipEventRecord");
		slot.pending = true;

		result* r = res.get();
		r->on_ready = [r, ev, blk, h, n_cells, n_fin, n_fp, o_wv, o_ent, o_fk, o_fv, o_fp, fp_stride, t0, W, packed,
					   n_x, pk_bits, esc_min]() {
			hip_check(hipEventSynchronize(ev->copied), "hipEventSynchronize");
			float ms = 0, sim_ms = 0;
			hip_check(hipEventElapsedTime(&ms, ev->e0, ev->e1), "hipEventElapsedTime");
			hip_check(hipEventElapsedTime(&sim_ms, ev->e0, ev->es), "hipEventElapsedTime");
			r->kernel_ms = ms;
			r->simulate_ms = sim_ms;
			r->win.hold = blk;
			if (packed)
			{
				uint64_t placed = 0;
				std::memcpy(&placed, h + o_wv + align256(n_x * 8), 8);
				if (placed != n_x)
					throw error(MBG_E_INTERNAL, "packed window cells: the side list holds " + std::to_string(placed)
													+ " values, the compaction counted " + std::to_string(n_x));
				r->win.rec = (const uint32_t*)h;
				r->win.kbits = pk_bits;
				r->win.esc_min = esc_min;
				r->win.xval = (const uint64_t*)(h + o_wv);
				r->win.nx = n_x;
			}
			else
			{
				r->win.key = (const uint64_t*)h;
				r->win.val = (const uint64_t*)(h + o_wv);
			}
			r->win.n = n_cells;
			std::memcpy(r->ent.data(), h + o_ent, r->ent.size() * 8);
			r->fin.hold = blk;
			r->fin.key = (const uint64_t*)(h + o_fk);
			r->fin.val = (const uint64_t*)(h + o_fv);
			r->fin.n = n_fin;
			const uint32_t* rec = (const uint32_t*)(h + o_fp);
			r->fp_words.resize(n_fp * (size_t)W);
			r->fp_count.resize(n_fp);
			for (uint64_t i = 0; i < n_fp; i++)
			{
				const uint32_t* e = rec + i * (fp_stride - 1);
				r->fp_count[i] = e[0];
				std::memcpy(r->fp_words.data() + i * (size_t)W, e + 1, 4 * (size_t)W);
			}
			r->index_cells();
			r->sort_fixed();
			r->total_ms = (wall_seconds() - t0) * 1e3;
		};
		return res;
	}
}

} // namespace

std::unique_ptr<result> simulate_async(program& p, const sim_params& sp, hipStream_t stream)
{
	return run_impl(p, sp, stream, nullptr, nullptr);
}

std::unique_ptr<result> simulate(program& p, const sim_params& sp, hipStream_t stream)
{
	auto r = run_impl(p, sp, stream, nullptr, nullptr);
	r->wait();
	return r;
}

void window_probs_device(program& p, const sim_params& sp, const float* rates, double* probs, hipStream_t stream)
{
	if (!probs)
		throw error(MBG_E_ARG, "null probability buffer");
	if (!sp.rates.empty())
		throw error(MBG_E_ARG, "device window probabilities take their rates from device memory");
	(void)run_impl(p, sp, stream, rates, probs);
}

} // namespace mbg
