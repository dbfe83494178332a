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

plan make_plan(const model& m, const sim_params& sp, uint64_t n_traj, int windows, int n_obs, bool wave_kernel)
{
	plan p;
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
	if (p.win_mode < 0)
		p.win_mode = win_dense_bytes <= 32 * 1024 ? MBG_HIST_LDS
					 : win_dense_bytes <= k_dense_limit ? MBG_HIST_DENSE
														: MBG_HIST_HASH;
	if (p.win_mode < 0 || p.win_mode > MBG_HIST_SORT)
		throw error(MBG_E_ARG, "unknown window mode");
	if ((p.win_mode == MBG_HIST_DENSE || p.win_mode == MBG_HIST_LDS) && win_dense_bytes > k_dense_limit)
		throw error(MBG_E_LIMIT, "dense window histogram too large; use the hashed or sorted mode");
	if (p.win_mode == MBG_HIST_LDS)
	{
		if (win_dense_bytes > k_lds_budget - lds)
			throw error(MBG_E_LIMIT, "window histogram does not fit in LDS");
		p.win_off = (int)(lds / 8);
		lds += win_dense_bytes;
	}

	p.fin_mode = sp.final_mode;
	if (p.fin_mode > MBG_HIST_LDS)
		throw error(MBG_E_ARG, "unknown final-state mode");
	if (p.fin_mode < 0)
		p.fin_mode = fin_dense_bytes <= k_lds_budget - lds ? MBG_HIST_LDS
					 : fin_dense_bytes <= (1ull << 30)	 ? MBG_HIST_DENSE
														 : MBG_HIST_HASH;
	if (p.fin_mode != MBG_HIST_HASH && fin_dense_bytes > k_dense_limit)
		throw error(MBG_E_LIMIT, "dense final-state histogram too large; use the hashed mode");
	if (p.fin_mode == MBG_HIST_LDS)
	{
		if (fin_dense_bytes > k_lds_budget - lds)
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
		p.tail_slots = cache_slots(k_lds_budget - lds, 16, wave_kernel ? 128 : 512, 64);
		p.tail_off = (int)(lds / 8);
		lds += (size_t)p.tail_slots * 16;
	}
	{
		// the wave kernel keeps a smaller cache: its LDS holds the node tables
		const size_t slot = 4 * (2 + (size_t)m.n_words());
		p.fp_slots = wave_kernel ? cache_slots(k_lds_budget - lds, slot, 16, 8) : cache_slots(k_lds_budget - lds, slot, 128, 16);
		p.fp_off = (int)(lds / 8);
		lds += ((size_t)p.fp_slots * slot + 7) & ~(size_t)7;
	}
	// block-private cache in front of a hashed final-state table: trajectories
	// that end in the same state add to one LDS slot
	if (p.fin_mode == MBG_HIST_HASH)
	{
		p.fin_cache_slots = cache_slots(k_lds_budget - lds, 16, wave_kernel ? 128 : 512, 64);
		p.fin_cache_off = (int)(lds / 8);
		lds += (size_t)p.fin_cache_slots * 16;
	}
	p.lds_bytes = (int)lds;

	if (p.win_mode == MBG_HIST_HASH || p.win_mode == MBG_HIST_SORT)
		if (n_obs + bits_for((uint64_t)windows) > 63)
			throw error(MBG_E_LIMIT, "observed nodes plus window bits exceed the 63-bit cell key");
	if (p.win_mode == MBG_HIST_HASH)
		p.win_cap = sp.window_capacity ? next_pow2(sp.window_capacity)
									   : next_pow2(std::min<uint64_t>(std::max<uint64_t>(n_traj * windows, 1 << 16),
																	   1ull << 28));
	else if (p.win_mode != MBG_HIST_SORT) // the sorted mode's pair capacity is the runner's
		p.win_cells = (uint64_t)windows << n_obs;
	if (p.fin_mode == MBG_HIST_HASH)
	{
		if (n_obs > 63)
			throw error(MBG_E_LIMIT, "more than 63 observed nodes");
		p.fin_cap = sp.final_capacity ? next_pow2(sp.final_capacity)
									  : next_pow2(std::min<uint64_t>(std::max<uint64_t>(2 * n_traj, 1 << 12), 1ull << 27));
	}
	else
		p.fin_cells = 1ull << n_obs;
	p.fp_cap = sp.fixed_capacity ? next_pow2(sp.fixed_capacity)
								 : next_pow2(std::min<uint64_t>(std::max<uint64_t>(2 * n_traj, 1 << 12), 1ull << 24));
	(void)m;
	return p;
}

} // namespace

void default_storage_modes(const model& m, int& win_mode, int& fin_mode)
{
	const plan pl = make_plan(m, sim_params {}, (uint64_t)m.sample_count(), m.windows(), m.n_observed(), false);
	win_mode = pl.win_mode;
	fin_mode = pl.fin_mode;
}

std::unique_ptr<result> simulate(program& p, const sim_params& sp, hipStream_t stream)
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
	for (auto& n : m.nodes)
		res->names.push_back(n.name);
	res->observed = m.observed_nodes();

	const bool use_wave = sp.kernel == 1 || (sp.kernel < 0 && W > max_lane_words);
	if (!use_wave && W > max_lane_words)
		throw error(MBG_E_LIMIT, "the lane kernel serves up to " + std::to_string(32 * max_lane_words) + " nodes");
	if (use_wave && p.opts.entropy == entropy_formula::reference)
		throw error(MBG_E_ARG, "the wave kernel computes the log-sum entropy only");
	if (!sp.rates.empty() && use_wave)
		throw error(MBG_E_LIMIT, "run-time rates are served by the lane kernel (up to 256 nodes)");
	if (!sp.rates.empty() && sp.rates.size() != p.gen.param_names.size())
		throw error(MBG_E_ARG, "expected " + std::to_string(p.gen.param_names.size()) + " rates");
	// run-time rates: a parameter view {rate, rate*log2(rate), enable masks} in device memory
	const mbg_param_view* d_view = nullptr;
	if (!sp.rates.empty())
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
		d_view = (const mbg_param_view*)base;
	}
	res->kernel = use_wave ? 1 : 0;
	plan pl = make_plan(m, sp, n_traj, windows, n_obs, use_wave);
	hipFunction_t fn = use_wave ? p.wave_function() : p.function(sp.rates.empty() ? 0 : 2, pl.win_mode, pl.fin_mode);
	// tail records for window storage in HBM (continuous time)
	const bool tails = !m.discrete_time() && pl.win_mode != MBG_HIST_LDS;
	uint64_t tail_cap = tails ? next_pow2(std::min<uint64_t>(std::max<uint64_t>(n_traj / 4, 1 << 12), 1 << 20)) : 0;
	// sorted window mode: pair capacity from the caller, the last run, or windows per trajectory
	uint64_t pair_cap = 0;
	if (pl.win_mode == MBG_HIST_SORT)
	{
		const uint64_t per = p.pairs_per_traj ? p.pairs_per_traj + p.pairs_per_traj / 8 + 1
											  : std::max<uint64_t>((uint64_t)windows, 16);
		pair_cap = sp.window_capacity ? sp.window_capacity : std::max<uint64_t>(n_traj * per, 1 << 16);
	}
	p.load_util();

	for (int attempt = 0;; attempt++)
	{
		// ---- storage
		auto* counters = (unsigned long long*)p.b_counters.ensure(MBG_NCOUNTERS * 8);
		hip_check(hipMemsetAsync(counters, 0, MBG_NCOUNTERS * 8, stream), "hipMemsetAsync");
		auto* ent = (unsigned long long*)p.b_ent.ensure((size_t)windows * 8);
		hip_check(hipMemsetAsync(ent, 0, (size_t)windows * 8, stream), "hipMemsetAsync");
		unsigned long long *win_vals = nullptr, *win_keys = nullptr, *fin_vals = nullptr, *fin_keys = nullptr;
		unsigned long long *pk = nullptr, *pv = nullptr;
		if (pl.win_mode == MBG_HIST_SORT)
		{
			pk = (unsigned long long*)p.b_pk.ensure(pair_cap * 8);
			pv = (unsigned long long*)p.b_pv2.ensure(pair_cap * 8);
		}
		else if (pl.win_mode == MBG_HIST_HASH)
		{
			win_keys = (unsigned long long*)p.b_win_keys.ensure(pl.win_cap * 16);
			hip_fill_slots(p, win_keys, pl.win_cap, stream);
		}
		else
		{
			win_vals = (unsigned long long*)p.b_win_vals.ensure(pl.win_cells * 8);
			hip_check(hipMemsetAsync(win_vals, 0, pl.win_cells * 8, stream), "hipMemsetAsync");
		}
		if (pl.fin_mode == MBG_HIST_HASH)
		{
			fin_keys = (unsigned long long*)p.b_fin_keys.ensure(pl.fin_cap * 16);
			hip_fill_slots(p, fin_keys, pl.fin_cap, stream);
		}
		else
		{
			fin_vals = (unsigned long long*)p.b_fin_vals.ensure(pl.fin_cells * 8);
			hip_check(hipMemsetAsync(fin_vals, 0, pl.fin_cells * 8, stream), "hipMemsetAsync");
		}
		unsigned long long* tail_tab = nullptr;
		if (tails)
		{
			tail_tab = (unsigned long long*)p.b_tail.ensure(tail_cap * 16);
			hip_fill_slots(p, tail_tab, tail_cap, stream);
		}
		const size_t fp_stride = 2 + (size_t)W;
		auto* fp = (unsigned int*)p.b_fp.ensure(pl.fp_cap * fp_stride * 4);
		hip_check(hipMemsetAsync(fp, 0, pl.fp_cap * fp_stride * 4, stream), "hipMemsetAsync");

		std::vector<float> istate;
		for (auto& n : m.nodes)
			istate.push_back(n.istate);
		auto* d_istate = (float*)p.b_istate.ensure(istate.size() * 4);
		hip_check(hipMemcpyAsync(d_istate, istate.data(), istate.size() * 4, hipMemcpyHostToDevice, stream),
				  "hipMemcpyAsync");

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
			hip_check(hipMemcpyAsync(img, wt.image.data(), wt.image.size(), hipMemcpyHostToDevice, stream),
					  "hipMemcpyAsync");
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
		int per_cu = sp.blocks_per_cu;
		if (per_cu <= 0)
		{
			hip_check(hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, block, lds_total),
					  "hipModuleOccupancyMaxActiveBlocksPerMultiprocessor");
			per_cu = std::max(per_cu, 1);
		}
		uint64_t blocks_needed = (n_traj + work_per_block - 1) / work_per_block;
		uint64_t grid = std::min<uint64_t>((uint64_t)p.num_cus * per_cu, std::max<uint64_t>(blocks_needed, 1));

		hipEvent_t e0, e1, es;
		hip_check(hipEventCreate(&e0), "hipEventCreate");
		hip_check(hipEventCreate(&e1), "hipEventCreate");
		hip_check(hipEventCreate(&es), "hipEventCreate");
		hip_check(hipEventRecord(e0, stream), "hipEventRecord");
		if (n_traj > 0)
		{
			void* args[] = { &a };
			hip_check(hipModuleLaunchKernel(fn, (unsigned)grid, 1, 1, block, 1, 1, lds_total, stream, args, nullptr),
					  "hipModuleLaunchKernel(simulate)");
			hip_check(hipEventRecord(es, stream), "hipEventRecord");
			if (tails)
			{
				const unsigned eg = (unsigned)std::min<uint64_t>(1024, (tail_cap + 255) / 256);
				hip_check(hipModuleLaunchKernel(p.fn_expand_tails, eg, 1, 1, 256, 1, 1, 0, stream, args, nullptr),
						  "hipModuleLaunchKernel(mbg_expand_tails)");
			}
		}
		std::vector<unsigned long long> ctr(MBG_NCOUNTERS);
		unsigned long long *uk = nullptr, *uv = nullptr, *ucount = nullptr;
		bool pair_over = false;
		if (pl.win_mode == MBG_HIST_SORT)
		{
			// the pair count sizes the sort
			hip_check(hipMemcpyAsync(ctr.data(), counters, MBG_NCOUNTERS * 8, hipMemcpyDeviceToHost, stream),
					  "hipMemcpyAsync");
			hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize");
			const uint64_t n_pairs = ctr[MBG_CTR_PAIRS];
			pair_over = n_pairs > pair_cap;
			if (!pair_over && n_pairs)
			{
				auto* ka = (unsigned long long*)p.b_pk_alt.ensure(pair_cap * 8);
				auto* va = (unsigned long long*)p.b_pv_alt.ensure(pair_cap * 8);
				ucount = (unsigned long long*)p.b_ucount.ensure(8);
				sort_reduce_pairs(p.b_sort_tmp, pk, pv, ka, va, n_pairs,
								  n_obs + bits_for((uint64_t)windows - 1) + 1, // + the padding bit
								  ucount, stream, &uk, &uv);
			}
			if (n_traj)
				p.pairs_per_traj = (n_pairs + n_traj - 1) / n_traj;
		}
		hip_check(hipEventRecord(e1, stream), "hipEventRecord");

		// ---- gather
		hip_check(hipMemcpyAsync(ctr.data(), counters, MBG_NCOUNTERS * 8, hipMemcpyDeviceToHost, stream),
				  "hipMemcpyAsync");
		hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize");
		float ms = 0, sim_ms = 0;
		hip_check(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
		if (n_traj > 0)
			hip_check(hipEventElapsedTime(&sim_ms, e0, es), "hipEventElapsedTime");
		(void)hipEventDestroy(e0);
		(void)hipEventDestroy(e1);
		(void)hipEventDestroy(es);

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
				pair_cap = ctr[MBG_CTR_PAIRS] + ctr[MBG_CTR_PAIRS] / 16 + 1;
			}
			if (win_over)
			{
				if (sp.window_capacity)
					throw error(MBG_E_OVERFLOW, "window hash table overflow; raise window_capacity");
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

		res->kernel_ms = ms;
		res->simulate_ms = sim_ms;
		res->window_pairs = pl.win_mode == MBG_HIST_SORT ? ctr[MBG_CTR_PAIRS] : 0;
		res->transitions = ctr[MBG_CTR_TRANSITIONS];
		res->finished = ctr[MBG_CTR_FINISHED];
		res->truncated = ctr[MBG_CTR_TRUNCATED];
		res->window_mode = pl.win_mode;
		res->final_mode = pl.fin_mode;
		res->window_capacity = pl.win_mode == MBG_HIST_SORT ? pair_cap : pl.win_cap;

		if (sp.device_only)
		{
			res->ent.assign(windows, 0);
			res->win.assign(windows, {});
			break;
		}

		res->ent.resize(windows);
		hip_check(hipMemcpy(res->ent.data(), ent, (size_t)windows * 8, hipMemcpyDeviceToHost), "hipMemcpy");

		res->win.assign(windows, {});
		if (pl.win_mode == MBG_HIST_SORT)
		{
			unsigned long long nu = 0;
			if (ucount)
				hip_check(hipMemcpy(&nu, ucount, 8, hipMemcpyDeviceToHost), "hipMemcpy");
			std::vector<unsigned long long> keys(nu), vals(nu);
			if (nu)
			{
				hip_check(hipMemcpy(keys.data(), uk, nu * 8, hipMemcpyDeviceToHost), "hipMemcpy");
				hip_check(hipMemcpy(vals.data(), uv, nu * 8, hipMemcpyDeviceToHost), "hipMemcpy");
			}
			// keys arrive sorted by (window, observed state)
			const uint64_t omask = n_obs >= 64 ? ~0ull : ((1ull << n_obs) - 1);
			for (uint64_t i = 0; i < nu; i++)
				if (keys[i] != MBG_PAIR_PAD_KEY)
					res->win[keys[i] >> n_obs].emplace_back(keys[i] & omask, vals[i]);
		}
		else if (pl.win_mode == MBG_HIST_HASH)
		{
			// the occupied cells, compacted and sorted by key on the device
			auto* tk = (unsigned long long*)p.b_pk.ensure(pl.win_cap * 8);
			auto* tv = (unsigned long long*)p.b_pv2.ensure(pl.win_cap * 8);
			auto* ka = (unsigned long long*)p.b_pk_alt.ensure(pl.win_cap * 8);
			auto* va = (unsigned long long*)p.b_pv_alt.ensure(pl.win_cap * 8);
			a.pair_keys = tk;
			a.pair_vals = tv;
			a.pair_cap = pl.win_cap;
			hip_check(hipMemsetAsync(counters + MBG_CTR_PAIRS, 0, 8, stream), "hipMemsetAsync");
			{
				void* args[] = { &a };
				const unsigned g = (unsigned)((pl.win_cap + 256 * 16 - 1) / (256 * 16)); // 16 slots per thread
				hip_check(hipModuleLaunchKernel(p.fn_table_pairs, g, 1, 1, 256, 1, 1, 0, stream, args, nullptr),
						  "hipModuleLaunchKernel(mbg_table_pairs)");
			}
			unsigned long long n_cells = 0;
			hip_check(hipMemcpyAsync(&n_cells, counters + MBG_CTR_PAIRS, 8, hipMemcpyDeviceToHost, stream),
					  "hipMemcpyAsync");
			hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize");
			unsigned long long nu = 0, *hk = nullptr, *hv = nullptr;
			if (n_cells)
			{
				auto* hcount = (unsigned long long*)p.b_ucount.ensure(8);
				sort_reduce_pairs(p.b_sort_tmp, tk, tv, ka, va, n_cells, n_obs + bits_for((uint64_t)windows - 1),
								  hcount, stream, &hk, &hv);
				hip_check(hipMemcpy(&nu, hcount, 8, hipMemcpyDeviceToHost), "hipMemcpy");
			}
			std::vector<unsigned long long> keys(nu), vals(nu);
			if (nu)
			{
				hip_check(hipMemcpy(keys.data(), hk, nu * 8, hipMemcpyDeviceToHost), "hipMemcpy");
				hip_check(hipMemcpy(vals.data(), hv, nu * 8, hipMemcpyDeviceToHost), "hipMemcpy");
			}
			const uint64_t omask = n_obs >= 64 ? ~0ull : ((1ull << n_obs) - 1);
			for (uint64_t i = 0; i < nu; i++)
				res->win[keys[i] >> n_obs].emplace_back(keys[i] & omask, vals[i]);
		}
		else
		{
			std::vector<unsigned long long> vals(pl.win_cells);
			hip_check(hipMemcpy(vals.data(), win_vals, pl.win_cells * 8, hipMemcpyDeviceToHost), "hipMemcpy");
			const uint64_t per = 1ull << n_obs;
			for (int w = 0; w < windows; w++)
				for (uint64_t s = 0; s < per; s++)
					if (uint64_t v = vals[(uint64_t)w * per + s])
						res->win[w].emplace_back(s, v);
		}

		if (pl.fin_mode == MBG_HIST_HASH)
		{
			std::vector<unsigned long long> tab(pl.fin_cap * 2);
			hip_check(hipMemcpy(tab.data(), fin_keys, pl.fin_cap * 16, hipMemcpyDeviceToHost), "hipMemcpy");
			for (uint64_t i = 0; i < pl.fin_cap; i++)
				if (tab[2 * i] != MBG_EMPTY_KEY && tab[2 * i + 1])
					res->fin.emplace_back(tab[2 * i], tab[2 * i + 1]);
			std::sort(res->fin.begin(), res->fin.end());
		}
		else
		{
			std::vector<unsigned long long> vals(pl.fin_cells);
			hip_check(hipMemcpy(vals.data(), fin_vals, pl.fin_cells * 8, hipMemcpyDeviceToHost), "hipMemcpy");
			for (uint64_t s = 0; s < pl.fin_cells; s++)
				if (vals[s])
					res->fin.emplace_back(s, vals[s]);
		}

		{
			std::vector<unsigned int> tab(pl.fp_cap * fp_stride);
			hip_check(hipMemcpy(tab.data(), fp, tab.size() * 4, hipMemcpyDeviceToHost), "hipMemcpy");
			for (uint64_t i = 0; i < pl.fp_cap; i++)
			{
				const unsigned int* e = &tab[i * fp_stride];
				if (e[0] == 2u && e[1])
					res->fixed.emplace_back(std::vector<uint32_t>(e + 2, e + 2 + W), e[1]);
			}
			// reference order: most significant word first (src/state.h:43-49)
			std::sort(res->fixed.begin(), res->fixed.end(), [](const auto& x, const auto& y) {
				for (int w = (int)x.first.size() - 1; w >= 0; w--)
					if (x.first[w] != y.first[w])
						return x.first[w] < y.first[w];
				return false;
			});
		}
		break;
	}
	res->total_ms = (wall_seconds() - t0) * 1e3;
	return res;
}

} // namespace mbg
