// Batched form: many variants of one model -- different rate vectors, mutant
// nodes forced on or off, trajectory counts and horizons -- simulated in one
// launch of the parametric lane kernel (simulate_batched).
//
// The reference has no counterpart: rates are baked into its generated code
// (src/generator.cpp:104-106, src/parser/expressions.cpp:211-214), one model
// runs per process (src/main.cpp:154-164) and mutants are a TODO
// (README.md:33).  Variant v draws from the stream tagged v, so its result is
// bit-identical to a plain run with variant = v and the same rates.
//
// The launch (statistics cleared by fill kernels, then the batched kernel) is
// captured in a hipGraph and replayed while the batch layout is unchanged.
// The launch shape (block size, blocks per CU, trajectories per lane of a
// work chunk) is chosen by timing the candidates on the batch itself, once
// per batch layout.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <numeric>
#include <string>
#include <sys/stat.h>
#include <unistd.h>

#include "../device/mbg_kernel_args.h"
#include "../device/mbg_numeric.h"
#include "engine.h"
#include "timer.h"

namespace mbg
{

namespace
{

uint64_t next_pow2(uint64_t v)
{
	uint64_t p = 1;
	while (p < v)
		p <<= 1;
	return p;
}

uint64_t mix(uint64_t h, uint64_t v)
{
	h ^= v + 0x9e3779b97f4a7c15ull + (h << 6) + (h >> 2);
	return h;
}

struct batch_layout
{
	int nv = 0;
	std::vector<int> windows;
	std::vector<float> max_time;
	std::vector<uint64_t> n, win_off, fin_off;
	std::vector<unsigned> ent_off;
	uint64_t win_cells = 0, ent_cells = 0, fin_cells = 0, fp_cap = 0;
	size_t fp_stride = 0;
};

void fill(program& p, hipStream_t stream, void* ptr, uint64_t words, uint64_t value)
{
	if (!words)
		return;
	void* args[] = { &ptr, &words, &value };
	const unsigned grid = (unsigned)std::min<uint64_t>(2048, (words + 255) / 256);
	hip_check(hipModuleLaunchKernel(p.fn_fill, grid, 1, 1, 256, 1, 1, 0, stream, args, nullptr),
			  "hipModuleLaunchKernel(mbg_fill_u64)");
}

uint64_t fnv(const std::string& s, uint64_t h)
{
	h ^= 1469598103934665603ull;
	for (unsigned char c : s)
	{
		h ^= c;
		h *= 1099511628211ull;
	}
	return h;
}

// The tuning cache: one line "block per_cu traj_per_lane lds_cells" per
// (generated code, batch shape, device); a missing or unreadable file is a miss.
bool read_tuned(const std::string& path, program::tuned& t)
{
	std::ifstream f(path);
	int b = 0, pc = 0, tpl = 0, sl = -1;
	if (!(f >> b >> pc >> tpl >> sl))
		return false;
	if (b <= 0 || b > 256 || b % 64 || pc <= 0 || tpl <= 0 || sl < 0 || sl > 4096)
		return false;
	t.block = b;
	t.per_cu = pc;
	t.tpl = tpl;
	t.slots = sl;
	return true;
}

void write_tuned(const std::string& path, const program::tuned& t)
{
	const std::string dir = path.substr(0, path.rfind('/'));
	(void)::mkdir(dir.c_str(), 0755);
	const std::string tmp = path + ".tmp" + std::to_string((unsigned long long)::getpid());
	{
		std::ofstream f(tmp);
		if (!f)
			return;
		f << t.block << " " << t.per_cu << " " << t.tpl << " " << t.slots << "\n";
		if (!f)
			return;
	}
	std::rename(tmp.c_str(), path.c_str());
}

} // namespace

std::vector<std::unique_ptr<result>> simulate_batched(program& p, const batch_params& bp,
													  const std::vector<variant_spec>& vs, hipStream_t stream,
													  batch_timing* timing)
{
	scope_timer tm("batched> simulate");
	const double t0 = wall_seconds();
	const model& m = *p.mdl;
	p.order_after_last_run(stream);
	const int W = m.n_words(), N = m.n_nodes(), n_obs = m.n_observed();
	if (W > max_form_words)
		throw error(MBG_E_LIMIT, "the batched form serves models of up to 256 nodes");
	if (n_obs > 20)
		throw error(MBG_E_LIMIT, "the batched form keeps dense per-variant windows: at most 20 observed nodes");
	std::vector<std::unique_ptr<result>> out;
	if (vs.empty())
		return out;
	hipFunction_t fn = p.function(1, 0, 0);
	const int np = (int)p.gen.param_names.size();
	const float tick = m.time_tick();
	const uint64_t seed = bp.seed < 0 ? m.seed() : (uint64_t)bp.seed;

	// ---- layout of the per-variant statistics
	batch_layout L;
	L.nv = (int)vs.size();
	uint64_t total_traj = 0;
	for (int v = 0; v < L.nv; v++)
	{
		const float mt = vs[v].max_time > 0 ? vs[v].max_time : m.max_time();
		const int win = (int)std::ceil(mt / tick);
		if (vs[v].n_traj < 0 || win <= 0 || win > (1 << 20))
			throw error(MBG_E_ARG, "variant " + std::to_string(v) + ": bad trajectory count or horizon");
		if (!vs[v].rates.empty() && (int)vs[v].rates.size() != np)
			throw error(MBG_E_ARG, "variant " + std::to_string(v) + ": expected " + std::to_string(np) + " rates");
		if (!vs[v].mut.empty() && (int)vs[v].mut.size() != N)
			throw error(MBG_E_ARG, "variant " + std::to_string(v) + ": expected " + std::to_string(N) + " mutations");
		L.max_time.push_back(mt);
		L.windows.push_back(win);
		L.n.push_back((uint64_t)vs[v].n_traj);
		L.win_off.push_back(L.win_cells);
		L.ent_off.push_back((unsigned)L.ent_cells);
		L.fin_off.push_back(L.fin_cells);
		L.win_cells += (uint64_t)win << n_obs;
		L.ent_cells += (uint64_t)win;
		L.fin_cells += 1ull << n_obs;
		total_traj += (uint64_t)vs[v].n_traj;
	}
	L.fp_stride = 3 + (size_t)W;
	L.fp_cap = next_pow2(std::min<uint64_t>(std::max<uint64_t>(2 * total_traj, 4096), 1ull << 24));

	// ---- per-variant tables: descriptors, then {rate, rate_lr, up_en, dn_en, istate} blocks
	std::vector<float> base_rates(np);
	for (int k = 0; k < np; k++)
		base_rates[k] = m.variables.at(p.gen.param_names[k]);
	const size_t block_bytes = (((size_t)2 * np + N + 2 * W) * 4 + 15) & ~(size_t)15;
	const size_t desc_bytes = ((size_t)L.nv * sizeof(mbg_batch_variant) + 15) & ~(size_t)15;
	std::vector<char> tab(desc_bytes + (size_t)L.nv * block_bytes, 0);
	char* dbase = (char*)p.b_bvar.ensure(tab.size());
	for (int v = 0; v < L.nv; v++)
	{
		const size_t off = desc_bytes + (size_t)v * block_bytes;
		float* rate = (float*)(tab.data() + off);
		float* rate_lr = rate + np;
		uint32_t* up_en = (uint32_t*)(rate_lr + np);
		uint32_t* dn_en = up_en + W;
		float* istate = (float*)(dn_en + W);
		const std::vector<float>& r = vs[v].rates.empty() ? base_rates : vs[v].rates;
		for (int k = 0; k < np; k++)
		{
			rate[k] = r[k];
			rate_lr[k] = rate_log2_rate(r[k]);
		}
		for (int w = 0; w < W; w++)
			up_en[w] = dn_en[w] = 0xffffffffu;
		for (int i = 0; i < N; i++)
		{
			istate[i] = m.nodes[i].istate;
			const int8_t mu = vs[v].mut.empty() ? 0 : vs[v].mut[i];
			if (mu > 0) // forced on: never switches off, starts on
			{
				dn_en[i / 32] &= ~(1u << (i % 32));
				istate[i] = 1.0f;
			}
			else if (mu < 0) // forced off
			{
				up_en[i / 32] &= ~(1u << (i % 32));
				istate[i] = 0.0f;
			}
		}
		mbg_batch_variant d;
		std::memset(&d, 0, sizeof(d));
		d.n_traj = L.n[v];
		d.win_off = L.win_off[v];
		d.fin_off = L.fin_off[v];
		d.ent_off = L.ent_off[v];
		d.windows = L.windows[v];
		d.max_time = L.max_time[v];
		d.ctr_off = 3u * (unsigned)v;
		d.istate = (const float*)(dbase + ((char*)istate - tab.data()));
		d.pv.rate = (const float*)(dbase + ((char*)rate - tab.data()));
		d.pv.rate_lr = (const float*)(dbase + ((char*)rate_lr - tab.data()));
		d.pv.up_en = (const uint32_t*)(dbase + ((char*)up_en - tab.data()));
		d.pv.dn_en = (const uint32_t*)(dbase + ((char*)dn_en - tab.data()));
		std::memcpy(tab.data() + (size_t)v * sizeof(mbg_batch_variant), &d, sizeof(d));
	}
	hip_check(hipMemcpyAsync(dbase, tab.data(), tab.size(), hipMemcpyHostToDevice, stream), "hipMemcpyAsync");

	// ---- storage
	auto* win = (unsigned long long*)p.b_bwin.ensure(std::max<uint64_t>(L.win_cells, 1) * 8);
	auto* ent = (unsigned long long*)p.b_bent.ensure(std::max<uint64_t>(L.ent_cells, 1) * 8);
	auto* fin = (unsigned long long*)p.b_bfin.ensure(std::max<uint64_t>(L.fin_cells, 1) * 8);
	const uint64_t fp_words = (L.fp_cap * L.fp_stride * 4 + 7) / 8;
	auto* fp = (unsigned int*)p.b_bfp.ensure(fp_words * 8);
	auto* vctr = (unsigned long long*)p.b_bctr.ensure((size_t)L.nv * 3 * 8);
	auto* counters = (unsigned long long*)p.b_counters.ensure(MBG_NCOUNTERS * 8);

	// ---- launch shape
	uint64_t tune_key = mix(mix(mix(0, (uint64_t)L.nv), total_traj), (uint64_t)p.num_cus);
	for (int v = 0; v < L.nv; v++)
		tune_key = mix(tune_key, ((uint64_t)L.windows[v] << 32) ^ L.n[v]);

	std::vector<mbg_batch_chunk> chunks;
	auto make_chunks = [&](int tpl) {
		chunks.clear();
		const uint64_t cs = 64ull * (uint64_t)tpl;
		std::vector<int> order(L.nv);
		std::iota(order.begin(), order.end(), 0);
		// longest horizons first, so the tail of the launch is short work
		std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return L.max_time[x] > L.max_time[y]; });
		for (int v : order)
			for (uint64_t b = 0; b < L.n[v]; b += cs)
				chunks.push_back(mbg_batch_chunk { (unsigned)v, 0u, b, std::min(L.n[v], b + cs) });
	};

	mbg_sim_args a;
	std::memset(&a, 0, sizeof(a));
	a.seed_lo = (uint32_t)seed;
	a.seed_hi = (uint32_t)(seed >> 32);
	a.max_steps = bp.max_steps ? bp.max_steps : (1u << 30);
	a.n_nodes = N;
	a.n_obs = n_obs;
	a.discrete = m.discrete_time();
	a.tick = tick;
	a.qscale = a.discrete ? 0x1p32f : 0x1p32f / tick;
	a.win_vals = win;
	a.win_ent = ent;
	a.fin_vals = fin;
	a.fp_table = fp;
	a.fp_mask = L.fp_cap - 1;
	a.counters = counters;
	a.refill_lanes = default_refill_lanes(N);
	a.bv = (const mbg_batch_variant*)dbase;
	a.vctr = vctr;

	auto enqueue = [&](int block, int per_cu, int tpl) {
		make_chunks(tpl);
		auto* dch = (mbg_batch_chunk*)p.b_bchunk.ensure(std::max<size_t>(chunks.size(), 1) * sizeof(mbg_batch_chunk));
		hip_check(hipMemcpyAsync(dch, chunks.data(), chunks.size() * sizeof(mbg_batch_chunk), hipMemcpyHostToDevice,
								 stream),
				  "hipMemcpyAsync");
		a.chunks = dch;
		a.n_chunks = (unsigned)chunks.size();
		const uint64_t waves = (uint64_t)p.num_cus * per_cu * (block / 64);
		const uint64_t grid = std::max<uint64_t>(1, std::min<uint64_t>((uint64_t)p.num_cus * per_cu,
																	   (chunks.size() + block / 64 - 1) / (block / 64)));
		(void)waves;
		return grid;
	};
	// the per-wave LDS cell cache keys cells by their 32-bit index in a variant
	const int max_windows = L.windows.empty() ? 0 : *std::max_element(L.windows.begin(), L.windows.end());
	const bool cache_ok = ((uint64_t)max_windows << n_obs) < 0xffffffffull;
	// per-wave LDS entropy sums when the longest horizon has at most 256 windows
	// (2 KB per wave: with the largest tuned cell cache, 1024 slots x 12 B, a
	// block of 4 waves stays within 56 KB)
	a.bent_windows = max_windows <= 256 ? max_windows : 0;
	auto cache_lds = [&](int block, int slots) {
		return (size_t)(block / 64) * ((size_t)slots * 12 + (size_t)a.bent_windows * 8);
	};
	auto launch = [&](int block, uint64_t grid, int slots) {
		a.bcache_slots = (unsigned)slots;
		fill(p, stream, counters, MBG_NCOUNTERS, 0);
		fill(p, stream, win, L.win_cells, 0);
		fill(p, stream, ent, L.ent_cells, 0);
		fill(p, stream, fin, L.fin_cells, 0);
		fill(p, stream, fp, fp_words, 0);
		fill(p, stream, vctr, (uint64_t)L.nv * 3, 0);
		if (a.n_chunks)
		{
			void* args[] = { &a };
			hip_check(hipModuleLaunchKernel(fn, (unsigned)grid, 1, 1, block, 1, 1, (unsigned)cache_lds(block, slots),
											stream, args, nullptr),
					  "hipModuleLaunchKernel(simulate_batched)");
		}
	};
	auto occupancy = [&](int block, int slots) {
		int per = 0;
		hip_check(hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&per, fn, block, cache_lds(block, slots)),
				  "hipModuleOccupancyMaxActiveBlocksPerMultiprocessor");
		return std::max(per, 1);
	};
	auto timed = [&](auto&& body) {
		hipEvent_t e0, e1;
		hip_check(hipEventCreate(&e0), "hipEventCreate");
		hip_check(hipEventCreate(&e1), "hipEventCreate");
		hip_check(hipEventRecord(e0, stream), "hipEventRecord");
		body();
		hip_check(hipEventRecord(e1, stream), "hipEventRecord");
		hip_check(hipEventSynchronize(e1), "hipEventSynchronize");
		float ms = 0;
		hip_check(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
		(void)hipEventDestroy(e0);
		(void)hipEventDestroy(e1);
		return (double)ms;
	};

	if (bp.lds_cells > 0 && (bp.lds_cells < 64 || bp.lds_cells > 4096 || (bp.lds_cells & (bp.lds_cells - 1))))
		throw error(MBG_E_ARG, "lds_cells must be -1, 0 or a power of two from 64 to 4096");
	if (bp.lds_cells > 0 && !cache_ok)
		throw error(MBG_E_LIMIT, "the LDS cell cache needs windows x 2^n_obs below 2^32 per variant");
	int block = bp.block, per_cu = bp.per_cu, tpl = bp.tpl, slots = bp.lds_cells < 0 ? 0 : bp.lds_cells;
	bool tuned_cached = false;
	if (!block || !tpl)
	{
		// the tuned shape of this batch: this program's, else the on-disk
		// tuning cache's (keyed by the generated code and the batch shape),
		// else a timed search over block x trajectories per lane x LDS layout
		uint64_t fkey = fnv(p.gen.source, tune_key ^ (cache_ok ? 1u : 0u));
		char fname[64];
		std::snprintf(fname, sizeof(fname), "/%016llx.tune", (unsigned long long)fkey);
		const std::string path = code_cache_dir() + fname;
		if (p.batch_tuned.key != tune_key)
		{
			program::tuned t;
			if (read_tuned(path, t))
			{
				t.key = tune_key;
				t.cached = true;
				p.batch_tuned = t;
			}
		}
		if (p.batch_tuned.key != tune_key)
		{
			// autotuner: time every candidate on this batch, keep the fastest
			double best = 1e300;
			for (int sl : { 0, 256, 1024 })
			{
				if (sl && !cache_ok)
					continue;
				for (int b : { 64, 128, 256 })
					for (int t : { 1, 2, 4, 6, 8, 12 })
					{
						const int pc = occupancy(b, sl);
						const uint64_t g = enqueue(b, pc, t);
						const double ms = timed([&] { launch(b, g, sl); });
						if (ms < best)
						{
							best = ms;
							p.batch_tuned = program::tuned { tune_key, b, pc, t, sl, false };
						}
					}
			}
			write_tuned(path, p.batch_tuned);
		}
		tuned_cached = p.batch_tuned.cached;
		block = block ? block : p.batch_tuned.block;
		tpl = tpl ? tpl : p.batch_tuned.tpl;
		per_cu = per_cu ? per_cu : p.batch_tuned.per_cu;
		if (!bp.lds_cells)
			slots = cache_ok ? p.batch_tuned.slots : 0;
	}
	if (block <= 0 || block > 256 || block % 64)
		throw error(MBG_E_ARG, "batched block size must be 64, 128, 192 or 256");
	if (!per_cu)
		per_cu = occupancy(block, slots);
	const uint64_t grid = enqueue(block, per_cu, tpl);

	double ms = 0;
	bool used_graph = false;
	if (bp.graph)
	{
		uint64_t gkey = mix(mix(mix(mix(tune_key, (uint64_t)block), (uint64_t)per_cu), (uint64_t)tpl), (uint64_t)slots);
		gkey = mix(mix(mix(gkey, (uint64_t)(uintptr_t)win), (uint64_t)(uintptr_t)fp), (uint64_t)(uintptr_t)dbase);
		gkey = mix(mix(gkey, (uint64_t)(uintptr_t)a.chunks), seed);
		gkey = mix(gkey, a.max_steps);
		if (!p.batch_graph || p.batch_graph_key != gkey)
		{
			if (p.batch_graph)
			{
				(void)hipGraphExecDestroy(p.batch_graph);
				p.batch_graph = nullptr;
			}
			hipStream_t cs;
			hip_check(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking), "hipStreamCreateWithFlags");
			hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize");
			hipGraph_t g;
			hip_check(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
			hipStream_t saved = stream;
			stream = cs;
			launch(block, grid, slots);
			stream = saved;
			hip_check(hipStreamEndCapture(cs, &g), "hipStreamEndCapture");
			hip_check(hipGraphInstantiate(&p.batch_graph, g, nullptr, nullptr, 0), "hipGraphInstantiate");
			(void)hipGraphDestroy(g);
			(void)hipStreamDestroy(cs);
			p.batch_graph_key = gkey;
		}
		ms = timed([&] { hip_check(hipGraphLaunch(p.batch_graph, stream), "hipGraphLaunch"); });
		used_graph = true;
	}
	else
		ms = timed([&] { launch(block, grid, slots); });

	// ---- occupied cells compacted on the device (in key order: the dense
	// arrays are compacted in index order), then one copy into pinned memory
	auto* cnt = (unsigned long long*)p.b_bcnt.ensure(4 * 8);
	unsigned long long *ck = nullptr, *cv = nullptr, *fk = nullptr, *fv = nullptr;
	unsigned int* fpc = nullptr;
	const size_t rec = L.fp_stride - 1; // {count, state words, variant}
	if (!bp.device_only)
	{
		hip_check(hipMemsetAsync(cnt, 0, 4 * 8, stream), "hipMemsetAsync");
		ck = (unsigned long long*)p.b_bck.ensure(std::max<uint64_t>(L.win_cells, 1) * 8);
		cv = (unsigned long long*)p.b_bcv.ensure(std::max<uint64_t>(L.win_cells, 1) * 8);
		compact_dense(p.b_sort_tmp, win, L.win_cells, ck, cv, cnt + 0, stream);
		fk = (unsigned long long*)p.b_bfk.ensure(std::max<uint64_t>(L.fin_cells, 1) * 8);
		fv = (unsigned long long*)p.b_bfv.ensure(std::max<uint64_t>(L.fin_cells, 1) * 8);
		compact_dense(p.b_sort_tmp, fin, L.fin_cells, fk, fv, cnt + 1, stream);
		fpc = (unsigned int*)p.b_bfpc.ensure(L.fp_cap * rec * 4);
		p.load_util();
		{
			int stride = (int)L.fp_stride;
			unsigned long long* c2 = cnt + 2;
			void* args[] = { &fp, &L.fp_cap, &stride, &fpc, &L.fp_cap, &c2 };
			const unsigned g = (unsigned)std::min<uint64_t>(4096, (L.fp_cap + 255) / 256);
			hip_check(hipModuleLaunchKernel(p.fn_fp_compact, g, 1, 1, 256, 1, 1, 0, stream, args, nullptr),
					  "hipModuleLaunchKernel(mbg_fp_compact)");
		}
	}
	if (!p.mailbox)
		p.mailbox = p.pool->get(4096);
	auto* mb = (unsigned long long*)p.mailbox.get();
	std::vector<unsigned long long> vc((size_t)L.nv * 3);
	std::shared_ptr<char> vblk = p.pool->get(vc.size() * 8 + 256);
	hip_check(hipMemcpyAsync(mb, counters, MBG_NCOUNTERS * 8, hipMemcpyDeviceToHost, stream), "hipMemcpyAsync");
	hip_check(hipMemcpyAsync(mb + 16, cnt, 4 * 8, hipMemcpyDeviceToHost, stream), "hipMemcpyAsync");
	hip_check(hipMemcpyAsync(vblk.get(), vctr, vc.size() * 8, hipMemcpyDeviceToHost, stream), "hipMemcpyAsync");
	hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize");
	std::memcpy(vc.data(), vblk.get(), vc.size() * 8);
	unsigned long long ctr[MBG_NCOUNTERS];
	std::memcpy(ctr, mb, sizeof(ctr));
	if (ctr[MBG_CTR_FP_TIMEOUT])
		throw error(MBG_E_INTERNAL, "fixed-point table: a slot was never published");
	if (ctr[MBG_CTR_FP_OVERFLOW])
		throw error(MBG_E_OVERFLOW, "fixed-point table overflow in the batched run");

	uint64_t all_trans = 0;
	for (int v = 0; v < L.nv; v++)
		all_trans += vc[3 * v];
	if (timing)
	{
		timing->kernel_ms = ms;
		timing->transitions = all_trans;
		timing->block = block;
		timing->per_cu = per_cu;
		timing->tpl = tpl;
		timing->lds_cells = slots;
		timing->graph = used_graph;
		timing->tuned_cached = tuned_cached;
	}

	const uint64_t n_cells = bp.device_only ? 0 : mb[16], n_fin = bp.device_only ? 0 : mb[17],
				   n_fp = bp.device_only ? 0 : mb[18];
	const size_t o_cv = (n_cells * 8 + 255) & ~(size_t)255, o_ent = o_cv + ((n_cells * 8 + 255) & ~(size_t)255);
	const size_t o_fk = o_ent + ((L.ent_cells * 8 + 255) & ~(size_t)255), o_fv = o_fk + ((n_fin * 8 + 255) & ~(size_t)255);
	const size_t o_fp = o_fv + ((n_fin * 8 + 255) & ~(size_t)255), total_bytes = o_fp + n_fp * rec * 4;
	std::shared_ptr<char> blk;
	if (!bp.device_only)
	{
		blk = p.pool->get(total_bytes);
		char* h = blk.get();
		auto d2h = [&](size_t off, const void* src, size_t bytes) {
			if (bytes)
				hip_check(hipMemcpyAsync(h + off, src, bytes, hipMemcpyDeviceToHost, stream), "hipMemcpyAsync");
		};
		d2h(0, ck, n_cells * 8);
		d2h(o_cv, cv, n_cells * 8);
		d2h(o_ent, ent, L.ent_cells * 8);
		d2h(o_fk, fk, n_fin * 8);
		d2h(o_fv, fv, n_fin * 8);
		d2h(o_fp, fpc, n_fp * rec * 4);
		hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize");
	}
	p.mark_run_done(stream);

	// ---- gather per variant: views into the shared block
	cell_view all_win, all_fin;
	std::vector<std::vector<uint64_t>> fp_of(L.nv);
	if (!bp.device_only)
	{
		const char* h = blk.get();
		all_win.hold = blk;
		all_win.key = (const uint64_t*)h;
		all_win.val = (const uint64_t*)(h + o_cv);
		all_win.n = n_cells;
		all_fin.hold = blk;
		all_fin.key = (const uint64_t*)(h + o_fk);
		all_fin.val = (const uint64_t*)(h + o_fv);
		all_fin.n = n_fin;
		const uint32_t* r = (const uint32_t*)(h + o_fp);
		for (uint64_t i = 0; i < n_fp; i++)
		{
			const uint32_t v = r[i * rec + 1 + W];
			if (v < (unsigned)L.nv)
				fp_of[v].push_back(i);
		}
	}
	auto sub_view = [](const cell_view& all, uint64_t first, uint64_t end) {
		return all.slice(all.lower_bound(first), all.lower_bound(end), first);
	};
	const uint64_t per = 1ull << n_obs;
	for (int v = 0; v < L.nv; v++)
	{
		auto r = std::make_unique<result>();
		r->windows = L.windows[v];
		r->n_obs = n_obs;
		r->n_words = W;
		r->discrete = m.discrete_time();
		r->tick = tick;
		r->max_time = L.max_time[v];
		r->n_traj = L.n[v];
		r->norm = (double)L.n[v];
		for (auto& nd : m.nodes)
			r->names.push_back(nd.name);
		r->observed = m.observed_nodes();
		r->transitions = vc[3 * v];
		r->finished = vc[3 * v + 1];
		r->truncated = vc[3 * v + 2];
		r->kernel_ms = ms;
		r->window_mode = MBG_HIST_DENSE;
		r->final_mode = MBG_HIST_DENSE;
		r->ent.assign(L.windows[v], 0);
		if (!bp.device_only)
		{
			r->win = sub_view(all_win, L.win_off[v], L.win_off[v] + ((uint64_t)L.windows[v] << n_obs));
			r->fin = sub_view(all_fin, L.fin_off[v], L.fin_off[v] + per);
			std::memcpy(r->ent.data(), blk.get() + o_ent + 8 * (size_t)L.ent_off[v], 8 * (size_t)L.windows[v]);
			const uint32_t* rr = (const uint32_t*)(blk.get() + o_fp);
			for (uint64_t i : fp_of[v])
			{
				r->fp_count.push_back(rr[i * rec]);
				r->fp_words.insert(r->fp_words.end(), rr + i * rec + 1, rr + i * rec + 1 + W);
			}
			r->sort_fixed();
		}
		r->index_cells();
		out.push_back(std::move(r));
	}
	const double total = (wall_seconds() - t0) * 1e3;
	for (auto& r : out)
		r->total_ms = total;
	if (timing)
		timing->total_ms = total;
	return out;
}

} // namespace mbg
