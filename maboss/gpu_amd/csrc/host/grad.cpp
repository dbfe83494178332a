// Gradient form: launches the likelihood-ratio kernel (mbg_grad_kernel.h)
// over the trajectories of a forward run.  No counterpart in the reference
// (its rates are compile-time constants and it has no derivative of any
// statistic).
#include <algorithm>
#include <cstring>

#include "../device/mbg_kernel_args.h"
#include "engine.h"
#include "timer.h"

namespace mbg
{

namespace
{

void fill_u64(program& p, hipStream_t stream, void* ptr, uint64_t words, uint64_t value)
{
	if (!words)
		return;
	p.load_util();
	void* args[] = { &ptr, &words, &value };
	const unsigned grid = (unsigned)std::min<uint64_t>(1024, (words + 255) / 256);
	hip_check(hipModuleLaunchKernel(p.fn_fill_u64, grid, 1, 1, 256, 1, 1, 0, stream, args, nullptr),
			  "hipModuleLaunchKernel(mbg_fill_u64)");
}

// LDS of one block: accumulators and 1/theta (doubles), the G table when
// staged, the jump table and the per-lane 16-bit jump counts.
size_t grad_lds_bytes(int np, uint64_t cells, bool g_lds, int n_nodes, int block)
{
	const size_t count_words = (size_t)(((np + 1) / 2) | 1); // per lane, odd (the kernel's layout)
	return 8 * (size_t)(2 * np) + (g_lds ? 8 * (size_t)cells : 0) + 2 * (size_t)((2 * n_nodes + 1) & ~1)
		   + 4 * count_words * (size_t)block;
}

} // namespace

const grad_host_tables& program::grad_tables()
{
	if (gtab.host)
		return gtab;
	const grad_code& gc = grad();
	const model& m = *mdl;
	const int N = m.n_nodes(), W = m.n_words(), np = gc.n_params;
	// enable masks (all on), istate, jump parameters, then the model's rates,
	// in pinned memory: the upload of each call is a plain DMA read of it
	gtab.en_bytes = ((size_t)W * 4 + 15) & ~(size_t)15;
	gtab.ist_bytes = ((size_t)N * 4 + 15) & ~(size_t)15;
	gtab.jp_bytes = (gc.jparam.size() * 2 + 15) & ~(size_t)15;
	gtab.bytes = gtab.en_bytes + gtab.ist_bytes + gtab.jp_bytes + (size_t)np * 4;
	std::shared_ptr<char> h = pool->get(gtab.bytes);
	char* t = h.get();
	std::memset(t, 0, gtab.bytes);
	std::memset(t, 0xff, (size_t)W * 4);
	for (int i = 0; i < N; i++)
		std::memcpy(t + gtab.en_bytes + 4 * (size_t)i, &m.nodes[i].istate, 4);
	std::memcpy(t + gtab.en_bytes + gtab.ist_bytes, gc.jparam.data(), gc.jparam.size() * 2);
	for (int k = 0; k < np; k++)
	{
		const float v = m.variables.at(gen.param_names[k]);
		std::memcpy(t + gtab.en_bytes + gtab.ist_bytes + gtab.jp_bytes + 4 * (size_t)k, &v, 4);
	}
	gtab.host = h;
	return gtab;
}

void simulate_grad(program& p, const grad_params& gp, const float* rates, const double* gprob, double* grad,
				   hipStream_t stream, grad_timing* timing, void* scratch)
{
	scope_timer tm("grad> simulate");
	const model& m = *p.mdl;
	if (scratch && !p.fn_grad)
		throw error(MBG_E_ARG, "the gradient kernel is not loaded: call mbg_grad_check first");
	const grad_code& gc = p.grad();
	if (m.n_words() > max_form_words)
		throw error(MBG_E_LIMIT, "the gradient form serves models of up to 256 nodes");
	const int n_obs = m.n_observed(), windows = m.windows(), N = m.n_nodes();
	if (n_obs > 24)
		throw error(MBG_E_LIMIT, "the gradient form takes a dense dL/dP: at most 24 observed nodes");
	if (!gprob || !grad)
		throw error(MBG_E_ARG, "null gradient buffer");
	const uint64_t n_traj = gp.n_traj < 0 ? (uint64_t)m.sample_count() : (uint64_t)gp.n_traj;
	const uint64_t seed = gp.seed < 0 ? m.seed() : (uint64_t)gp.seed;
	const int np = gc.n_params;
	const uint64_t cells = (uint64_t)windows << n_obs;
	hipFunction_t fn = p.grad_function();

	const grad_host_tables& gt = p.grad_tables();
	const size_t en_bytes = gt.en_bytes, ist_bytes = gt.ist_bytes, jp_bytes = gt.jp_bytes, tab_bytes = gt.bytes;
	// device workspace: the counters, then the tables; from the library's
	// static scratch array (no allocation) or from the program's buffers
	const size_t ctr_bytes = 256;
	char* ws = nullptr;
	if (scratch)
	{
		if (ctr_bytes + tab_bytes > k_noalloc_scratch_bytes)
			throw error(MBG_E_LIMIT, "the gradient workspace (" + std::to_string(ctr_bytes + tab_bytes)
										 + " bytes) exceeds the " + std::to_string(k_noalloc_scratch_bytes)
										 + "-byte scratch of the no-allocation entry point");
		ws = (char*)scratch;
	}
	else
		ws = (char*)p.b_gen.ensure(ctr_bytes + tab_bytes);
	auto* counters = (unsigned long long*)ws;
	char* dtab = ws + ctr_bytes;
	hip_check(hipMemcpyAsync(dtab, gt.host.get(), tab_bytes, hipMemcpyHostToDevice, stream), "hipMemcpyAsync");
	if (!rates)
		rates = (const float*)(dtab + en_bytes + ist_bytes + jp_bytes);
	fill_u64(p, stream, counters, MBG_NCOUNTERS, 0);
	fill_u64(p, stream, grad, (uint64_t)np, 0);

	// launch shape: the G table goes to LDS when small; block size from the LDS budget
	if (!p.lds_per_block)
	{
		hipDeviceProp_t prop;
		hip_check(hipGetDeviceProperties(&prop, p.device), "hipGetDeviceProperties");
		p.lds_per_block = prop.sharedMemPerBlock;
	}
	const size_t lds_max = std::min<size_t>(p.lds_per_block, 64 * 1024);
	const bool g_lds = cells * 8 <= 32 * 1024;
	int block = gp.block > 0 ? gp.block : 256;
	if (block % 64 != 0 || block > 256)
		throw error(MBG_E_ARG, "block_size must be a multiple of 64 up to 256 for the gradient kernel");
	if (gp.block <= 0)
		while (block > 64 && grad_lds_bytes(np, cells, g_lds, N, block) > lds_max)
			block -= 64;
	const size_t lds = grad_lds_bytes(np, cells, g_lds, N, block);
	if (lds > lds_max)
		throw error(MBG_E_LIMIT, "gradient kernel LDS exceeds the per-block limit");
	int per_cu = gp.per_cu;
	if (per_cu <= 0)
	{
		hip_check(hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, block, lds),
				  "hipModuleOccupancyMaxActiveBlocksPerMultiprocessor");
		per_cu = std::max(per_cu, 1);
	}
	const uint64_t need = (n_traj + (uint64_t)block - 1) / (uint64_t)block;
	const unsigned grid = (unsigned)std::max<uint64_t>(1, std::min<uint64_t>(need, (uint64_t)per_cu * p.num_cus));

	mbg_grad_args a;
	std::memset(&a, 0, sizeof(a));
	a.n_traj = n_traj;
	a.id_base = gp.first_id;
	a.seed_lo = (uint32_t)seed;
	a.seed_hi = (uint32_t)(seed >> 32);
	a.variant = gp.variant;
	a.max_steps = gp.max_steps ? gp.max_steps : (1u << 30);
	a.n_nodes = N;
	a.n_obs = n_obs;
	a.windows = windows;
	a.n_params = np;
	a.max_time = m.max_time();
	a.tick = m.time_tick();
	a.g_lds = g_lds ? 1 : 0;
	// 16 measured best on grad_30 (3.63 ms; 12: 3.67, 24: 3.67, 8: 3.84, 4: 4.42)
	a.fold_lanes = gp.fold_lanes > 0 ? std::min(gp.fold_lanes, 64) : 16;
	a.lds_bytes = (int)lds;
	a.istate = (const float*)(dtab + en_bytes);
	a.pv.rate = rates;
	a.pv.rate_lr = rates; // the entropy is not used by the gradient
	a.pv.up_en = (const uint32_t*)dtab;
	a.pv.dn_en = (const uint32_t*)dtab;
	a.gprob = gprob;
	a.jparam = (const unsigned short*)(dtab + en_bytes + ist_bytes);
	a.grad = grad;
	a.counters = counters;
	a.scale = 1.0 / ((double)n_traj * (double)a.tick);

	hipEvent_t e0 = nullptr, e1 = nullptr;
	if (timing)
	{
		hip_check(hipEventCreate(&e0), "hipEventCreate");
		hip_check(hipEventCreate(&e1), "hipEventCreate");
		hip_check(hipEventRecord(e0, stream), "hipEventRecord");
	}
	if (n_traj)
	{
		void* args[] = { &a };
		hip_check(hipModuleLaunchKernel(fn, grid, 1, 1, (unsigned)block, 1, 1, (unsigned)lds, stream, args, nullptr),
				  "hipModuleLaunchKernel(simulate_grad)");
	}
	if (timing)
	{
		hip_check(hipEventRecord(e1, stream), "hipEventRecord");
		unsigned long long ctr[MBG_NCOUNTERS];
		hip_check(hipMemcpyAsync(ctr, counters, sizeof(ctr), hipMemcpyDeviceToHost, stream), "hipMemcpyAsync");
		hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize");
		float ms = 0;
		hip_check(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
		(void)hipEventDestroy(e0);
		(void)hipEventDestroy(e1);
		timing->kernel_ms = ms;
		timing->transitions = ctr[MBG_CTR_TRANSITIONS];
		timing->block = block;
		timing->grid = (int)grid;
	}
}

} // namespace mbg
