// Fused lane-per-trajectory simulate kernels for gfx950.
//
// One lane owns one trajectory at a time, with its W state words, time and
// step counter in VGPRs.  Each SSA step (src/jit_kernels/simulation.cu:
// 106-148) feeds its dwell slice straight into the window histogram
// (src/jit_kernels/window_average_small.cu:50-136), without writing a
// trajectory record to HBM; the end of a trajectory feeds the final-state
// histogram (src/jit_kernels/final_states.cu:9-21) and the fixed-point table
// (src/statistics/fixed_states.cu:84-125).
//
// All histogram cells are unsigned 64-bit fixed point (dwell * 2^32 / tick,
// rounded to nearest), so the result does not depend on the order in which
// lanes add: it is the same bits on every run, on any launch shape, and in
// the host oracle.  Consecutive contributions of one lane to the same
// (window, observed state) cell are summed in registers before one atomic.
//
// `simulate` (plain form) is persistent: lanes pull trajectory ids from a
// wave-aggregated atomic work queue until it is drained, so uneven lengths do
// not leave lanes idle.  `simulate_batched` (parametric build) runs many
// model variants in one launch: a wave takes a chunk of ids of one variant,
// so the variant's rate table is wave-uniform, and its lanes refill from the
// chunk until it is exhausted.
//
// Compiled by hipRTC per model, in one translation unit with the generated
// mbg_eval / mbg_eval_p / mbg_obs_index, for MBG_KW state words and a
// flavour MBG_PARAM: 0 plain (rates baked into mbg_eval), 1 batched,
// 2 plain with the rates read at run time (mbg_eval_p through a.pv).
#if !defined(__HIPCC_RTC__)
#include <hip/hip_runtime.h>
#endif

#include "mbg_kernel_args.h"
#include "mbg_lane_core.h"
#include "mbg_stats_device.h"

#ifndef MBG_KW
#error "MBG_KW (state words) must be defined"
#endif

#ifndef MBG_PARAM
#define MBG_PARAM 0
#endif

using namespace mbg_dev;

// Occupancy the kernel is compiled for (waves per SIMD), set per model.
#if defined(MBG_LANE_WAVES) && MBG_LANE_WAVES > 0
#define MBG_LANE_OCCUPANCY __attribute__((amdgpu_waves_per_eu(MBG_LANE_WAVES, 8)))
#else
#define MBG_LANE_OCCUPANCY
#endif

namespace
{

constexpr int W = MBG_KW;
// trajectory ids a wave reserves at a time (one queue atomic per chunk)
constexpr uint64_t k_id_chunk = 64;

// Wave-reduced counter update: transitions, finished and truncated
// trajectories go to their three destinations.
__device__ __forceinline__ void wave_add_counters(unsigned long long* d_trans, unsigned long long* d_fin,
												  unsigned long long* d_trunc, uint64_t ntrans, uint64_t nfin,
												  uint64_t ntrunc)
{
	for (int off = 32; off > 0; off >>= 1)
	{
		ntrans += __shfl_xor(ntrans, off);
		nfin += __shfl_xor(nfin, off);
		ntrunc += __shfl_xor(ntrunc, off);
	}
	if (__lane_id() == 0)
	{
		if (ntrans)
			atomicAdd(d_trans, (unsigned long long)ntrans);
		if (nfin)
			atomicAdd(d_fin, (unsigned long long)nfin);
		if (ntrunc)
			atomicAdd(d_trunc, (unsigned long long)ntrunc);
	}
}

} // namespace

#if MBG_PARAM != 1

// The plain kernel is compiled for the run's window and final-state storage
// (MBG_HIST_*): only that storage's code is in it.
#if !defined(MBG_WIN_MODE) || !defined(MBG_FIN_MODE)
#error "MBG_WIN_MODE and MBG_FIN_MODE select the storage compiled into the plain kernel"
#endif

#if MBG_PARAM == 2
#define MBG_EVAL_STEP(s, u2, e) mbg_eval_p(s, u2, &pv, e, nullptr)
#else
#define MBG_EVAL_STEP(s, u2, e) mbg_eval(s, u2, e)
#endif

extern "C" __global__ void __launch_bounds__(256) MBG_LANE_OCCUPANCY simulate(mbg_sim_args a)
{
	extern __shared__ unsigned long long lds[];
	for (int i = threadIdx.x; i < a.lds_bytes / 8; i += blockDim.x)
		lds[i] = 0ull;
	// initial-state probabilities in LDS: a refill then reads no global
	// memory, so it does not wait for the window-cell atomics in flight
	__shared__ float istate[32 * W];
	for (int i = threadIdx.x; i < a.n_nodes; i += blockDim.x)
		istate[i] = a.istate[i];
#if MBG_PARAM == 0
	mbg_eval_setup();
#else
	// the view's pointers in SGPRs (scalar load of the read-only view)
	const mbg_param_view pv = *(const MBG_CONST mbg_param_view*)a.pv;
	mbg_eval_p_setup(&pv);
#endif
	__syncthreads();

	const lane_hist_t<MBG_WIN_MODE, MBG_FIN_MODE> hist { lds, a };
	const int lane = __lane_id();
	const int wave = threadIdx.x >> 6;
	const uint32_t B0 = ((uint32_t)a.n_nodes + 3u) >> 2;
	auto fe = [&](int w, uint64_t q) { hist.ent(w, q); };
	auto ft = [&](int w, uint64_t o) { hist.tail(w, o); };
	const bool use_tails = a.tail_table != nullptr;
	// window cells completed by the wave's records: staged in LDS, 64 at a
	// time to the sink (sorted mode: pair chunks in HBM; hashed mode: one table
	// insert per lane, run as steps that each wait for the previous batch's
	// memory operation only), or straight into the histogram
	unsigned long long* cur = lds + a.pair_lds_off + 4 * wave;
	uint8_t* const own = (uint8_t*)(lds + a.own_lds_off) + 64 * wave;
	const bool hashed = MBG_WIN_MODE == MBG_HIST_HASH;
	const bool staged = (MBG_WIN_MODE == MBG_HIST_SORT || hashed) && a.pair_stage;
	pair_stage_reg stage { lds + a.pair_stage_off + 2 * a.pair_stage * wave, a.pair_stage };
	hash_inflight ins;
	struct
	{
		const mbg_sim_args& a;
		unsigned long long* cur;
		hash_inflight& ins;
		__device__ void step()
		{
			if (MBG_WIN_MODE == MBG_HIST_HASH)
				ins.step(a);
		}
		__device__ bool ready() const { return MBG_WIN_MODE != MBG_HIST_HASH || !ins.busy(); }
		__device__ void take(uint64_t key, uint64_t q)
		{
			if (MBG_WIN_MODE == MBG_HIST_HASH)
				ins.issue(a.win_table, a.win_mask, key, q);
			else if (MBG_WIN_MODE == MBG_HIST_SORT)
				append_pair_chunked(cur, a.pair_keys, a.pair_vals, a.pair_cap, a.counters + MBG_CTR_PAIRS,
									(uint64_t)a.pair_chunk, key, q);
		}
	} sink { a, cur, ins };
	auto emit = [&](bool has, int w, uint64_t o, uint64_t q) {
		if (staged)
			stage.add(has, ((uint64_t)w << a.n_obs) | o, q, sink);
		else if (has)
			hist.win(w, o, q);
	};

	uint32_t s[W];
#pragma unroll
	for (int w = 0; w < W; ++w)
		s[w] = 0u;
	float t = 0.0f;
	uint32_t k = 0;
	uint64_t id = 0, obs = 0; // obs: observed index of s, kept up to date at each flip
	bool active = false, exhausted = false;
	pending pend;
	// Models of up to 64 nodes: the lane's last fixed point and the
	// trajectories that ended in it since, inserted when the lane ends in
	// another one and at the end (trajectories of such models often end in
	// one attractor, and each insert runs one lane of the wave at a time).
	constexpr bool fp_pend = W <= 2;
	uint32_t pfs[W];
#pragma unroll
	for (int w = 0; w < W; ++w)
		pfs[w] = 0u;
	uint32_t pcnt = 0;
	// the block's {transitions, finished, truncated} counters, after the
	// waves' item-owner bytes; a trajectory adds its steps when it ends
	unsigned long long* const bctr = (unsigned long long*)((uint8_t*)(lds + a.own_lds_off) + 64 * (blockDim.x >> 6));

	id_chunks queue;
	for (;;)
	{
		// Refill lanes whose trajectory ended, from the wave's chunk of ids, in
		// whole rounds: one round of the initial state's Philox blocks starts
		// up to 64 / ceil(N/4) trajectories for the cost of one, so the wave
		// waits until refill_lanes lanes are idle, or none is running.
		const uint64_t need = __ballot(!active && !exhausted);
		if (need && ((uint64_t)__popcll(need) >= (uint64_t)a.refill_lanes || __ballot(active) == 0))
		{
			const uint64_t mine = queue.take(need, a.counters + MBG_CTR_QUEUE, a.n_traj, k_id_chunk);
			bool fresh = false;
			if (!active && !exhausted)
			{
				if (mine < a.n_traj)
				{
					id = a.id_base + mine;
					t = 0.0f;
					k = 0;
					active = true;
					fresh = true;
				}
				else
					exhausted = true;
			}
			init_states_wave<W>(s, fresh, id, a.variant, a.seed_lo, a.seed_hi, a.n_nodes, istate);
			if (fresh)
				obs = mbg_obs_index(s);
		}
		if (__ballot(active) == 0)
			break;

		bool ended = false, fixed = false;
		mbg_eval_t e;
		float tp = 0.0f, H = 0.0f;
		if (active)
		{
			float u1, u2;
			step_uniforms(a.discrete, B0, k, id, a.variant, a.seed_lo, a.seed_hi, u1, u2);
			MBG_EVAL_STEP(s, u2, &e);
			tp = t;
			t = advance_time(a.discrete, t, e.total, u1, a.tick, a.max_time, fixed);
			H = fixed ? 0.0f : e.H;
		}
		record_wave(active, a.discrete, tp, t, obs, H, a.tick, a.qscale, a.windows, pend, emit, fe, ft,
					use_tails && t >= a.max_time, own);
		if (active)
		{
			k++;
			if (t >= a.max_time || k >= a.max_steps)
			{
				ended = true;
				atomicAdd(bctr, (unsigned long long)k);
				atomicAdd(bctr + (t >= a.max_time ? 1 : 2), 1ull);
			}
			else
			{
#pragma unroll
				for (int w = 0; w < W; ++w)
					s[w] ^= ((e.flip >> 5) == w) ? (1u << (e.flip & 31)) : 0u;
				obs ^= mbg_obs_flip(e.flip);
			}
		}

		// Trajectory end: final state and fixed points (uniform region).
		if (__ballot(ended))
		{
			if (ended && t >= a.max_time)
				hist.fin(obs);
			if (fp_pend)
			{
				const bool fx = ended && fixed;
				bool same = fx && pcnt != 0u;
#pragma unroll
				for (int w = 0; w < W; ++w)
					same = same && pfs[w] == s[w];
				uint64_t fm = __ballot(fx && !same && pcnt != 0u);
				while (fm)
				{
					const int l = __ffsll((unsigned long long)fm) - 1;
					if (lane == l)
						hist.fixed<W>(pfs, pcnt);
					fm &= fm - 1ull;
				}
				if (same)
					pcnt++;
				else if (fx)
				{
#pragma unroll
					for (int w = 0; w < W; ++w)
						pfs[w] = s[w];
					pcnt = 1u;
				}
			}
			else
			{
				uint64_t fm = __ballot(ended && fixed);
				while (fm)
				{
					const int l = __ffsll((unsigned long long)fm) - 1;
					if (lane == l)
						hist.fixed<W>(s);
					fm &= fm - 1ull;
				}
			}
			if (ended)
				active = false;
		}
	}

	if (fp_pend)
	{
		uint64_t fm = __ballot(pcnt != 0u);
		while (fm)
		{
			const int l = __ffsll((unsigned long long)fm) - 1;
			if (lane == l)
				hist.fixed<W>(pfs, pcnt);
			fm &= fm - 1ull;
		}
	}

	// the lanes' last pending cells, then the rest of the stage / the inserts in flight
	emit(pend.pw >= 0, pend.pw, pend.pobs, pend.pq);
	if (pend.pew >= 0)
		fe(pend.pew, pend.peq);
	if (staged)
		stage.drain(sink);
	if (hashed)
		ins.finish(a);
	if (MBG_WIN_MODE == MBG_HIST_SORT)
		pad_pair_chunk(lds + a.pair_lds_off + 4 * (threadIdx.x >> 6), a.pair_keys, a.pair_vals, a.pair_cap);

	// Flush the block's counters, LDS histograms and caches.
	__syncthreads();
	if (threadIdx.x < 3 && bctr[threadIdx.x])
		atomicAdd(a.counters + (threadIdx.x == 0 ? MBG_CTR_TRANSITIONS
											: threadIdx.x == 1 ? MBG_CTR_FINISHED : MBG_CTR_TRUNCATED),
				  bctr[threadIdx.x]);
	flush_block_hist(lds, a);
	flush_block_fp<W>(lds, a);
}

#else // MBG_PARAM == 1

#define MBG_BCACHE_EMPTY 0xffffffffu

// Batched variants.  Statistics of variant v live at its offsets in the
// dense window / entropy / final arrays; fixed points share one table keyed
// by (state, variant); the stream of variant v is the plain run's stream
// with variant tag v, so variant v reproduces a plain run bit for bit.
extern "C" __global__ void __launch_bounds__(256) MBG_LANE_OCCUPANCY simulate_batched(mbg_sim_args a)
{
	const int lane = __lane_id();
	const uint32_t B0 = ((uint32_t)a.n_nodes + 3u) >> 2;
	// per wave: the lane that owns each item of a cooperative window pass
	__shared__ uint8_t own_tab[4][64];
	uint8_t* const own = own_tab[(threadIdx.x >> 6) & 3];
	// per wave: the initial-state probabilities of its current variant (a
	// refill then reads no global memory, so it does not wait for the
	// histogram atomics in flight)
	__shared__ float ist_tab[4][32 * W];
	float* const istate = ist_tab[(threadIdx.x >> 6) & 3];
	// per wave in dynamic LDS: the window-cell cache, S values (u64) then S
	// keys (u32), cell index (w << n_obs) | obs within the chunk's variant;
	// then EW entropy sums of the chunk's windows, added to the variant's
	// slice when the chunk is done (every lane of a wave adds to the same few
	// windows, which as global atomics cost 16 % of the batched launch)
	extern __shared__ unsigned long long bcache[];
	const uint32_t S = a.bcache_slots;
	const uint32_t EW = (uint32_t)a.bent_windows;
	const int sbits = S ? 31 - __builtin_clz(S) : 0;
	unsigned long long* const cval = bcache + (size_t)(threadIdx.x >> 6) * (S + S / 2 + EW);
	uint32_t* const ckey = (uint32_t*)(cval + S);
	unsigned long long* const went = cval + S + S / 2;
	for (uint32_t i = (uint32_t)lane; i < S; i += 64u)
	{
		cval[i] = 0ull;
		ckey[i] = MBG_BCACHE_EMPTY;
	}
	for (uint32_t i = (uint32_t)lane; i < EW; i += 64u)
		went[i] = 0ull;
	MBG_WAVE_SYNC();
	uint32_t s[W];
#pragma unroll
	for (int w = 0; w < W; ++w)
		s[w] = 0u;

	for (;;)
	{
		uint32_t c = 0;
		if (lane == 0)
			c = atomicAdd((unsigned int*)(a.counters + MBG_CTR_QUEUE), 1u);
		c = __builtin_amdgcn_readfirstlane(c);
		if (c >= a.n_chunks)
			break;
		const mbg_batch_chunk ch = ((const MBG_CONST mbg_batch_chunk*)a.chunks)[c];
		const uint32_t v = __builtin_amdgcn_readfirstlane(ch.variant);
		const MBG_CONST mbg_batch_variant& vd = ((const MBG_CONST mbg_batch_variant*)a.bv)[v];
		const float max_time = vd.max_time;
		const int windows = vd.windows;
		const unsigned long long win_off = vd.win_off, fin_off = vd.fin_off;
		const unsigned ent_off = vd.ent_off;
		for (int i = lane; i < a.n_nodes; i += 64)
			istate[i] = vd.istate[i];
		MBG_WAVE_SYNC();
		const mbg_param_view pv = vd.pv;
		mbg_eval_p_setup(&pv);

		// window cells through the wave's LDS cache when it has one: the
		// chunk's trajectories (one variant) add to a slot keyed by the cell,
		// and a cell whose slot another cell holds goes to HBM directly
		auto fw = [&](int w, uint64_t o, uint64_t q) {
			const uint32_t cell = ((uint32_t)w << a.n_obs) | (uint32_t)o;
			if (S)
			{
				const uint32_t slot = (cell * 0x9E3779B1u) >> (32 - sbits);
				const uint32_t k = atomicCAS(ckey + slot, MBG_BCACHE_EMPTY, cell);
				if (k == MBG_BCACHE_EMPTY || k == cell)
				{
					atomicAdd(cval + slot, (unsigned long long)q);
					return;
				}
			}
			atomicAdd(a.win_vals + win_off + cell, (unsigned long long)q);
		};
		auto fe = [&](int w, uint64_t q) {
			if (EW)
				atomicAdd(went + w, (unsigned long long)q);
			else
				atomicAdd(a.win_ent + ent_off + w, (unsigned long long)q);
		};
		auto emit = [&](bool has, int w, uint64_t o, uint64_t q) {
			if (has)
				fw(w, o, q);
		};

		uint64_t next = ch.begin;
		const uint64_t end = ch.end;
		float t = 0.0f;
		uint32_t k = 0;
		uint64_t id = 0, obs = 0; // obs: observed index of s, kept up to date at each flip
		bool active = false;
		pending pend;
		uint64_t ntrans = 0, nfinished = 0, ntrunc = 0;

		for (;;)
		{
			// new trajectories in whole initial-state rounds, as in the plain kernel
			const uint64_t need = __ballot(!active);
			if (need && next < end && ((uint64_t)__popcll(need) >= (uint64_t)a.refill_lanes || __ballot(active) == 0))
			{
				const uint64_t mine = next + __popcll(need & ((1ull << lane) - 1ull));
				bool fresh = false;
				if (!active && mine < end)
				{
					id = mine;
					t = 0.0f;
					k = 0;
					active = true;
					fresh = true;
				}
				init_states_wave<W>(s, fresh, id, v, a.seed_lo, a.seed_hi, a.n_nodes, istate);
				if (fresh)
					obs = mbg_obs_index(s);
				next = min(end, next + (uint64_t)__popcll(need));
			}
			if (__ballot(active) == 0)
				break;

			bool ended = false, fixed = false;
			mbg_eval_t e;
			float tp = 0.0f, H = 0.0f;
			if (active)
			{
				float u1, u2;
				step_uniforms(a.discrete, B0, k, id, v, a.seed_lo, a.seed_hi, u1, u2);
				mbg_eval_p(s, u2, &pv, &e, nullptr);
				tp = t;
				t = advance_time(a.discrete, t, e.total, u1, a.tick, max_time, fixed);
				H = fixed ? 0.0f : e.H;
				ntrans++;
			}
			record_wave(active, a.discrete, tp, t, obs, H, a.tick, a.qscale, windows, pend, emit, fe,
						[](int, uint64_t) {}, false, own);
			if (active)
			{
				k++;
				if (t >= max_time)
				{
					ended = true;
					nfinished++;
				}
				else if (k >= a.max_steps)
				{
					ended = true;
					ntrunc++;
				}
				else
				{
#pragma unroll
					for (int w = 0; w < W; ++w)
						s[w] ^= ((e.flip >> 5) == w) ? (1u << (e.flip & 31)) : 0u;
					obs ^= mbg_obs_flip(e.flip);
				}
			}

			if (__ballot(ended))
			{
				if (ended && t >= max_time)
					atomicAdd(a.fin_vals + fin_off + obs, 1ull);
				uint64_t fm = __ballot(ended && fixed);
				while (fm)
				{
					const int l = __ffsll((unsigned long long)fm) - 1;
					if (lane == l)
					{
						uint32_t key[W + 1];
#pragma unroll
						for (int w = 0; w < W; ++w)
							key[w] = s[w];
						key[W] = v;
						fp_insert<W + 1>(a.fp_table, a.fp_mask, key, 1u, a.counters);
					}
					fm &= fm - 1ull;
				}
				if (ended)
					active = false;
			}
		}
		pend.flush_all(fw, fe);
		if (S)
		{
			// the chunk is done: its cached cells to HBM, the slots emptied for the next chunk
			MBG_WAVE_SYNC();
			for (uint32_t i = (uint32_t)lane; i < S; i += 64u)
			{
				const uint32_t k = ckey[i];
				if (k != MBG_BCACHE_EMPTY)
				{
					atomicAdd(a.win_vals + win_off + k, cval[i]);
					ckey[i] = MBG_BCACHE_EMPTY;
					cval[i] = 0ull;
				}
			}
			MBG_WAVE_SYNC();
		}
		if (EW)
		{
			// the chunk's entropy sums to the variant's slice, the LDS sums emptied
			MBG_WAVE_SYNC();
			for (uint32_t i = (uint32_t)lane; i < (uint32_t)windows; i += 64u)
			{
				const unsigned long long v = went[i];
				if (v)
				{
					atomicAdd(a.win_ent + ent_off + i, v);
					went[i] = 0ull;
				}
			}
			MBG_WAVE_SYNC();
		}
		wave_add_counters(a.vctr + vd.ctr_off, a.vctr + vd.ctr_off + 1, a.vctr + vd.ctr_off + 2, ntrans, nfinished,
						  ntrunc);
	}
}

#endif

// Fills n 64-bit words with a value; used instead of memset nodes inside
// captured graphs.
extern "C" __global__ void __launch_bounds__(256) mbg_fill_u64(unsigned long long* p, unsigned long long n,
															   unsigned long long value)
{
	for (unsigned long long i = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
		 i += (unsigned long long)gridDim.x * blockDim.x)
		p[i] = value;
}
