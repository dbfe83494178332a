// Fused lane-per-trajectory simulate kernel for gfx950.
//
// One lane owns one trajectory at a time, with its W state words, time and
// step counter in VGPRs.  The kernel is persistent: lanes pull trajectory ids
// from a wave-aggregated atomic work queue until the queue is drained, so
// uneven trajectory lengths do not leave lanes idle.  Each SSA step
// (src/jit_kernels/simulation.cu:106-148) feeds its dwell slice straight into
// the window histogram (src/jit_kernels/window_average_small.cu:50-136),
// without writing a trajectory record to HBM; the end of a trajectory feeds
// the final-state histogram (src/jit_kernels/final_states.cu:9-21) and the
// fixed-point table (src/statistics/fixed_states.cu:84-125).
//
// All histogram cells are unsigned 64-bit fixed point (dwell * 2^32 / tick,
// rounded to nearest), so the result does not depend on the order in which
// lanes add: it is the same bits on every run, on any launch shape, and in
// the host oracle.  Consecutive contributions of one lane to the same
// (window, observed state) cell are summed in registers before one atomic.
//
// Built once per state-word count MBG_KW into LLVM bitcode; hipRTC links it
// with the generated mbg_eval / mbg_eval_p / mbg_obs_index of a model.
#include <hip/hip_runtime.h>

#include "mbg_kernel_args.h"
#include "mbg_stats_device.h"

#ifndef MBG_KW
#error "MBG_KW (state words) must be defined"
#endif

#ifndef MBG_PARAM
#define MBG_PARAM 0
#endif

using namespace mbg_dev;

namespace
{

constexpr int W = MBG_KW;

// Fixed-point table insert.  Called by one lane of a wave at a time, so the
// only wait (for a slot another wave is publishing) is on a wave that has
// already claimed the slot and only has stores left; the wait is bounded.
__device__ void fp_insert(const mbg_sim_args& a, const uint32_t* s)
{
	constexpr int stride = 2 + W;
	uint64_t h = 0x9e3779b97f4a7c15ull;
#pragma unroll
	for (int w = 0; w < W; ++w)
		h = mix64(h ^ s[w]);
	uint64_t slot = h & a.fp_mask;
	for (uint64_t p = 0; p <= a.fp_mask; ++p)
	{
		unsigned int* e = a.fp_table + slot * stride;
		unsigned int tag = __hip_atomic_load(e, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
		if (tag == 0u)
		{
			tag = atomicCAS(e, 0u, 1u);
			if (tag == 0u)
			{
#pragma unroll
				for (int w = 0; w < W; ++w)
					e[2 + w] = s[w];
				__hip_atomic_store(e, 2u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
				atomicAdd(e + 1, 1u);
				return;
			}
		}
		for (int spin = 0; tag == 1u && spin < (1 << 22); ++spin)
		{
			__builtin_amdgcn_s_sleep(2);
			tag = __hip_atomic_load(e, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
		}
		if (tag != 2u)
		{
			atomicAdd(a.counters + MBG_CTR_FP_TIMEOUT, 1ull);
			return;
		}
		bool same = true;
#pragma unroll
		for (int w = 0; w < W; ++w)
			same &= __hip_atomic_load(e + 2 + w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == s[w];
		if (same)
		{
			atomicAdd(e + 1, 1u);
			return;
		}
		slot = (slot + 1) & a.fp_mask;
	}
	atomicAdd(a.counters + MBG_CTR_FP_OVERFLOW, 1ull);
}

} // namespace

extern "C" __global__ void __launch_bounds__(256) simulate(mbg_sim_args a)
{
	extern __shared__ unsigned long long lds[];
	for (int i = threadIdx.x; i < a.lds_bytes / 8; i += blockDim.x)
		lds[i] = 0ull;
	__syncthreads();

	const lane_hist hist { lds, a };
	const int lane = __lane_id();
	const uint32_t B0 = ((uint32_t)a.n_nodes + 3u) >> 2;

	uint32_t s[W];
#pragma unroll
	for (int w = 0; w < W; ++w)
		s[w] = 0u;
	float t = 0.0f;
	uint32_t k = 0;
	uint64_t id = 0;
	bool active = false, exhausted = false;

	// register-combined window / entropy contributions
	int pw = -1, pew = -1;
	uint64_t pobs = 0, pq = 0, peq = 0;
	uint64_t ntrans = 0, nfinished = 0, ntrunc = 0;

#if MBG_PARAM
	const mbg_param_view* pv = a.pv;
#endif

	for (;;)
	{
		// Refill lanes whose trajectory ended: one atomic per wave.
		const uint64_t need = __ballot(!active && !exhausted);
		if (need)
		{
			const int leader = __ffsll((unsigned long long)need) - 1;
			uint64_t base = 0;
			if (lane == leader)
				base = atomicAdd(a.counters + MBG_CTR_QUEUE, (unsigned long long)__popcll(need));
			base = bcast64(base, leader);
			if (!active && !exhausted)
			{
				const uint64_t mine = base + __popcll(need & ((1ull << lane) - 1ull));
				if (mine < a.n_traj)
				{
					id = a.id_base + mine;
					// Initial state (src/jit_kernels/simulation.cu:34-71):
					// bit i = (u_i <= istate_i), u_i = draw i of the stream.
#pragma unroll
					for (int w = 0; w < W; ++w)
					{
						uint32_t word = 0u;
						for (int b = 0; b < 32; b += 4)
						{
							const int i = w * 32 + b;
							if (i >= a.n_nodes)
								break;
							const mbg_u32x4 r = mbg_philox4x32_10(
								mbg_u32x4 { (uint32_t)(i >> 2), (uint32_t)id, (uint32_t)(id >> 32), a.variant },
								a.seed_lo, a.seed_hi);
							const uint32_t rw[4] = { r.x, r.y, r.z, r.w };
#pragma unroll
							for (int j = 0; j < 4; ++j)
								if (i + j < a.n_nodes && mbg_u01(rw[j]) <= a.istate[i + j])
									word |= 1u << (b + j);
						}
						s[w] = word;
					}
					t = 0.0f;
					k = 0;
					active = true;
				}
				else
					exhausted = true;
			}
		}
		if (__ballot(active) == 0)
			break;

		bool ended = false, fixed = false;
		if (active)
		{
			// Uniforms of step k (see mbg_numeric.h for the stream layout).
			float u1 = 1.0f, u2;
			if (a.discrete)
			{
				const mbg_u32x4 r = mbg_philox4x32_10(
					mbg_u32x4 { B0 + (k >> 2), (uint32_t)id, (uint32_t)(id >> 32), a.variant }, a.seed_lo, a.seed_hi);
				u2 = mbg_u01(pick(r, k & 3u));
			}
			else
			{
				const mbg_u32x4 r = mbg_philox4x32_10(
					mbg_u32x4 { B0 + (k >> 1), (uint32_t)id, (uint32_t)(id >> 32), a.variant }, a.seed_lo, a.seed_hi);
				u1 = mbg_u01((k & 1u) ? r.z : r.x);
				u2 = mbg_u01((k & 1u) ? r.w : r.y);
			}

			mbg_eval_t e;
#if MBG_PARAM
			mbg_eval_p(s, u2, pv, &e);
#else
			mbg_eval(s, u2, &e);
#endif
			const float tp = t;
			float H = 0.0f;
			if (e.total == 0.0f)
			{
				fixed = true;
				t = a.max_time;
			}
			else
			{
				if (a.discrete)
				{
					t = t + a.tick;
					t = (a.max_time - a.tick < t) ? a.max_time : t;
				}
				else
					t = t + (-mbg_logf(u1)) / e.total;
				t = fminf(t, a.max_time);
				H = e.H;
			}
			ntrans++;

			// The record (state during [tp, t), H) goes into the windows.
			const uint64_t obs = mbg_obs_index(s);
			if (a.discrete)
			{
				const int w = (int)roundf(tp / a.tick);
				if (w < a.windows)
				{
					if (w == pw && obs == pobs)
						pq += 1ull;
					else
					{
						if (pw >= 0)
							hist.win(pw, pobs, pq);
						pw = w;
						pobs = obs;
						pq = 1ull;
					}
					const uint64_t qe = (uint64_t)rintf(H * a.qscale);
					if (qe)
					{
						if (w == pew)
							peq += qe;
						else
						{
							if (pew >= 0)
								hist.ent(pew, peq);
							pew = w;
							peq = qe;
						}
					}
				}
			}
			else
			{
				int w = (int)floorf(tp / a.tick);
				float b = tp;
				while (t > b)
				{
					const float wend = (float)(w + 1) * a.tick;
					const float en = fminf(t, wend);
					const float sl = en - b;
					if (w < a.windows)
					{
						const uint64_t q = (uint64_t)rintf(sl * a.qscale);
						if (q)
						{
							if (w == pw && obs == pobs)
								pq += q;
							else
							{
								if (pw >= 0)
									hist.win(pw, pobs, pq);
								pw = w;
								pobs = obs;
								pq = q;
							}
						}
						const uint64_t qe = (uint64_t)rintf((H * sl) * a.qscale);
						if (qe)
						{
							if (w == pew)
								peq += qe;
							else
							{
								if (pew >= 0)
									hist.ent(pew, peq);
								pew = w;
								peq = qe;
							}
						}
					}
					w++;
					b = en;
				}
			}

			k++;
			if (t >= a.max_time)
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
			}
		}

		// Trajectory end: final state and fixed points (uniform region).
		const uint64_t endm = __ballot(ended);
		if (endm)
		{
			if (ended && t >= a.max_time)
				hist.fin(mbg_obs_index(s));
			uint64_t fm = __ballot(ended && fixed);
			while (fm)
			{
				const int l = __ffsll((unsigned long long)fm) - 1;
				if (lane == l)
					fp_insert(a, s);
				fm &= fm - 1ull;
			}
			if (ended)
				active = false;
		}
	}

	if (pw >= 0)
		hist.win(pw, pobs, pq);
	if (pew >= 0)
		hist.ent(pew, peq);

	// Wave totals of the counters, one atomic per wave.
	for (int off = 32; off > 0; off >>= 1)
	{
		ntrans += __shfl_xor(ntrans, off);
		nfinished += __shfl_xor(nfinished, off);
		ntrunc += __shfl_xor(ntrunc, off);
	}
	if (lane == 0)
	{
		atomicAdd(a.counters + MBG_CTR_TRANSITIONS, (unsigned long long)ntrans);
		atomicAdd(a.counters + MBG_CTR_FINISHED, (unsigned long long)nfinished);
		if (ntrunc)
			atomicAdd(a.counters + MBG_CTR_TRUNCATED, (unsigned long long)ntrunc);
	}

	// Flush the block's LDS histograms.
	__syncthreads();
	if (a.lds_bytes)
	{
		const uint64_t nwin = a.win_mode == MBG_HIST_LDS ? ((uint64_t)a.windows << a.n_obs) : 0;
		for (uint64_t i = threadIdx.x; i < nwin; i += blockDim.x)
		{
			const unsigned long long v = lds[a.win_lds_off + i];
			if (v)
				atomicAdd(a.win_vals + i, v);
		}
		if (a.ent_lds)
			for (int i = threadIdx.x; i < a.windows; i += blockDim.x)
			{
				const unsigned long long v = lds[a.ent_lds_off + i];
				if (v)
					atomicAdd(a.win_ent + i, v);
			}
		const uint64_t nfin = a.fin_mode == MBG_HIST_LDS ? (1ull << a.n_obs) : 0;
		for (uint64_t i = threadIdx.x; i < nfin; i += blockDim.x)
		{
			const unsigned long long v = lds[a.fin_lds_off + i];
			if (v)
				atomicAdd(a.fin_vals + i, v);
		}
	}
}
