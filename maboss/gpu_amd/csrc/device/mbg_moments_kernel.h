// Per-trajectory second moments of the window statistics, for the ErrorTH
// and ErrorProba columns (the reference prints both as 0,
// src/statistics/window_average_small.cpp:235, :244-245).  Compiled by hipRTC
// with a model's generated code, which provides MBG_W, mbg_eval_p and
// mbg_obs_index.
//
// One lane per trajectory replays the run's trajectories: the same Philox
// stream (seed, id, variant), the same rates (run-time values through
// mbg_eval_p, which gives the plain kernel's floats), the same time, flip and
// window-slice arithmetic, so its fixed-point slices are the ones the window
// cells were summed from.  Per lane it keeps the window the trajectory is in:
// up to MBG_MOM_K observed states with the trajectory's fixed-point time in
// each (all visits to a state in the window together) and its entropy sum T.
// When the trajectory leaves the window, or ends, each time P adds P^2 to
// sq[(w << n_obs) | obs] and T adds T^2 to esq[w], exactly, as 128-bit sums
// (low word with the returned old value giving the carry, then the high
// word).  A window that sees more than MBG_MOM_K states appends every one of
// its partial times, as (id, w, obs, P) records, to the spill list; the host
// adds those up per (id, w, obs) and squares them.
#pragma once

#include "mbg_kernel_args.h"
#include "mbg_lane_core.h"

#ifndef MBG_MOM_K
#define MBG_MOM_K 8
#endif

namespace mbg_mom
{

__device__ __forceinline__ void add_u128(unsigned long long* p, uint64_t lo, uint64_t hi)
{
	const uint64_t old = atomicAdd(p, (unsigned long long)lo);
	hi += (old + lo < old) ? 1u : 0u;
	if (hi)
		atomicAdd(p + 1, (unsigned long long)hi);
}

__device__ __forceinline__ void add_square(unsigned long long* p, uint64_t v)
{
	add_u128(p, v * v, __umul64hi(v, v));
}

struct window_acc
{
	int w = -1, n = 0;
	bool spilled = false;
	uint64_t T = 0;
	uint64_t obs[MBG_MOM_K], q[MBG_MOM_K];
};

__device__ __forceinline__ void spill(const mbg_moments_args& a, uint64_t id, int w, uint64_t obs, uint64_t q)
{
	const unsigned long long i = atomicAdd(a.counters + MBG_MOM_CTR_SPILL, 1ull);
	if (i < a.spill_cap)
	{
		unsigned long long* r = a.spill + 4 * i;
		r[0] = id;
		r[1] = (unsigned long long)w;
		r[2] = obs;
		r[3] = q;
	}
}

__device__ __forceinline__ void flush(const mbg_moments_args& a, window_acc& c, uint64_t id)
{
	if (c.w >= 0)
	{
#pragma unroll
		for (int j = 0; j < MBG_MOM_K; ++j)
			if (j < c.n)
			{
				if (c.spilled)
					spill(a, id, c.w, c.obs[j], c.q[j]);
				else
					add_square(a.sq + 2 * ((((uint64_t)c.w) << a.n_obs) | c.obs[j]), c.q[j]);
			}
		if (c.T)
			add_square(a.esq + 2 * (uint64_t)c.w, c.T);
	}
	c.w = -1;
	c.n = 0;
	c.T = 0;
	c.spilled = false;
}

__device__ __forceinline__ void at(const mbg_moments_args& a, window_acc& c, int w, uint64_t id)
{
	if (w != c.w)
	{
		flush(a, c, id);
		c.w = w;
	}
}

__device__ __forceinline__ void add(const mbg_moments_args& a, window_acc& c, uint64_t obs, uint64_t q, uint64_t id)
{
	bool done = false;
#pragma unroll
	for (int j = 0; j < MBG_MOM_K; ++j)
		if (!done && j < c.n && c.obs[j] == obs)
		{
			c.q[j] += q;
			done = true;
		}
	if (done)
		return;
	if (c.n < MBG_MOM_K)
	{
#pragma unroll
		for (int j = 0; j < MBG_MOM_K; ++j)
			if (j == c.n)
			{
				c.obs[j] = obs;
				c.q[j] = q;
			}
		c.n++;
	}
	else
	{
		spill(a, id, c.w, obs, q);
		c.spilled = true;
	}
}

} // namespace mbg_mom

extern "C" __global__ void __launch_bounds__(256) simulate_moments(mbg_moments_args a)
{
	constexpr int W = MBG_W;
	mbg_eval_p_setup(&a.pv);
	__syncthreads();

	const uint32_t B0 = ((uint32_t)a.n_nodes + 3u) >> 2;
	uint32_t s[W];
#pragma unroll
	for (int w = 0; w < W; ++w)
		s[w] = 0u;
	float t = 0.0f;
	uint32_t k = 0;
	uint64_t id = 0;
	bool active = false, exhausted = false;
	mbg_mom::window_acc c;

	mbg_dev::id_chunks queue;
	for (;;)
	{
		// new trajectories in whole initial-state rounds, as in the simulate kernel
		const uint64_t need = __ballot(!active && !exhausted);
		if (need && ((uint64_t)__popcll(need) >= (uint64_t)a.refill_lanes || __ballot(active) == 0))
		{
			const uint64_t mine = queue.take(need, a.counters + MBG_MOM_CTR_QUEUE, a.n_traj, 64);
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
			mbg_dev::init_states_wave<W>(s, fresh, id, a.variant, a.seed_lo, a.seed_hi, a.n_nodes, a.istate);
		}
		if (__ballot(active) == 0)
			break;
		if (active)
		{
			float u1, u2;
			mbg_dev::step_uniforms(a.discrete, B0, k, id, a.variant, a.seed_lo, a.seed_hi, u1, u2);
			mbg_eval_t e;
			mbg_eval_p(s, u2, &a.pv, &e, nullptr);
			const float tp = t;
			bool fixed = false;
			t = mbg_dev::advance_time(a.discrete, t, e.total, u1, a.tick, a.max_time, fixed);
			const float H = fixed ? 0.0f : e.H;
			const uint64_t obs = mbg_obs_index(s);
			if (a.discrete)
			{
				const int w = (int)roundf(tp / a.tick);
				if (w < a.windows)
				{
					mbg_mom::at(a, c, w, id);
					mbg_mom::add(a, c, obs, 1ull, id);
					c.T += (uint64_t)rintf(H * a.qscale);
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
						mbg_mom::at(a, c, w, id);
						if (q)
							mbg_mom::add(a, c, obs, q, id);
						c.T += (uint64_t)rintf((H * sl) * a.qscale);
					}
					++w;
					b = en;
				}
			}
			k++;
			if (t >= a.max_time || k >= a.max_steps)
			{
				mbg_mom::flush(a, c, id);
				active = false;
			}
			else
			{
				const int f = e.flip;
#pragma unroll
				for (int w2 = 0; w2 < W; ++w2)
					s[w2] ^= ((f >> 5) == w2) ? (1u << (f & 31)) : 0u;
			}
		}
	}
}
