// Likelihood-ratio gradient kernel for gfx950 (no counterpart in the
// reference).  Compiled by hipRTC together with a model's generated code,
// which provides MBG_W, MBG_GRAD_NP, mbg_eval_p, mbg_obs_index and
// mbg_grad_dwell, so the per-parameter accumulators are statically indexed
// registers.
//
// Estimator.  P_w(x) = E[(1/tick) * time spent in window w with observed
// state x].  For an upstream gradient G[w][x] = dL/dP_w(x),
//   dL/dtheta_j = 1/(n tick) * sum_traj sum_w Gw * S_j(e_w)
// where Gw = sum of G[w][obs] * slice over the trajectory's dwell slices in
// window w, e_w is the end of window w (or of the trajectory) and S_j is the
// score of the path up to e_w:
//   S_j(t) = sum_{jumps before t} 1/theta_j [jump uses theta_j]
//            - integral_0^t c_j(s(u)) du,   c_j = d total / d theta_j.
// Gw of later windows is not known when a jump or dwell happens, so the
// kernel uses the suffix form of the same sum: with C(u) = sum of Gw over the
// windows completed by time u and Gtot = C at the trajectory's end,
//   sum_w Gw S_j(e_w) = Gtot * (n_j / theta_j - T_j)
//                       - sum_{jumps of j} C(t_k) / theta_j + BT_j,
//   T_j = integral c_j du,  BT_j = integral c_j C du.
// Per lane and parameter T_j lives in a register in double precision (the
// per-trajectory score accumulator), fed with each record's dwell summed in
// double over its window slices; n_j lives in LDS (u16, lane private).  BT_j
// is accumulated by events: with D(t) = integral_0^t C du (a double per
// lane), BT_j = sum over the intervals where node direction k, whose rate is
// theta_j, is enabled of D(end) - D(start), so when an enabled indicator
// turns on (off) at t, -D(t) (+D(t)) goes straight to a per-block double
// accumulator in LDS, and the intervals still open at the trajectory's end
// close there.  Ended trajectories are folded together: each lane's
// Gtot * (n_j / theta_j - T_j) + (open intervals) * D is summed over the
// wave (DPP) and added to the block accumulator, which also takes the
// -C(t_k) / theta_j term of every jump; the block totals go to the output
// with one atomic per parameter.  Only T occupies registers per parameter
// (a lane keeping BT_j, or a running total of its trajectories, as well
// needs twice as many).

// The trajectory is the forward kernel's: same Philox stream (seed, id,
// variant), same rates (runtime values through mbg_eval_p), same time and
// flip arithmetic, so the sample is the one the forward statistics came from.
#pragma once

#include "mbg_kernel_args.h"
#include "mbg_lane_core.h"

#ifndef MBG_GRAD_NP
#error "MBG_GRAD_NP must be defined by the generated code"
#endif

// Waves per SIMD the registers are budgeted for (set by the host from the
// parameter count; the per-parameter accumulators take most registers).
#ifndef MBG_GRAD_WAVES
#define MBG_GRAD_WAVES 1
#endif

// Trajectories of a wave that have ended wait, without stepping, until
// a.fold_lanes of the wave's lanes have ended (or no lane is still running),
// and are then folded together: the fold touches every parameter and sums
// over the wave, so doing it for one lane at a time would run it on most
// steps with one or two lanes active.

namespace mbg_grad
{

// Sum of v over the wave's 64 lanes, in lane 63 (every lane must call): the
// inclusive DPP scan of mbg_lane_core.h on the two halves of a double.
template <int CTRL, int ROWS>
__device__ __forceinline__ double add_dpp_f64(double v)
{
	const uint64_t b = __double_as_longlong(v);
	const uint32_t lo = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)b, CTRL, ROWS, 0xF, false);
	const uint32_t hi = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)(b >> 32), CTRL, ROWS, 0xF, false);
	return v + __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}

__device__ __forceinline__ double wave_sum_f64(double v)
{
	v = add_dpp_f64<0x111, 0xF>(v);
	v = add_dpp_f64<0x112, 0xF>(v);
	v = add_dpp_f64<0x114, 0xF>(v);
	v = add_dpp_f64<0x118, 0xF>(v);
	v = add_dpp_f64<0x142, 0xA>(v);
	v = add_dpp_f64<0x143, 0xC>(v);
	return v;
}

} // namespace mbg_grad

extern "C" __global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(MBG_GRAD_WAVES)))
simulate_grad(mbg_grad_args a)
{
	constexpr int NP = MBG_GRAD_NP;
	constexpr int W = MBG_W;
	// u16 jump counts per lane, lane-major: an odd number of 32-bit words per
	// lane, so the lanes' counts of one parameter fall in distinct banks
	constexpr int CW = ((NP + 1) / 2) | 1;
	extern __shared__ double glds[];
	// LDS: [NP] block accumulators, [NP] 1/theta, [cells] G (optional),
	// [2N] jump parameters (u16), [blockDim][2 CW] jump counts (u16)
	double* acc = glds;
	double* invt = glds + NP;
	const uint64_t cells = (uint64_t)a.windows << a.n_obs;
	double* gtab = glds + 2 * NP;
	unsigned short* jp = (unsigned short*)(gtab + (a.g_lds ? cells : 0));
	unsigned short* cnt = jp + ((2 * a.n_nodes + 1) & ~1);
	for (int i = threadIdx.x; i < NP; i += blockDim.x)
	{
		acc[i] = 0.0;
		const float r = a.pv.rate[i];
		invt[i] = r > 0.0f ? 1.0 / (double)r : 0.0;
	}
	if (a.g_lds)
		for (uint64_t i = threadIdx.x; i < cells; i += blockDim.x)
			gtab[i] = a.gprob[i];
	for (int i = threadIdx.x; i < 2 * a.n_nodes; i += blockDim.x)
		jp[i] = a.jparam[i];
	for (int i = threadIdx.x; i < 2 * CW * (int)blockDim.x; i += blockDim.x)
		cnt[i] = 0;
	mbg_eval_p_setup(&a.pv);
	__syncthreads();
	const double* G = a.g_lds ? gtab : a.gprob;
	unsigned short* const mycnt = cnt + 2 * CW * threadIdx.x; // count of parameter j at mycnt[j]

	const int lane = __lane_id();
	const uint32_t B0 = ((uint32_t)a.n_nodes + 3u) >> 2;
	uint32_t s[W];
	// enabled up (cu) / down (cd) directions of the current record's state;
	// all clear before a trajectory's first record
	uint32_t cu[W], cd[W];
#pragma unroll
	for (int w = 0; w < W; ++w)
	{
		s[w] = 0u;
		cu[w] = 0u;
		cd[w] = 0u;
	}
	// T[j]: the current trajectory's integral of c_j
	double T[NP];
#pragma unroll
	for (int j = 0; j < NP; ++j)
		T[j] = 0.0;
	float t = 0.0f;
	uint32_t k = 0;
	uint64_t id = 0;
	bool active = false, waiting = false, exhausted = false;
	// completed windows' G; G of the window in progress; D = integral of C
	double C = 0.0, gw = 0.0, D = 0.0;
	uint64_t ntrans = 0;

	mbg_dev::id_chunks queue;
	for (;;)
	{
		// fold the ended trajectories: Gtot * (n_j / theta_j - T_j) and the
		// intervals still open (D at the end) join the block accumulators
		const uint64_t wm = __ballot(waiting);
		if (wm && (__popcll(wm) >= a.fold_lanes || __ballot(active) == 0ull))
		{
			const double gt = C + gw;
#pragma unroll
			for (int j = 0; j < NP; ++j)
			{
				double v = 0.0;
				if (waiting)
				{
					v = fma(gt, fma((double)mycnt[j], invt[j], -T[j]), (double)mbg_grad_on(j, cu, cd) * D);
					mycnt[j] = 0;
					T[j] = 0.0;
				}
				v = mbg_grad::wave_sum_f64(v);
				if (lane == 63 && v != 0.0)
					atomicAdd(acc + j, v);
			}
			if (waiting)
			{
#pragma unroll
				for (int w = 0; w < W; ++w)
				{
					cu[w] = 0u;
					cd[w] = 0u;
				}
				waiting = false;
			}
		}

		const uint64_t need = __ballot(!active && !waiting && !exhausted);
		if (need)
		{
			const uint64_t mine = queue.take(need, a.counters + MBG_CTR_QUEUE, a.n_traj, 64);
			bool fresh = false;
			if (!active && !waiting && !exhausted)
			{
				if (mine < a.n_traj)
				{
					id = a.id_base + mine;
					t = 0.0f;
					k = 0;
					C = 0.0;
					gw = 0.0;
					D = 0.0;
					active = true;
					fresh = true;
				}
				else
					exhausted = true;
			}
			mbg_dev::init_states_wave<W>(s, fresh, id, a.variant, a.seed_lo, a.seed_hi, a.n_nodes, a.istate);
		}
		if (__ballot(active || waiting) == 0)
			break;

		if (active)
		{
			float u1, u2;
			mbg_dev::step_uniforms(false, B0, k, id, a.variant, a.seed_lo, a.seed_hi, u1, u2);
			mbg_eval_t e;
			uint32_t en[W]; // enabled bits: the rate of each node without re-evaluating its logic
			mbg_eval_p(s, u2, &a.pv, &e, en);
			const float tp = t;
			bool fixed = false;
			t = mbg_dev::advance_time(false, t, e.total, u1, a.tick, a.max_time, fixed);
			ntrans++;

			// enabled directions that turned on (-D) or off (+D) at tp; none
			// while D is 0 (a trajectory's first records)
#pragma unroll
			for (int w = 0; w < W; ++w)
			{
				const uint32_t nu = en[w] & ~s[w], nd = en[w] & s[w];
				uint64_t tog = D != 0.0 ? ((uint64_t)(nd ^ cd[w]) << 32 | (nu ^ cu[w])) : 0ull;
				while (tog)
				{
					const int bit = __ffsll((unsigned long long)tog) - 1;
					tog &= tog - 1ull;
					const int dir = bit >> 5, node = 32 * w + (bit & 31);
					const unsigned p = jp[2 * node + dir];
					if (p != 0xffffu)
						atomicAdd(acc + p, (((dir ? nd : nu) >> (bit & 31)) & 1u) ? -D : D);
				}
				cu[w] = nu;
				cd[w] = nd;
			}

			// dwell slices of [tp, t): G weights, window completions, C-weighted dwell
			const uint64_t obs = mbg_obs_index(s);
			int w = (int)floorf(tp / a.tick);
			float b = tp;
			double dsum = 0.0, cdt = 0.0;
			while (t > b)
			{
				const float wend = (float)(w + 1) * a.tick;
				const float en = fminf(t, wend);
				const float sl = en - b;
				cdt += C * (double)sl;
				if (w < a.windows)
					gw += G[((uint64_t)w << a.n_obs) | obs] * (double)sl;
				dsum += (double)sl;
				if (en == wend)
				{
					C += gw;
					gw = 0.0;
				}
				++w;
				b = en;
			}
			mbg_grad_dwell(s, en, dsum, T);
			D += cdt;

			k++;
			if (t >= a.max_time || k >= a.max_steps)
			{
				active = false;
				waiting = true;
			}
			else
			{
				const int f = e.flip;
				const uint32_t xf = (s[f >> 5] >> (f & 31)) & 1u;
				const unsigned p = jp[2 * f + (int)xf];
				if (p != 0xffffu)
				{
					mycnt[p] += 1;
					atomicAdd(acc + p, -invt[p] * C);
				}
#pragma unroll
				for (int w2 = 0; w2 < W; ++w2)
					s[w2] ^= ((f >> 5) == w2) ? (1u << (f & 31)) : 0u;
				// keep the 16-bit counts exact on very long trajectories
				if ((k & 0x7fffu) == 0u)
				{
#pragma unroll
					for (int j = 0; j < NP; ++j)
					{
						T[j] -= (double)mycnt[j] * invt[j];
						mycnt[j] = 0;
					}
				}
			}
		}
	}

	for (int off = 32; off > 0; off >>= 1)
		ntrans += __shfl_xor(ntrans, off);
	if (lane == 0 && ntrans)
		atomicAdd(a.counters + MBG_CTR_TRANSITIONS, (unsigned long long)ntrans);
	__syncthreads();
	for (int j = threadIdx.x; j < NP; j += blockDim.x)
		if (acc[j] != 0.0)
			atomicAdd(a.grad + j, acc[j] * a.scale);
}
