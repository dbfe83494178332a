// Utility kernels of the runner (precompiled code object, model independent).
//
// mbg_fill_slots  -- empty hash slots {key = ~0, value = 0}, 16 bytes each,
//                    written with 16-byte stores (a 2D memset of the key
//                    halves is several times slower).
// mbg_fill_u64    -- fill n 64-bit words (used instead of memset nodes in
//                    captured graphs).
// mbg_expand_tails -- expands tail records into window cells.  A trajectory's
//                    last record spans [tp, max_time); every window after the
//                    first one it touches receives the same fixed-point
//                    quantum q(w) = rint((min(max_time, (w+1)*tick) - w*tick)
//                    * 2^32/tick) from every such trajectory, so the kernels
//                    count those trajectories per (first full window, observed
//                    state) and this kernel adds count * q(w) to each window w,
//                    which is exactly the sum of the per-window additions.
#include <hip/hip_runtime.h>

#include "mbg_kernel_args.h"
#include "mbg_stats_device.h"

using namespace mbg_dev;

extern "C" __global__ void __launch_bounds__(256) mbg_fill_slots(ulonglong2* p, unsigned long long n)
{
	for (unsigned long long i = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
		 i += (unsigned long long)gridDim.x * blockDim.x)
		p[i] = make_ulonglong2(MBG_EMPTY_KEY, 0ull);
}

extern "C" __global__ void __launch_bounds__(256) mbg_fill_u64(unsigned long long* p, unsigned long long n,
															   unsigned long long value)
{
	for (unsigned long long i = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
		 i += (unsigned long long)gridDim.x * blockDim.x)
		p[i] = value;
}

// q(w) of a window the final record covers from its start
__device__ __forceinline__ uint64_t full_window_quantum(int w, float tick, float max_time, float qscale)
{
	const float b = (float)w * tick;
	const float en = fminf(max_time, (float)(w + 1) * tick);
	// the record's window loop stops once a window starts at or after max_time
	if (en <= b)
		return 0;
	return (uint64_t)rintf((en - b) * qscale);
}

extern "C" __global__ void __launch_bounds__(256) mbg_expand_tails(mbg_sim_args a)
{
	const uint64_t cap = a.tail_mask + 1;
	const uint64_t omask = a.n_obs >= 64 ? ~0ull : ((1ull << a.n_obs) - 1ull);
	for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < cap; i += (uint64_t)gridDim.x * blockDim.x)
	{
		const unsigned long long key = a.tail_table[2 * i];
		const unsigned long long cnt = a.tail_table[2 * i + 1];
		if (key == MBG_EMPTY_KEY || cnt == 0)
			continue;
		const int w0 = (int)(key >> a.n_obs);
		const uint64_t obs = key & omask;
		for (int w = w0; w < a.windows; ++w)
		{
			const uint64_t q = full_window_quantum(w, a.tick, a.max_time, a.qscale);
			if (!q)
				continue;
			const uint64_t cell = ((uint64_t)w << a.n_obs) | obs;
			if (a.win_mode == MBG_HIST_DENSE)
				atomicAdd(a.win_vals + cell, (unsigned long long)(q * cnt));
			else if (a.win_mode == MBG_HIST_SORT)
				append_pair(a.pair_keys, a.pair_vals, a.pair_cap, a.counters + MBG_CTR_PAIRS, cell, q * cnt);
			else
				hash_add(a.win_table, a.win_mask, cell, q * cnt, a.counters + MBG_CTR_WIN_OVERFLOW);
		}
	}
}

// The occupied cells of a hashed window table as (cell key, value) pairs in
// the sorted mode's pair buffers (one wave-aggregated append per round), so
// the pair sort hands the host the cells in key order.
extern "C" __global__ void __launch_bounds__(256) mbg_table_pairs(mbg_sim_args a)
{
	const uint64_t cap = a.win_mask + 1;
	const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
	// every lane runs the same number of rounds, so each append sees its whole wave
	const uint64_t rounds = (cap + stride - 1) / stride;
	for (uint64_t r = 0; r < rounds; ++r)
	{
		const uint64_t i = r * stride + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
		const unsigned long long key = i < cap ? a.win_table[2 * i] : MBG_EMPTY_KEY;
		const unsigned long long val = i < cap ? a.win_table[2 * i + 1] : 0ull;
		if (key != MBG_EMPTY_KEY && val != 0ull)
			append_pair(a.pair_keys, a.pair_vals, a.pair_cap, a.counters + MBG_CTR_PAIRS, key, val);
	}
}
