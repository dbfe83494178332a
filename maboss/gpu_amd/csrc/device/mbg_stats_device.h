// Device helpers shared by the simulate kernels: hashing, wave broadcasts
// and the window / entropy / final-state histogram updates.
#pragma once

#if !defined(__HIPCC_RTC__)
#include <hip/hip_runtime.h>
#endif

#include "mbg_kernel_args.h"

namespace mbg_dev
{

__device__ __forceinline__ uint64_t mix64(uint64_t x)
{
	// splitmix64 finaliser (Steele, Lea, Flood, "Fast splittable pseudorandom
	// number generators", OOPSLA 2014)
	x ^= x >> 30;
	x *= 0xbf58476d1ce4e5b9ull;
	x ^= x >> 27;
	x *= 0x94d049bb133111ebull;
	x ^= x >> 31;
	return x;
}

__device__ __forceinline__ uint64_t bcast64(uint64_t v, int lane)
{
	const uint32_t lo = __builtin_amdgcn_readlane((uint32_t)v, lane);
	const uint32_t hi = __builtin_amdgcn_readlane((uint32_t)(v >> 32), lane);
	return ((uint64_t)hi << 32) | lo;
}

// Open-addressing insert-or-add into a table of 16-byte slots {key, value}
// (key and value share a cache line).  Linear probing, bounded by the
// capacity; a full table is reported through the overflow counter.
__device__ __forceinline__ void hash_add(unsigned long long* table, uint64_t mask, uint64_t key, uint64_t q,
										 unsigned long long* overflow)
{
	uint64_t slot = mix64(key) & mask;
	for (uint64_t p = 0; p <= mask; ++p)
	{
		unsigned long long* e = table + 2 * slot;
		const unsigned long long k = atomicCAS(e, MBG_EMPTY_KEY, (unsigned long long)key);
		if (k == MBG_EMPTY_KEY || k == key)
		{
			atomicAdd(e + 1, (unsigned long long)q);
			return;
		}
		slot = (slot + 1) & mask;
	}
	atomicAdd(overflow, 1ull);
}

struct lane_hist
{
	unsigned long long* lds;
	const mbg_sim_args& a;

	__device__ void win(int w, uint64_t obs, uint64_t q) const
	{
		const uint64_t cell = ((uint64_t)w << a.n_obs) | obs;
		if (a.win_mode == MBG_HIST_LDS)
			atomicAdd(lds + a.win_lds_off + cell, (unsigned long long)q);
		else if (a.win_mode == MBG_HIST_DENSE)
			atomicAdd(a.win_vals + cell, (unsigned long long)q);
		else
			hash_add(a.win_table, a.win_mask, cell, q, a.counters + MBG_CTR_WIN_OVERFLOW);
	}

	__device__ void ent(int w, uint64_t q) const
	{
		if (a.ent_lds)
			atomicAdd(lds + a.ent_lds_off + w, (unsigned long long)q);
		else
			atomicAdd(a.win_ent + w, (unsigned long long)q);
	}

	// One more trajectory whose final record covers windows w0.. in state obs.
	__device__ void tail(int w0, uint64_t obs) const
	{
		hash_add(a.tail_table, a.tail_mask, ((uint64_t)w0 << a.n_obs) | obs, 1ull, a.counters + MBG_CTR_TAIL_OVERFLOW);
	}

	__device__ void fin(uint64_t obs) const
	{
		if (a.fin_mode == MBG_HIST_LDS)
			atomicAdd(lds + a.fin_lds_off + obs, 1ull);
		else if (a.fin_mode == MBG_HIST_DENSE)
			atomicAdd(a.fin_vals + obs, 1ull);
		else
			hash_add(a.fin_table, a.fin_mask, obs, 1ull, a.counters + MBG_CTR_FIN_OVERFLOW);
	}
};

__device__ __forceinline__ uint32_t pick(const mbg_u32x4& r, uint32_t i)
{
	return i == 0 ? r.x : i == 1 ? r.y : i == 2 ? r.z : r.w;
}


} // namespace mbg_dev
