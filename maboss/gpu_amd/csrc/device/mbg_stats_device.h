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

// Probe bound of the window-cell table: a cell that finds neither its key
// nor an empty slot within this many slots goes to the spill list.
#define MBG_WIN_PROBES 256

__device__ __forceinline__ void win_spill(const mbg_sim_args& a, uint64_t key, uint64_t q)
{
	const unsigned long long i = atomicAdd(a.counters + MBG_CTR_WIN_SPILL, 1ull);
	if (i < a.win_spill_cap)
	{
		a.win_spill_keys[i] = key;
		a.win_spill_vals[i] = q;
	}
	else
		atomicAdd(a.counters + MBG_CTR_WIN_OVERFLOW, 1ull);
}

// hash_add for the window-cell table, with the probe bound and the spill list
__device__ __forceinline__ void hash_add_win(const mbg_sim_args& a, uint64_t key, uint64_t q)
{
	uint64_t slot = mix64(key) & a.win_mask;
	for (uint64_t p = 0; p <= a.win_mask && p < MBG_WIN_PROBES; ++p)
	{
		unsigned long long* e = a.win_table + 2 * slot;
		const unsigned long long k = atomicCAS(e, MBG_EMPTY_KEY, (unsigned long long)key);
		if (k == MBG_EMPTY_KEY || k == key)
		{
			atomicAdd(e + 1, (unsigned long long)q);
			return;
		}
		slot = (slot + 1) & a.win_mask;
	}
	win_spill(a, key, q);
}

// One lane's window-cell insert into the hashed table as a sequence of
// non-blocking steps, so each memory round trip overlaps the lane's next
// steps: the slot's key is read first (a key already in the table, about half
// the inserts, then costs one read and the value add instead of a claim and
// the add; a read of a slot is also much cheaper than a failed claim), an
// empty slot is claimed with atomicCAS, a slot held by another key moves the
// probe on.  step() reads the result of the operation in flight and sends
// the next one, or the value add (fire and forget) when the slot is found.
// Same slots and sums as hash_add.
struct hash_inflight
{
	uint64_t slot = 0, key = 0, val = 0, ret = 0;
	uint32_t probes = 0;
	uint32_t state = 0; // 0 free, 1 key read in flight, 2 claim in flight

	__device__ __forceinline__ bool busy() const { return state != 0u; }

	__device__ __forceinline__ void issue(unsigned long long* table, uint64_t mask, uint64_t k, uint64_t q)
	{
		key = k;
		val = q;
		slot = mix64(k) & mask;
		probes = 0;
		ret = __hip_atomic_load(table + 2 * slot, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
		state = 1;
	}

	__device__ __forceinline__ void step(const mbg_sim_args& a)
	{
		unsigned long long* const table = a.win_table;
		const uint64_t mask = a.win_mask;
		if (state == 0u)
			return;
		if (ret == key || (state == 2u && ret == MBG_EMPTY_KEY))
		{
			atomicAdd(table + 2 * slot + 1, (unsigned long long)val);
			state = 0;
		}
		else if (state == 1u && ret == MBG_EMPTY_KEY)
		{
			// read as empty (or not yet seen filled): claim it
			ret = atomicCAS(table + 2 * slot, MBG_EMPTY_KEY, (unsigned long long)key);
			state = 2;
		}
		else if (++probes > mask || probes >= MBG_WIN_PROBES)
		{
			win_spill(a, key, val);
			state = 0;
		}
		else
		{
			slot = (slot + 1) & mask;
			ret = __hip_atomic_load(table + 2 * slot, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
			state = 1;
		}
	}

	__device__ __forceinline__ void finish(const mbg_sim_args& a)
	{
		while (state != 0u)
			step(a);
	}
};

// Append a (key, value) pair; the active lanes of the wave take consecutive
// slots with one atomic.  Pairs beyond the capacity are counted, not stored
// (the host then grows the buffers and runs again).
__device__ __forceinline__ void append_pair(unsigned long long* keys, unsigned long long* vals, uint64_t cap,
											unsigned long long* counter, uint64_t key, uint64_t q)
{
	const uint64_t m = __ballot(1);
	const int leader = __ffsll((unsigned long long)m) - 1;
	const int lane = __lane_id();
	uint64_t base = 0;
	if (lane == leader)
		base = atomicAdd(counter, (unsigned long long)__popcll(m));
	base = bcast64(base, leader);
	const uint64_t idx = base + __popcll(m & ((1ull << lane) - 1ull));
	if (idx < cap)
	{
		keys[idx] = key;
		vals[idx] = q;
	}
}

// Append through a per-wave chunk of reserved slots: cur = {next index,
// slots left} in LDS (one pair of words per wave).  The active lanes take
// consecutive slots; the leader reserves a new chunk (one global atomic per
// chunk) when the current one runs out, and lanes past the old chunk's end
// continue in the new one.
__device__ __forceinline__ void append_pair_chunked(unsigned long long* cur, unsigned long long* keys,
													unsigned long long* vals, uint64_t cap, unsigned long long* counter,
													uint64_t chunk, uint64_t key, uint64_t q)
{
	const uint64_t m = __ballot(1);
	const int leader = __ffsll((unsigned long long)m) - 1;
	const int lane = __lane_id();
	const uint64_t n = (uint64_t)__popcll(m);
	uint64_t start = 0, split = 0, nb = 0;
	if (lane == leader)
	{
		const uint64_t next = cur[0], left = cur[1];
		start = next;
		if (n <= left)
		{
			split = n;
			cur[0] = next + n;
			cur[1] = left - n;
		}
		else
		{
			split = left;
			nb = atomicAdd(counter, (unsigned long long)chunk);
			cur[0] = nb + (n - left);
			cur[1] = chunk - (n - left);
		}
	}
	start = bcast64(start, leader);
	split = bcast64(split, leader);
	nb = bcast64(nb, leader);
	const uint64_t rank = (uint64_t)__popcll(m & ((1ull << lane) - 1ull));
	const uint64_t idx = rank < split ? start + rank : nb + (rank - split);
	if (idx < cap)
	{
		keys[idx] = key;
		vals[idx] = q;
	}
}

// Window pairs of one wave staged in LDS with the fill count in a register.
// add() is called by the whole wave (lanes with has set contribute a pair).
// Once 64 pairs are staged they go to the sink, one per lane that is ready
// for one (sorted mode: every lane, HBM as whole lines; hashed mode: a lane
// whose previous insert has finished after one more step), and the rest move
// down; at most cap - 64 stay staged, so 64 new pairs always fit.
struct pair_stage_reg
{
	unsigned long long* keys; // [cap] then values [cap]
	int cap;
	unsigned int n = 0;

	// sink.step(): every lane advances its work in flight; sink.ready(): the
	// lane can take a pair; sink.take(key, value)
	template <typename SINK>
	__device__ void out(SINK&& sink)
	{
		const unsigned int lane = (unsigned int)__lane_id();
		sink.step();
		const bool ready = sink.ready();
		const uint64_t rm = __ballot(ready);
		const unsigned int nready = (unsigned int)__popcll(rm);
		const unsigned int take = n < nready ? n : nready;
		const unsigned int rank = (unsigned int)__popcll(rm & ((1ull << lane) - 1ull));
		if (ready && rank < take)
			sink.take(keys[rank], keys[cap + rank]);
		// the pairs [take, n) move to [0, n - take), 64 at a time from the bottom
		for (unsigned int base = 0; base + take < n; base += 64u)
		{
			unsigned long long k = 0, v = 0;
			const unsigned int src = take + base + lane;
			if (src < n)
			{
				k = keys[src];
				v = keys[cap + src];
			}
			__builtin_amdgcn_wave_barrier();
			if (src < n)
			{
				keys[base + lane] = k;
				keys[cap + base + lane] = v;
			}
			__builtin_amdgcn_wave_barrier();
		}
		n -= take;
	}

	template <typename SINK>
	__device__ void add(bool has, uint64_t key, uint64_t q, SINK&& sink)
	{
		const uint64_t m = __ballot(has);
		if (!m)
			return;
		if (n >= 64u)
			out(sink);
		while (n > (unsigned int)cap - 64u)
			out(sink);
		if (has)
		{
			const unsigned int i = n + (unsigned int)__popcll(m & ((1ull << __lane_id()) - 1ull));
			keys[i] = key;
			keys[cap + i] = q;
		}
		n += (unsigned int)__popcll(m);
	}

	template <typename SINK>
	__device__ void drain(SINK&& sink)
	{
		while (n > 0u)
			out(sink);
	}
};

// End of a wave's work: pad the rest of its chunk (all lanes of the wave).
__device__ __forceinline__ void pad_pair_chunk(unsigned long long* cur, unsigned long long* keys,
											   unsigned long long* vals, uint64_t cap)
{
	const uint64_t next = cur[0], left = cur[1];
	for (uint64_t i = __lane_id(); i < left; i += 64)
		if (next + i < cap)
		{
			keys[next + i] = MBG_PAIR_PAD_KEY; // sorts after every cell key
			vals[next + i] = 0ull;
		}
}

// Block-private count cache in LDS: slots {key + 1, count} (0 = empty),
// nslots a power of two, a few probes.  False when no slot was found (the
// caller then updates the global table directly).
__device__ __forceinline__ bool lds_count_add(unsigned long long* slots, int nslots, uint64_t key, uint64_t q)
{
	const unsigned long long k1 = (unsigned long long)key + 1ull;
	uint32_t s = (uint32_t)mix64(key) & (uint32_t)(nslots - 1);
	for (int p = 0; p < 8; ++p)
	{
		unsigned long long* e = slots + 2 * s;
		const unsigned long long old = atomicCAS(e, 0ull, k1);
		if (old == 0ull || old == k1)
		{
			atomicAdd(e + 1, (unsigned long long)q);
			return true;
		}
		s = (s + 1u) & (uint32_t)(nslots - 1);
	}
	return false;
}

// Fixed-point table insert of cnt trajectories.  Slots hold {tag, count, key
// words}; the key is the state (plus the variant in the batched form).
// Callers let one lane of a wave insert at a time, so the only wait (for a
// slot another wave is publishing) is on a wave that has claimed the slot
// and only has stores left; the wait is bounded.
template <int KW>
__device__ void fp_insert(unsigned int* table, uint64_t mask, const uint32_t (&key)[KW], unsigned int cnt,
						  unsigned long long* counters)
{
	constexpr int stride = 2 + KW;
	uint64_t h = 0x9e3779b97f4a7c15ull;
#pragma unroll
	for (int w = 0; w < KW; ++w)
		h = mix64(h ^ key[w]);
	uint64_t slot = h & mask;
	for (uint64_t p = 0; p <= mask; ++p)
	{
		unsigned int* e = table + slot * stride;
		unsigned int tag = __hip_atomic_load(e, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
		if (tag == 0u)
		{
			tag = atomicCAS(e, 0u, 1u);
			if (tag == 0u)
			{
#pragma unroll
				for (int w = 0; w < KW; ++w)
					e[2 + w] = key[w];
				__hip_atomic_store(e, 2u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
				atomicAdd(e + 1, cnt);
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
			atomicAdd(counters + MBG_CTR_FP_TIMEOUT, 1ull);
			return;
		}
		bool same = true;
#pragma unroll
		for (int w = 0; w < KW; ++w)
			same &= __hip_atomic_load(e + 2 + w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == key[w];
		if (same)
		{
			atomicAdd(e + 1, cnt);
			return;
		}
		slot = (slot + 1) & mask;
	}
	atomicAdd(counters + MBG_CTR_FP_OVERFLOW, 1ull);
}

// Block-private fixed-point cache in LDS, same slot layout as the global
// table; one lane of a wave at a time (as fp_insert).  Adds cnt trajectories;
// false when the probes find no slot.
template <int KW>
__device__ bool fp_lds_insert(unsigned int* t, int nslots, const uint32_t (&key)[KW], unsigned int cnt)
{
	constexpr int stride = 2 + KW;
	uint64_t h = 0x9e3779b97f4a7c15ull;
#pragma unroll
	for (int w = 0; w < KW; ++w)
		h = mix64(h ^ key[w]);
	uint32_t slot = (uint32_t)h & (uint32_t)(nslots - 1);
	for (int p = 0; p < 16; ++p)
	{
		unsigned int* e = t + slot * stride;
		unsigned int tag = atomicCAS(e, 0u, 1u);
		if (tag == 0u)
		{
#pragma unroll
			for (int w = 0; w < KW; ++w)
				e[2 + w] = key[w];
			__threadfence_block();
			atomicExch(e, 2u);
			atomicAdd(e + 1, cnt);
			return true;
		}
		for (int spin = 0; tag == 1u && spin < (1 << 16); ++spin)
			tag = __hip_atomic_load(e, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP);
		if (tag != 2u)
			return false;
		bool same = true;
#pragma unroll
		for (int w = 0; w < KW; ++w)
			same &= __hip_atomic_load(e + 2 + w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) == key[w];
		if (same)
		{
			atomicAdd(e + 1, cnt);
			return true;
		}
		slot = (slot + 1u) & (uint32_t)(nslots - 1);
	}
	return false;
}

// WM / FM: window and final-state storage (MBG_HIST_*) fixed at compile
// time, or -1 to follow the run's arguments.
template <int WM = -1, int FM = -1>
struct lane_hist_t
{
	unsigned long long* lds;
	const mbg_sim_args& a;

	__device__ __forceinline__ int win_mode() const { return WM >= 0 ? WM : a.win_mode; }
	__device__ __forceinline__ int fin_mode() const { return FM >= 0 ? FM : a.fin_mode; }

	__device__ void win(int w, uint64_t obs, uint64_t q) const
	{
		const uint64_t cell = ((uint64_t)w << a.n_obs) | obs;
		if (win_mode() == MBG_HIST_LDS)
			atomicAdd(lds + a.win_lds_off + cell, (unsigned long long)q);
		else if (win_mode() == MBG_HIST_DENSE)
			atomicAdd(a.win_vals + cell, (unsigned long long)q);
		else if (win_mode() == MBG_HIST_SORT)
			append_pair_chunked(lds + a.pair_lds_off + 4 * (threadIdx.x >> 6), a.pair_keys, a.pair_vals, a.pair_cap,
								a.counters + MBG_CTR_PAIRS, (uint64_t)a.pair_chunk, cell, q);
		else
			hash_add_win(a, cell, q);
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
		const uint64_t key = ((uint64_t)w0 << a.n_obs) | obs;
		if (a.tail_lds_slots && lds_count_add(lds + a.tail_lds_off, a.tail_lds_slots, key, 1ull))
			return;
		hash_add(a.tail_table, a.tail_mask, key, 1ull, a.counters + MBG_CTR_TAIL_OVERFLOW);
	}

	// cnt more trajectories ending in fixed point key (one lane of a wave at a time).
	template <int KW>
	__device__ void fixed(const uint32_t (&key)[KW], unsigned int cnt = 1u) const
	{
		if (a.fp_lds_slots && fp_lds_insert<KW>((unsigned int*)(lds + a.fp_lds_off), a.fp_lds_slots, key, cnt))
			return;
		fp_insert<KW>(a.fp_table, a.fp_mask, key, cnt, a.counters);
	}

	__device__ void fin(uint64_t obs) const
	{
		if (fin_mode() == MBG_HIST_LDS)
			atomicAdd(lds + a.fin_lds_off + obs, 1ull);
		else if (fin_mode() == MBG_HIST_DENSE)
			atomicAdd(a.fin_vals + obs, 1ull);
		else if (!a.fin_lds_cache_slots || !lds_count_add(lds + a.fin_lds_cache_off, a.fin_lds_cache_slots, obs, 1ull))
			hash_add(a.fin_table, a.fin_mask, obs, 1ull, a.counters + MBG_CTR_FIN_OVERFLOW);
	}
};

using lane_hist = lane_hist_t<>;

// End of a block: add its LDS histograms and caches to the global storage
// (after a barrier).
__device__ __forceinline__ void flush_block_hist(unsigned long long* lds, const mbg_sim_args& a)
{
	if (!a.lds_bytes)
		return;
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
	for (int i = threadIdx.x; i < a.tail_lds_slots; i += blockDim.x)
	{
		const unsigned long long k1 = lds[a.tail_lds_off + 2 * i];
		if (k1)
			hash_add(a.tail_table, a.tail_mask, k1 - 1ull, lds[a.tail_lds_off + 2 * i + 1],
					 a.counters + MBG_CTR_TAIL_OVERFLOW);
	}
	for (int i = threadIdx.x; i < a.fin_lds_cache_slots; i += blockDim.x)
	{
		const unsigned long long k1 = lds[a.fin_lds_cache_off + 2 * i];
		if (k1)
			hash_add(a.fin_table, a.fin_mask, k1 - 1ull, lds[a.fin_lds_cache_off + 2 * i + 1],
					 a.counters + MBG_CTR_FIN_OVERFLOW);
	}
}

// The block's cached fixed points to the global table: one slot per wave
// iteration, one lane inserting at a time.
template <int KW>
__device__ void flush_block_fp(unsigned long long* lds, const mbg_sim_args& a)
{
	if (!a.fp_lds_slots)
		return;
	const unsigned int* t = (const unsigned int*)(lds + a.fp_lds_off);
	const int lane = __lane_id(), wave = threadIdx.x >> 6, nwaves = blockDim.x >> 6;
	for (int base = wave * 64; base < a.fp_lds_slots; base += nwaves * 64)
	{
		const int i = base + lane;
		const unsigned int* e = t + (size_t)i * (2 + KW);
		const bool full = i < a.fp_lds_slots && e[0] == 2u && e[1] != 0u;
		uint64_t m = __ballot(full);
		while (m)
		{
			const int l = __ffsll((unsigned long long)m) - 1;
			if (lane == l)
			{
				uint32_t key[KW];
#pragma unroll
				for (int w = 0; w < KW; ++w)
					key[w] = e[2 + w];
				fp_insert<KW>(a.fp_table, a.fp_mask, key, e[1], a.counters);
			}
			m &= m - 1ull;
		}
	}
}

__device__ __forceinline__ uint32_t pick(const mbg_u32x4& r, uint32_t i)
{
	return i == 0 ? r.x : i == 1 ? r.y : i == 2 ? r.z : r.w;
}


} // namespace mbg_dev
