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
// mbg_table_compact -- occupied slots of a hash table as (key, value) pairs,
//                    and how many values are at least esc_min (the cells the
//                    packed copy sends through its side list).
// mbg_fp_compact   -- published slots of the fixed-point table as records.
// mbg_rate_view    -- parameter view of run-time rates held in device memory.
// mbg_cells_to_probs -- dense window cells to probabilities (double).
// mbg_pack_cells   -- key-sorted window cells as 12-byte records for the copy
//                    to the host.
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
				hash_add_win(a, cell, q * cnt);
		}
	}
}

// The occupied slots {key, value} of a hash table (window cells, final
// states) as (key, value) pairs, so that a sort hands the host the cells in
// key order.  A block takes 16 x 256 consecutive slots, thread t every 256th
// from its own (coalesced across the block), reads each slot once into
// registers, and the block's threads order their occupied slots with a wave
// scan and the waves' totals; one atomic per block reserves the output
// range.  Slots with a zero value are skipped.
constexpr int k_compact_per_thread = 16;
extern "C" __global__ void __launch_bounds__(256) mbg_table_compact(const unsigned long long* table,
																	unsigned long long cap, unsigned long long* keys,
																	unsigned long long* vals,
																	unsigned long long out_cap,
																	unsigned long long* count,
																	unsigned long long esc_min,
																	unsigned long long* esc_count)
{
	__shared__ unsigned int wtot[4], wesc[4];
	__shared__ unsigned long long base;
	const ulonglong2* slots = (const ulonglong2*)table;
	const uint64_t first = (uint64_t)blockIdx.x * blockDim.x * k_compact_per_thread + threadIdx.x;
	ulonglong2 e[k_compact_per_thread];
	unsigned int mine = 0, esc = 0;
#pragma unroll
	for (int j = 0; j < k_compact_per_thread; ++j)
	{
		const uint64_t i = first + (uint64_t)j * blockDim.x;
		e[j] = i < cap ? slots[i] : make_ulonglong2(MBG_EMPTY_KEY, 0ull);
		const bool used = e[j].x != MBG_EMPTY_KEY && e[j].y != 0ull;
		mine += used ? 1u : 0u;
		esc += (used && esc_min && e[j].y >= esc_min) ? 1u : 0u;
	}
	if (esc_min)
	{
		// values the packed copy sends through its side list (one atomic per block)
#pragma unroll
		for (int off = 32; off > 0; off >>= 1)
			esc += __shfl_xor(esc, off);
	}
	// inclusive scan over the wave, then the waves' totals in order
	const int lane = __lane_id(), wave = threadIdx.x >> 6;
	unsigned int incl = mine;
#pragma unroll
	for (int off = 1; off < 64; off <<= 1)
	{
		const unsigned int v = __shfl_up(incl, off);
		if (lane >= off)
			incl += v;
	}
	if (lane == 63)
		wtot[wave] = incl;
	if (lane == 0)
		wesc[wave] = esc;
	__syncthreads();
	if (threadIdx.x == 0)
	{
		unsigned int tot = 0;
		for (int w = 0; w < (int)(blockDim.x >> 6); ++w)
		{
			const unsigned int t = wtot[w];
			wtot[w] = tot;
			tot += t;
		}
		base = tot ? atomicAdd(count, (unsigned long long)tot) : 0ull;
		unsigned int esc_tot = 0;
		for (int w = 0; w < (int)(blockDim.x >> 6); ++w)
			esc_tot += wesc[w];
		if (esc_tot)
			atomicAdd(esc_count, (unsigned long long)esc_tot);
	}
	__syncthreads();
	uint64_t o = base + wtot[wave] + incl - mine;
#pragma unroll
	for (int j = 0; j < k_compact_per_thread; ++j)
		if (e[j].x != MBG_EMPTY_KEY && e[j].y != 0ull)
		{
			if (o < out_cap)
			{
				keys[o] = e[j].x;
				vals[o] = e[j].y;
			}
			o++;
		}
}

// The published slots of a fixed-point table (slots of `stride` words {tag,
// count, key words}; tag 2 = complete) as records of stride - 1 words {count,
// key words}, in no particular order (the host orders them).
extern "C" __global__ void __launch_bounds__(256) mbg_fp_compact(const unsigned int* table, unsigned long long cap,
																 int stride, unsigned int* out,
																 unsigned long long out_cap,
																 unsigned long long* count)
{
	for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < cap; i += (uint64_t)gridDim.x * blockDim.x)
	{
		const unsigned int* e = table + i * (uint64_t)stride;
		if (e[0] != 2u || e[1] == 0u)
			continue;
		const unsigned long long o = atomicAdd(count, 1ull);
		if (o >= out_cap)
			continue;
		unsigned int* r = out + o * (uint64_t)(stride - 1);
		for (int w = 1; w < stride; ++w)
			r[w - 1] = e[w];
	}
}

// Parameter view of run-time rates held in device memory: the view header
// (pointers into the same block), rate * log2(rate) of each rate (the same
// float code as the host's table), and enable masks with every bit set.
extern "C" __global__ void __launch_bounds__(256) mbg_rate_view(mbg_param_view* view, const float* rate, float* rate_lr,
															   uint32_t* en, int n_params, int n_words)
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i == 0)
	{
		mbg_param_view v;
		v.rate = (const MBG_RATE_AS float*)rate;
		v.rate_lr = (const MBG_RATE_AS float*)rate_lr;
		v.up_en = (const MBG_RATE_AS uint32_t*)en;
		v.dn_en = (const MBG_RATE_AS uint32_t*)(en + n_words);
		*view = v;
	}
	if (i < n_params)
		rate_lr[i] = mbg_rlog2r(rate[i]);
	if (i < 2 * n_words)
		en[i] = 0xffffffffu;
}

// Window probabilities from dense fixed-point cells: raw / denom in double,
// the division the host result applies (denom = n * 2^32 in continuous
// time, n in discrete time).
extern "C" __global__ void __launch_bounds__(256) mbg_cells_to_probs(const unsigned long long* cells,
																	unsigned long long n, double denom, double* out)
{
	for (unsigned long long i = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
		 i += (unsigned long long)gridDim.x * blockDim.x)
		out[i] = (double)cells[i] / denom;
}

// Key-sorted window cells (keys below 2^kbits, 1 <= kbits <= 60) as 12-byte
// records for the copy to the host: record i is the 96-bit integer
// key | field << kbits, stored as three 32-bit words, least significant
// first.  With esc_min = 0 (kbits <= 31, so the 96 - kbits bit field holds
// any 64-bit value) the field is the value.  Otherwise the field has
// vb = 96 - kbits bits and esc_min = 2^(vb - 1): a value below esc_min is the
// field itself, and a larger one goes to the side list xval at a position p
// the field records as esc_min + p.  Positions are handed out per block (one
// atomic reserves the block's range), so the list's order varies from run to
// run while every record decodes to the same value; *xcount ends at the
// number of escaped values, which the caller knows beforehand (the table
// compaction counts them) and sizes xval with.  A thread packs 4 consecutive
// cells with 16-byte loads and stores.
extern "C" __global__ void __launch_bounds__(256) mbg_pack_cells(const unsigned long long* keys,
																 const unsigned long long* vals, unsigned long long n,
																 unsigned int kbits, unsigned long long esc_min,
																 unsigned int* rec, unsigned long long* xval,
																 unsigned long long xcap, unsigned long long* xcount)
{
	__shared__ unsigned int wtot[4];
	__shared__ unsigned long long base;
	const uint64_t c0 = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) * 4;
	unsigned long long k[4], v[4];
	if (c0 + 3 < n)
	{
		const ulonglong2 ka = *(const ulonglong2*)(keys + c0), kb = *(const ulonglong2*)(keys + c0 + 2);
		const ulonglong2 va = *(const ulonglong2*)(vals + c0), vb = *(const ulonglong2*)(vals + c0 + 2);
		k[0] = ka.x, k[1] = ka.y, k[2] = kb.x, k[3] = kb.y;
		v[0] = va.x, v[1] = va.y, v[2] = vb.x, v[3] = vb.y;
	}
	else
	{
#pragma unroll
		for (int j = 0; j < 4; ++j)
		{
			k[j] = c0 + j < n ? keys[c0 + j] : 0ull;
			v[j] = c0 + j < n ? vals[c0 + j] : 0ull;
		}
	}
	bool x[4];
	unsigned int mine = 0;
#pragma unroll
	for (int j = 0; j < 4; ++j)
	{
		x[j] = c0 + j < n && esc_min && v[j] >= esc_min;
		mine += x[j] ? 1u : 0u;
	}
	// the block's escapes: wave scan of the per-thread counts, waves' totals in order
	const int lane = __lane_id(), wave = threadIdx.x >> 6;
	unsigned int incl = mine;
#pragma unroll
	for (int off = 1; off < 64; off <<= 1)
	{
		const unsigned int t = __shfl_up(incl, off);
		if (lane >= off)
			incl += t;
	}
	if (lane == 63)
		wtot[wave] = incl;
	__syncthreads();
	if (threadIdx.x == 0)
	{
		unsigned int tot = 0;
		for (int q = 0; q < (int)(blockDim.x >> 6); ++q)
		{
			const unsigned int t = wtot[q];
			wtot[q] = tot;
			tot += t;
		}
		base = tot ? atomicAdd(xcount, (unsigned long long)tot) : 0ull;
	}
	__syncthreads();
	uint64_t o = base + wtot[wave] + incl - mine;
	unsigned int w[12];
#pragma unroll
	for (int j = 0; j < 4; ++j)
	{
		unsigned long long f = v[j];
		if (x[j])
		{
			if (o < xcap)
				xval[o] = v[j];
			f = esc_min + o;
			o++;
		}
		const unsigned long long lo = k[j] | (f << kbits);
		w[3 * j] = (unsigned int)lo;
		w[3 * j + 1] = (unsigned int)(lo >> 32);
		w[3 * j + 2] = (unsigned int)(f >> (64 - kbits));
	}
	if (c0 + 3 < n)
	{
		uint4* out = (uint4*)(rec + c0 * 3);
		out[0] = make_uint4(w[0], w[1], w[2], w[3]);
		out[1] = make_uint4(w[4], w[5], w[6], w[7]);
		out[2] = make_uint4(w[8], w[9], w[10], w[11]);
	}
	else
	{
		for (int j = 0; j < 4; ++j)
			if (c0 + j < n)
			{
				rec[(c0 + j) * 3] = w[3 * j];
				rec[(c0 + j) * 3 + 1] = w[3 * j + 1];
				rec[(c0 + j) * 3 + 2] = w[3 * j + 2];
			}
	}
}
