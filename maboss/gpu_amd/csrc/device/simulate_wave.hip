// Wave-per-trajectory simulate kernel for large models (gfx950).
//
// One 64-lane wave owns one trajectory.  The N nodes are split into 64
// contiguous slices of npl = ceil(N/64) nodes, one slice per lane, and the
// state is distributed one 32-bit word per lane (word w lives in lane w, so
// N <= 2048).  A node's logic is a truth table over at most 6 parents whose
// bits are fetched from the owning lanes with ds_bpermute (__shfl); its rate
// is (state == logic) ? 0 : (state ? down : up), the reference's fast form
// (src/generator.cpp:93-107).
//
// Rates live in LDS (one [rows][64] array per wave, rows = npl rounded up to
// a power of two, the extra rows zero).  After a flip only the flipped node
// and its children are re-evaluated, lanes taking one of them each.  Every
// step each lane re-sums its slice sequentially, a DPP scan over the 64 lanes
// (row_scan) gives the total and the prefix of every slice, a ballot picks
// the slice holding u2*total, and the slice's own running sums pick the node
// (src/jit_kernels/simulation.cu:9-21 selects by prefix; here the canonical
// order is slice-then-scan, and zero-rate nodes are never chosen).  Time, entropy and the window, final
// and fixed-point statistics follow src/jit_kernels/simulation.cu:106-163 and
// src/jit_kernels/window_average_small.cu:50-136, as in the lane kernel.
//
// Random numbers: the same Philox stream as the lane kernel; each lane
// computes one block of a 64-block batch, so one Philox evaluation per lane
// serves 128 continuous (256 discrete) steps.
#include <hip/hip_runtime.h>

#include "mbg_kernel_args.h"
#include "mbg_stats_device.h"
#include "mbg_lane_core.h"

using namespace mbg_dev;

namespace
{

__device__ __forceinline__ float rdlane_f(float v, int l)
{
	return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}

__device__ __forceinline__ uint32_t rdlane_u(uint32_t v, int l) { return __builtin_amdgcn_readlane(v, l); }

// Rate of node i in the wave's state (sw = this lane's state word).  Called by
// all lanes together: every lane takes part in the parent-word shuffles.
__device__ __forceinline__ float node_rate(const mbg_wnode* nodes, int i, bool valid, uint32_t sw)
{
	const mbg_wnode& nd = nodes[valid ? i : 0];
	const int npar = valid ? nd.npar : 0;
	uint32_t idx = 0;
#pragma unroll
	for (int j = 0; j < MBG_WAVE_MAXPAR; ++j)
	{
		const unsigned p = nd.par[j];
		const uint32_t w = (uint32_t)__shfl((int)sw, (int)(p >> 5));
		idx |= (j < npar ? ((w >> (p & 31u)) & 1u) : 0u) << j;
	}
	const uint32_t ws = (uint32_t)__shfl((int)sw, i >> 5);
	const bool on = (ws >> (i & 31)) & 1u;
	const bool logic = (nd.tt >> idx) & 1ull;
	return (!valid || on == logic) ? 0.0f : (on ? nd.down : nd.up);
}

// v + (DPP-moved value of another lane, or 0 where the move has no source or
// the row is masked off): one v_add_f32_dpp
template <int CTRL, int ROWS>
__device__ __forceinline__ float add_dpp(float v)
{
	return v + __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, ROWS, 0xF, false));
}

// Inclusive scan over the 64 lanes in the DPP order: Kogge-Stone inside each
// row of 16 lanes (row_shr 1, 2, 4, 8), then rows 1 and 3 add the last lane
// of the row before (row_bcast:15), then rows 2 and 3 add lane 31
// (row_bcast:31).  Lanes without a term add 0.
__device__ __forceinline__ float row_scan(float v)
{
	v = add_dpp<0x111, 0xF>(v);
	v = add_dpp<0x112, 0xF>(v);
	v = add_dpp<0x114, 0xF>(v);
	v = add_dpp<0x118, 0xF>(v);
	v = add_dpp<0x142, 0xA>(v);
	v = add_dpp<0x143, 0xC>(v);
	return v;
}

// Running sums of row 0: lane k (< 16) ends with ((0 + r_0) + r_1) + ... + r_k,
// the sequential sum (each step moves the row right by one lane and adds r).
__device__ __forceinline__ float row0_running_sum(float r)
{
	float q = 0.0f + r;
#pragma unroll
	for (int j = 1; j < 16; ++j)
		q = r + __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(q), 0x111, 0xF, 0xF, false));
	return q;
}

__device__ __forceinline__ uint64_t fp_hash(uint32_t sw, int W)
{
	uint64_t h = 0x9e3779b97f4a7c15ull;
	for (int w = 0; w < W; ++w)
		h = mix64(h ^ rdlane_u(sw, w));
	return h;
}

// Add cnt trajectories ending in the fixed point whose word w lane w holds
// (sw; lanes >= W hold none) to the global fixed-point table: the slots,
// hash and probing of fp_insert, done by the whole wave (lane 0 claims,
// every lane writes or compares its word).
__device__ void wave_fp_global(const mbg_sim_args& a, uint32_t sw, int W, int lane, uint32_t cnt)
{
	const int stride = 2 + W;
	uint64_t slot = fp_hash(sw, W) & a.fp_mask;
	for (uint64_t p = 0; p <= a.fp_mask; ++p)
	{
		unsigned int* e = a.fp_table + slot * stride;
		uint32_t tag = 0, claimed = 0;
		if (lane == 0)
		{
			tag = __hip_atomic_load(e, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
			if (tag == 0u)
			{
				tag = atomicCAS(e, 0u, 1u);
				claimed = tag == 0u;
			}
		}
		claimed = rdlane_u(claimed, 0);
		if (claimed)
		{
			if (lane < W)
				e[2 + lane] = sw;
			__threadfence();
			if (lane == 0)
			{
				__hip_atomic_store(e, 2u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
				atomicAdd(e + 1, cnt);
			}
			return;
		}
		if (lane == 0)
			for (int spin = 0; tag == 1u && spin < (1 << 22); ++spin)
			{
				__builtin_amdgcn_s_sleep(2);
				tag = __hip_atomic_load(e, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
			}
		tag = rdlane_u(tag, 0);
		if (tag != 2u)
		{
			if (lane == 0)
				atomicAdd(a.counters + MBG_CTR_FP_TIMEOUT, 1ull);
			return;
		}
		__threadfence();
		const bool same = lane >= W || __hip_atomic_load(e + 2 + lane, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == sw;
		if (__ballot(!same) == 0ull)
		{
			if (lane == 0)
				atomicAdd(e + 1, cnt);
			return;
		}
		slot = (slot + 1) & a.fp_mask;
	}
	if (lane == 0)
		atomicAdd(a.counters + MBG_CTR_FP_OVERFLOW, 1ull);
}

// One more trajectory ending in that fixed point, counted in the block's LDS
// cache: nslots (a power of two) slots of (2 + W) words {tag, count, key},
// tag 0 free, 1 being written, 2 complete.  False when 16 probes find no
// slot (the caller then goes to the global table).
__device__ bool wave_fp_lds(unsigned int* t, int nslots, uint32_t sw, int W, int lane)
{
	const int stride = 2 + W;
	uint32_t slot = (uint32_t)fp_hash(sw, W) & (uint32_t)(nslots - 1);
	for (int p = 0; p < 16; ++p)
	{
		unsigned int* e = t + slot * stride;
		uint32_t tag = 0;
		if (lane == 0)
			tag = atomicCAS(e, 0u, 1u);
		tag = rdlane_u(tag, 0);
		if (tag == 0u)
		{
			if (lane < W)
				e[2 + lane] = sw;
			__threadfence_block();
			if (lane == 0)
			{
				atomicExch(e, 2u);
				atomicAdd(e + 1, 1u);
			}
			return true;
		}
		if (lane == 0)
			for (int spin = 0; tag == 1u && spin < (1 << 16); ++spin)
				tag = __hip_atomic_load(e, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP);
		tag = rdlane_u(tag, 0);
		if (tag != 2u)
			return false;
		const bool same =
			lane >= W || __hip_atomic_load(e + 2 + lane, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) == sw;
		if (__ballot(!same) == 0ull)
		{
			if (lane == 0)
				atomicAdd(e + 1, 1u);
			return true;
		}
		slot = (slot + 1u) & (uint32_t)(nslots - 1);
	}
	return false;
}

template <int NPLMAX>
__device__ void wave_main(const mbg_sim_args& a)
{
	extern __shared__ unsigned long long lds[];
	char* ldsb = (char*)lds;
	const int N = a.n_nodes;
	const int npl = a.wt.npl;
	const int W = (N + 31) >> 5;

	// stage the node tables, clear the LDS histograms
	{
		const uint4* src = (const uint4*)a.wt.nodes;
		uint4* dst = (uint4*)(ldsb + a.wave_lds_off);
		for (int i = threadIdx.x; i < a.wt.tab_bytes / 16; i += blockDim.x)
			dst[i] = src[i];
		for (int i = threadIdx.x; i < a.lds_bytes / 8; i += blockDim.x)
			lds[i] = 0ull;
	}
	__syncthreads();

	const mbg_wnode* nodes = (const mbg_wnode*)(ldsb + a.wave_lds_off);
	const unsigned* coff = (const unsigned*)(nodes + N);
	const unsigned short* child = (const unsigned short*)(coff + N + 1);
	const unsigned short* obsl = child + a.wt.n_child;
	const unsigned short* ra = obsl + a.n_obs; // rate slot of each node
	const unsigned short* oba = ra + N;        // rate slot of each observed node
	const float4* otab = (const float4*)(ldsb + a.wave_lds_off + a.wt.obs_tab_off);
	const int lane = __lane_id();
	const int wave = threadIdx.x >> 6;
	// the wave's rates, NPLMAX rows of 64 lanes; rows npl.. stay zero, so the
	// per-slot loops need no bounds
	float* rates = (float*)(ldsb + a.wave_lds_off + a.wt.tab_bytes) + (size_t)wave * NPLMAX * 64;
	for (int s = npl; s < NPLMAX; ++s)
		rates[s * 64 + lane] = 0.0f;

	const lane_hist hist { lds, a };
	const uint32_t B0 = ((uint32_t)N + 3u) >> 2;

	pending pend;
	auto fw = [&](int w, uint64_t o, uint64_t q) {
		if (lane == 0)
			hist.win(w, o, q);
	};
	auto fe = [&](int w, uint64_t q) {
		if (lane == 0)
			hist.ent(w, q);
	};
	auto ft = [&](int w, uint64_t o) {
		if (lane == 0)
			hist.tail(w, o);
	};
	const bool use_tails = a.tail_table != nullptr;
	uint64_t ntrans = 0, nfinished = 0, ntrunc = 0;

	for (;;)
	{
		uint64_t tid = 0;
		if (lane == 0)
			tid = atomicAdd(a.counters + MBG_CTR_QUEUE, 1ull);
		tid = bcast64(tid, 0);
		if (tid >= a.n_traj)
			break;
		const uint64_t id = a.id_base + tid;
		const uint32_t idlo = (uint32_t)id, idhi = (uint32_t)(id >> 32);

		// initial state, word `lane` (src/jit_kernels/simulation.cu:34-71)
		uint32_t sw = 0u;
		if (lane < W)
		{
			for (int b = 0; b < 8; ++b)
			{
				const int i0 = lane * 32 + b * 4;
				if (i0 >= N)
					break;
				const mbg_u32x4 r =
					mbg_philox4x32_10(mbg_u32x4 { (uint32_t)(i0 >> 2), idlo, idhi, a.variant }, a.seed_lo, a.seed_hi);
				const uint32_t rw[4] = { r.x, r.y, r.z, r.w };
#pragma unroll
				for (int j = 0; j < 4; ++j)
					if (i0 + j < N && mbg_u01(rw[j]) <= a.istate[i0 + j])
						sw |= 1u << (b * 4 + j);
			}
		}
		// observed index: bit k = observed node k
		uint64_t obs = 0;
		{
			const int o = lane < a.n_obs ? obsl[lane] : 0;
			const uint32_t w = (uint32_t)__shfl((int)sw, o >> 5);
			obs = __ballot(lane < a.n_obs && ((w >> (o & 31)) & 1u));
		}
		// all rates
#pragma unroll
		for (int k = 0; k < NPLMAX; ++k)
		{
			if (k < npl)
			{
				const int i = lane * npl + k;
				const float rv = node_rate(nodes, i, i < N, sw);
				rates[k * 64 + lane] = rv;
			}
		}

		float t = 0.0f;
		uint32_t k = 0;
		mbg_u32x4 rb { 0, 0, 0, 0 };
		bool fixed = false, ended_ok = false;
		for (;;)
		{
			// this lane's slice and the scan over lanes
			float L = 0.0f;
#pragma unroll
			for (int s = 0; s < NPLMAX; ++s)
				L = L + rates[s * 64 + lane];
			const float P = row_scan(L);
			const float total = rdlane_f(P, 63);

			// uniforms of step k: one Philox block per lane per batch
			float u1 = 1.0f, u2;
			if (a.discrete)
			{
				if ((k & 255u) == 0u)
					rb = mbg_philox4x32_10(mbg_u32x4 { B0 + ((k >> 8) << 6) + (uint32_t)lane, idlo, idhi, a.variant },
										   a.seed_lo, a.seed_hi);
				u2 = mbg_u01(rdlane_u(pick(rb, k & 3u), (int)((k >> 2) & 63u)));
			}
			else
			{
				if ((k & 127u) == 0u)
					rb = mbg_philox4x32_10(mbg_u32x4 { B0 + ((k >> 7) << 6) + (uint32_t)lane, idlo, idhi, a.variant },
										   a.seed_lo, a.seed_hi);
				const int src = (int)((k >> 1) & 63u);
				u1 = mbg_u01(rdlane_u((k & 1u) ? rb.z : rb.x, src));
				u2 = mbg_u01(rdlane_u((k & 1u) ? rb.w : rb.y, src));
			}

			const float tp = t;
			float H = 0.0f;
			if (total == 0.0f)
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
					t = t + (-mbg_logf(u1)) / total;
				t = fminf(t, a.max_time);
				// transition entropy, log-sum form: observed node j (< 64) on lane j
				float Sv = 0.0f, Tv = 0.0f, v = 0.0f;
				if (lane < a.n_obs)
				{
					v = rates[oba[lane]];
					const float4 ot = otab[lane]; // {up, down, r log2 r of each}
					Sv = Sv + v;
					Tv = Tv + (v == 0.0f ? 0.0f : (v == ot.x ? ot.z : ot.w));
				}
				const float S = rdlane_f(row_scan(Sv), 63);
				const float T = rdlane_f(row_scan(Tv), 63);
				const int cnt = __popcll(__ballot(v != 0.0f));
				const float h = mbg_log2f(S) - T / S;
				H = (cnt > 1 && h > 0.0f) ? h : 0.0f;
			}
			ntrans++;

			// record (state during [tp, t), H) into the windows; the values
			// are wave-uniform and lane 0 performs the flushes
			record(a.discrete, tp, t, obs, H, a.tick, a.qscale, a.windows, pend, fw, fe, ft,
				   use_tails && t >= a.max_time);

			k++;
			if (t >= a.max_time)
			{
				nfinished++;
				ended_ok = true;
				break;
			}
			if (k >= a.max_steps)
			{
				ntrunc++;
				break;
			}

			// flip selection: first lane whose scan reaches u2*total with
			// weight in its slice (else the last lane with weight), then the
			// first slot of that slice reaching the remainder with a nonzero
			// rate (else its last nonzero slot)
			const float rr = u2 * total;
			const unsigned long long c1 = __ballot(P >= rr && L > 0.0f);
			const unsigned long long c2 = __ballot(L > 0.0f);
			// c2 is nonzero whenever total > 0; the fallback only guards NaN rates
			const int ls = c1 ? __ffsll(c1) - 1 : (c2 ? 63 - __clzll(c2) : 0);
			const float rem = rr - (ls == 0 ? 0.0f : rdlane_f(P, ls - 1));
			int kstar;
			if (NPLMAX <= 16)
			{
				// the selected slice on lanes 0..15 of row 0, running sums by DPP
				const float rs = lane < NPLMAX ? rates[lane * 64 + ls] : 0.0f;
				const float q = row0_running_sum(rs);
				const bool pos = lane < 16 && rs > 0.0f;
				const uint64_t hit = __ballot(pos && q >= rem), any = __ballot(pos);
				kstar = hit ? __ffsll((unsigned long long)hit) - 1 : (any ? 63 - __clzll(any) : -1);
			}
			else
			{
				float qsum = 0.0f;
				int ks = -1, last = -1;
#pragma unroll
				for (int s = 0; s < NPLMAX; ++s)
				{
					const float rs = rates[s * 64 + lane];
					qsum = qsum + rs;
					const bool pos = rs > 0.0f;
					last = pos ? s : last;
					ks = (pos && ks < 0 && qsum >= rem) ? s : ks;
				}
				kstar = (int)rdlane_u((uint32_t)(ks >= 0 ? ks : last), ls);
			}
			const int f = min(max(ls * npl + kstar, 0), N - 1);

			if (lane == (f >> 5))
				sw ^= 1u << (f & 31);
			const unsigned op = nodes[f].obs_pos;
			if (op != MBG_WAVE_NOOBS)
				obs ^= 1ull << op;

			// re-evaluate the flipped node and its children
			const int c0 = (int)coff[f];
			const int items = 1 + (int)coff[f + 1] - c0;
			for (int base = 0; base < items; base += 64)
			{
				const int j = base + lane;
				const bool valid = j < items;
				const int node = !valid ? 0 : (j == 0 ? f : (int)child[c0 + j - 1]);
				const float rv = node_rate(nodes, node, valid, sw);
				if (valid)
					rates[ra[node]] = rv;
			}
			__builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
		}

		// trajectory end: final state and fixed points
		if (ended_ok)
		{
			if (lane == 0)
				hist.fin(obs);
			if (fixed && !(a.fp_lds_slots && wave_fp_lds((unsigned int*)(lds + a.fp_lds_off), a.fp_lds_slots, sw, W, lane)))
				wave_fp_global(a, sw, W, lane, 1u);
		}
	}

	pend.flush_all(fw, fe);
	if (lane == 0)
	{
		atomicAdd(a.counters + MBG_CTR_TRANSITIONS, (unsigned long long)ntrans);
		atomicAdd(a.counters + MBG_CTR_FINISHED, (unsigned long long)nfinished);
		if (ntrunc)
			atomicAdd(a.counters + MBG_CTR_TRUNCATED, (unsigned long long)ntrunc);
	}

	if (a.win_mode == MBG_HIST_SORT)
		pad_pair_chunk(lds + a.pair_lds_off + 4 * wave, a.pair_keys, a.pair_vals, a.pair_cap);
	__syncthreads();
	flush_block_hist(lds, a);
	// the block's cached fixed points: one slot per wave
	if (a.fp_lds_slots)
	{
		const unsigned int* t = (const unsigned int*)(lds + a.fp_lds_off);
		for (int i = wave; i < a.fp_lds_slots; i += (int)(blockDim.x >> 6))
		{
			const unsigned int* e = t + (size_t)i * (2 + W);
			const uint32_t tag = rdlane_u(e[0], 0), cnt = rdlane_u(e[1], 0);
			if (tag == 2u && cnt != 0u)
				wave_fp_global(a, lane < W ? e[2 + lane] : 0u, W, lane, cnt);
		}
	}
}

} // namespace

#define MBG_WAVE_KERNEL(n)                                                                                             \
	extern "C" __global__ void __launch_bounds__(512) simulate_wave_n##n(mbg_sim_args a) { wave_main<n>(a); }

MBG_WAVE_KERNEL(1)
MBG_WAVE_KERNEL(2)
MBG_WAVE_KERNEL(4)
MBG_WAVE_KERNEL(8)
MBG_WAVE_KERNEL(16)
MBG_WAVE_KERNEL(32)
