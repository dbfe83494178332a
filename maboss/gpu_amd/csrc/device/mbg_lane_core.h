// Per-trajectory pieces of the lane-per-trajectory kernels (plain and
// batched): initial state, the uniforms of a step, the time advance and the
// folding of one record into the window statistics.  Semantics in
// src/jit_kernels/simulation.cu:34-71, :106-163 and
// src/jit_kernels/window_average_small.cu:50-136; stream layout in
// mbg_numeric.h.
#pragma once

#include "mbg_stats_device.h"

namespace mbg_dev
{

// bit i = (u_i <= istate_i), u_i = draw i of trajectory id's stream
template <int W>
__device__ __forceinline__ void init_state(uint32_t (&s)[W], uint64_t id, uint32_t variant, uint32_t seed_lo,
										   uint32_t seed_hi, int n_nodes, const float* istate)
{
#pragma unroll
	for (int w = 0; w < W; ++w)
	{
		uint32_t word = 0u;
		for (int b = 0; b < 32; b += 4)
		{
			const int i = w * 32 + b;
			if (i >= n_nodes)
				break;
			const mbg_u32x4 r = mbg_philox4x32_10(
				mbg_u32x4 { (uint32_t)(i >> 2), (uint32_t)id, (uint32_t)(id >> 32), variant }, seed_lo, seed_hi);
			const uint32_t rw[4] = { r.x, r.y, r.z, r.w };
#pragma unroll
			for (int j = 0; j < 4; ++j)
				if (i + j < n_nodes && mbg_u01(rw[j]) <= istate[i + j])
					word |= 1u << (b + j);
		}
		s[w] = word;
	}
}

// Initial states for the lanes with `take` set, computed by the whole wave
// (every lane must call): a trajectory needs ceil(N/4) Philox blocks, so
// instead of one lane running them all, each lane runs one block of one
// refilling lane's trajectory (64 / blocks trajectories per round), turns
// its 4 uniforms into 4 state bits at bit 4*(block % 8) of its word, and the
// words are gathered with cross-lane shuffles.  Same draws and comparisons
// as init_state.  Lanes without `take` keep s.
template <int W>
__device__ void init_states_wave(uint32_t (&s)[W], bool take, uint64_t id, uint32_t variant, uint32_t seed_lo,
								 uint32_t seed_hi, int n_nodes, const float* istate)
{
	const int lane = __lane_id();
	const int B = (n_nodes + 3) >> 2; // Philox blocks per trajectory, at most 8 * W
	uint64_t todo = __ballot(take);
	if (B > 64)
	{
		// above 256 nodes: one trajectory per round, its blocks b = lane +
		// 64 * pass (at most 2 passes for 16 words); the 8 blocks of state
		// word w = 8 * pass + (lane >> 3) meet in lane 8 * (w & 7)
		while (todo)
		{
			const int target = __ffsll((unsigned long long)todo) - 1;
			const uint32_t lo = __shfl((uint32_t)id, target), hi = __shfl((uint32_t)(id >> 32), target);
			for (int pass = 0; 64 * pass < B; ++pass)
			{
				const int b = lane + 64 * pass;
				uint32_t v = 0u;
				if (b < B)
				{
					const mbg_u32x4 r = mbg_philox4x32_10(mbg_u32x4 { (uint32_t)b, lo, hi, variant }, seed_lo, seed_hi);
					const uint32_t rw[4] = { r.x, r.y, r.z, r.w };
#pragma unroll
					for (int j = 0; j < 4; ++j)
					{
						const int i = 4 * b + j;
						if (i < n_nodes && mbg_u01(rw[j]) <= istate[i])
							v |= 1u << (4 * (b & 7) + j);
					}
				}
#pragma unroll
				for (int off = 1; off < 8; off <<= 1)
				{
					const uint32_t o = __shfl_down(v, off);
					if ((b & 7) + off < 8 && b + off < B)
						v |= o;
				}
#pragma unroll
				for (int w = 0; w < W; ++w)
					if ((w >> 3) == pass)
					{
						const uint32_t word = __shfl(v, 8 * (w & 7));
						if (lane == target)
							s[w] = 8 * w < B ? word : 0u;
					}
			}
			todo &= todo - 1ull;
		}
		return;
	}
	const int per = 64 / B; // trajectories per round
	const int q = lane / B, b = lane - q * B;
	while (todo)
	{
		// this round: the first `per` lanes of todo
		uint64_t sel = 0, rest = todo;
		for (int k = 0; k < per && rest; ++k)
		{
			sel |= rest & (0ull - rest);
			rest &= rest - 1ull;
		}
		// lane L of this round's q-th trajectory
		uint64_t u = sel;
		for (int k = 0; k < q && u; ++k)
			u &= u - 1ull;
		const int target = u ? __ffsll((unsigned long long)u) - 1 : 0;
		const bool work = u != 0ull && q < per;
		const uint32_t lo = __shfl((uint32_t)id, target), hi = __shfl((uint32_t)(id >> 32), target);
		uint32_t v = 0u;
		if (work)
		{
			const mbg_u32x4 r = mbg_philox4x32_10(mbg_u32x4 { (uint32_t)b, lo, hi, variant }, seed_lo, seed_hi);
			const uint32_t rw[4] = { r.x, r.y, r.z, r.w };
#pragma unroll
			for (int j = 0; j < 4; ++j)
			{
				const int i = 4 * b + j;
				if (i < n_nodes && mbg_u01(rw[j]) <= istate[i])
					v |= 1u << (4 * (b & 7) + j);
			}
		}
		// OR the 8 blocks of each state word into the lane holding its first block
#pragma unroll
		for (int off = 1; off < 8; off <<= 1)
		{
			const uint32_t o = __shfl_down(v, off);
			if ((b & 7) + off < 8 && b + off < B)
				v |= o;
		}
		const bool mine = (sel >> lane) & 1ull;
		const int myq = __popcll(sel & ((1ull << lane) - 1ull));
#pragma unroll
		for (int w = 0; w < W; ++w)
		{
			const uint32_t word = __shfl(v, (myq * B + 8 * w) & 63);
			if (mine)
				s[w] = 8 * w < B ? word : 0u;
		}
		todo &= ~sel;
	}
}

// Trajectory indices for the lanes of a wave from a queue shared by all
// waves, reserved in chunks: the wave keeps its current chunk [next, end)
// (wave-uniform) and takes one atomic per chunk instead of one per refill.
// Chunks are guided: about 1/4 of the wave's share of what the queue had
// left at its last reservation, at most max_chunk, so waves run out of work
// at nearly the same time.  Lanes in `need` receive consecutive indices in
// lane order; an index at or beyond n_total means the queue is drained.
// Every lane must call.
struct id_chunks
{
	uint64_t next = 0, end = 0;

	__device__ uint64_t take(uint64_t need, unsigned long long* queue, uint64_t n_total, uint64_t max_chunk)
	{
		const int lane = __lane_id();
		const uint64_t n = (uint64_t)__popcll(need);
		const uint64_t r = (uint64_t)__popcll(need & ((1ull << lane) - 1ull));
		const uint64_t avail = end - next;
		uint64_t mine = next + r;
		if (n > avail)
		{
			const uint64_t waves = (uint64_t)gridDim.x * (blockDim.x >> 6);
			const uint64_t left = end < n_total ? n_total - end : 0;
			uint64_t chunk = left / (4 * waves);
			chunk = chunk > max_chunk ? max_chunk : chunk;
			const uint64_t want = n - avail > chunk ? n - avail : chunk;
			const int leader = __ffsll((unsigned long long)need) - 1;
			uint64_t base = 0;
			if (lane == leader)
				base = atomicAdd(queue, (unsigned long long)want);
			base = bcast64(base, leader);
			if (r >= avail)
				mine = base + (r - avail);
			next = base + (n - avail);
			end = base + want;
		}
		else
			next += n;
		return mine;
	}
};

// u1 (dwell) and u2 (flip) of step k; discrete time draws u2 only
__device__ __forceinline__ void step_uniforms(bool discrete, uint32_t B0, uint32_t k, uint64_t id, uint32_t variant,
											  uint32_t seed_lo, uint32_t seed_hi, float& u1, float& u2)
{
	if (discrete)
	{
		const mbg_u32x4 r = mbg_philox4x32_10(mbg_u32x4 { B0 + (k >> 2), (uint32_t)id, (uint32_t)(id >> 32), variant },
											  seed_lo, seed_hi);
		u1 = 1.0f;
		u2 = mbg_u01(pick(r, k & 3u));
	}
	else
	{
		const mbg_u32x4 r = mbg_philox4x32_10(mbg_u32x4 { B0 + (k >> 1), (uint32_t)id, (uint32_t)(id >> 32), variant },
											  seed_lo, seed_hi);
		u1 = mbg_u01((k & 1u) ? r.z : r.x);
		u2 = mbg_u01((k & 1u) ? r.w : r.y);
	}
}

// New time after a step with this total rate (simulation.cu:114-131)
__device__ __forceinline__ float advance_time(bool discrete, float t, float total, float u1, float tick, float max_time,
											  bool& fixed)
{
	if (total == 0.0f)
	{
		fixed = true;
		return max_time;
	}
	if (discrete)
	{
		t = t + tick;
		t = (max_time - tick < t) ? max_time : t;
	}
	else
		t = t + (-mbg_logf(u1)) / total;
	return fminf(t, max_time);
}

// Register-combined contributions of one lane to the window cells and
// entropy sums; flushed through the callbacks when the cell changes.
struct pending
{
	int pw = -1, pew = -1;
	uint64_t pobs = 0, pq = 0, peq = 0;

	template <typename FW>
	__device__ __forceinline__ void win(int w, uint64_t obs, uint64_t q, FW&& flush)
	{
		if (w == pw && obs == pobs)
			pq += q;
		else
		{
			if (pw >= 0)
				flush(pw, pobs, pq);
			pw = w;
			pobs = obs;
			pq = q;
		}
	}

	template <typename FE>
	__device__ __forceinline__ void ent(int w, uint64_t q, FE&& flush)
	{
		if (w == pew)
			peq += q;
		else
		{
			if (pew >= 0)
				flush(pew, peq);
			pew = w;
			peq = q;
		}
	}

	template <typename FW, typename FE>
	__device__ __forceinline__ void flush_all(FW&& fw, FE&& fe)
	{
		if (pw >= 0)
			fw(pw, pobs, pq);
		if (pew >= 0)
			fe(pew, peq);
		pw = pew = -1;
	}
};

// Fold the record (state obs during [tp, t), entropy H) into the windows:
// continuous time splits the dwell over windows (window_average_small.cu:72-91),
// discrete time counts one step at window round(tp / tick) (:123-129).
// With final_tail (the trajectory's last record, t == max_time), the windows
// after the first one get no per-window addition: ft(w, obs) records them
// once, to be expanded with the per-window quantum after the run.
template <typename FW, typename FE, typename FT>
__device__ __forceinline__ void record(bool discrete, float tp, float t, uint64_t obs, float H, float tick, float qscale,
									   int windows, pending& p, FW&& fw, FE&& fe, FT&& ft, bool final_tail)
{
	if (discrete)
	{
		const int w = (int)roundf(tp / tick);
		if (w < windows)
		{
			p.win(w, obs, 1ull, fw);
			const uint64_t qe = (uint64_t)rintf(H * qscale);
			if (qe)
				p.ent(w, qe, fe);
		}
		return;
	}
	int w = (int)floorf(tp / tick);
	const int w_first = w;
	bool tail_done = false;
	float b = tp;
	while (t > b)
	{
		const float wend = (float)(w + 1) * tick;
		const float en = fminf(t, wend);
		const float sl = en - b;
		if (w < windows)
		{
			if (final_tail && w > w_first)
			{
				if (!tail_done)
				{
					ft(w, obs);
					tail_done = true;
				}
			}
			else
			{
				const uint64_t q = (uint64_t)rintf(sl * qscale);
				if (q)
					p.win(w, obs, q, fw);
			}
			const uint64_t qe = (uint64_t)rintf((H * sl) * qscale);
			if (qe)
				p.ent(w, qe, fe);
		}
		w++;
		b = en;
		// once the tail is recorded, the later windows only add entropy: none when H is 0
		if (tail_done && H == 0.0f)
			break;
	}
}

// Inclusive prefix sum over the wave's 64 lanes with DPP moves (all lanes
// active): within rows of 16 by row_shr 1, 2, 4, 8, then rows 1 and 3 add
// lane 15 of the row before (row_bcast:15) and rows 2 and 3 add lane 31
// (row_bcast:31).  A move with no source lane adds 0.
template <int CTRL, int ROWS>
__device__ __forceinline__ int add_dpp_i(int v)
{
	return v + __builtin_amdgcn_update_dpp(0, v, CTRL, ROWS, 0xF, false);
}

__device__ __forceinline__ int wave_incl_scan(int v)
{
	v = add_dpp_i<0x111, 0xF>(v);
	v = add_dpp_i<0x112, 0xF>(v);
	v = add_dpp_i<0x114, 0xF>(v);
	v = add_dpp_i<0x118, 0xF>(v);
	v = add_dpp_i<0x142, 0xA>(v);
	v = add_dpp_i<0x143, 0xC>(v);
	return v;
}

// record() for the whole wave at once: every lane calls (on: the lane has a
// record) and every window cell completed reaches emit(has, w, obs, q) at a
// point where the whole wave is active (emit can then stage it with
// wave-wide operations).  Continuous time: the lane itself handles the
// record's first and last windows (the ones its pending cell and pending
// entropy may merge with); the full windows in between are spread over the
// wave's lanes, one window per lane, so a record that spans many windows
// does not keep the rest of the wave waiting.  Same cell and entropy sums as
// record() (the same slices: for each window w after the first, the slice is
// min(t, (w + 1) tick) - w tick in FP32), added in another order.
template <typename EMIT, typename FE, typename FT>
__device__ __forceinline__ void record_wave(bool on, bool discrete, float tp, float t, uint64_t obs, float H, float tick,
											float qscale, int windows, pending& p, EMIT&& emit, FE&& fe, FT&& ft,
											bool final_tail, uint8_t* own)
{
	// the lane's pending cell when the record moves to another one
	auto merge = [&](int w, uint64_t q, bool& has, int& ow, uint64_t& oo, uint64_t& oq) {
		if (w == p.pw && obs == p.pobs)
			p.pq += q;
		else
		{
			if (p.pw >= 0)
			{
				has = true;
				ow = p.pw;
				oo = p.pobs;
				oq = p.pq;
			}
			p.pw = w;
			p.pobs = obs;
			p.pq = q;
		}
	};
	if (discrete)
	{
		bool has = false;
		int ow = 0;
		uint64_t oo = 0, oq = 0;
		if (on)
		{
			const int w = (int)roundf(tp / tick);
			if (w < windows)
			{
				merge(w, 1ull, has, ow, oo, oq);
				const uint64_t qe = (uint64_t)rintf(H * qscale);
				if (qe)
					p.ent(w, qe, fe);
			}
		}
		emit(has, ow, oo, oq);
		return;
	}

	bool has0 = false, has1 = false, cells = false;
	int ow0 = 0, ow1 = 0;
	uint64_t oo0 = 0, oq0 = 0, oo1 = 0, oq1 = 0;
	int ra = 0, len = 0; // windows [ra, ra + len) go to the wave
	if (on && t > tp)
	{
		// first window: slice [tp, min(t, end of window))
		const int w0 = (int)floorf(tp / tick);
		const float en0 = fminf(t, (float)(w0 + 1) * tick);
		if (w0 < windows)
		{
			const float sl = en0 - tp;
			const uint64_t q = (uint64_t)rintf(sl * qscale);
			if (q)
				merge(w0, q, has0, ow0, oo0, oq0);
			const uint64_t qe = (uint64_t)rintf((H * sl) * qscale);
			if (qe)
				p.ent(w0, qe, fe);
		}
		if (t > en0)
		{
			// the record reaches windows w0 + 1 .. wl: wl is the largest w with
			// w tick < t in FP32 (record() visits w while its start is below t)
			int wl = (int)floorf(t / tick);
			while (wl > w0 + 1 && (float)wl * tick >= t)
				--wl;
			if (wl < w0 + 1)
				wl = w0 + 1;
			while ((float)(wl + 1) * tick < t)
				++wl;
			if (final_tail)
			{
				// the trajectory's last record: the tail table takes the
				// windows' dwell; they still add entropy
				if (w0 + 1 < windows)
				{
					ft(w0 + 1, obs);
					if (H != 0.0f)
					{
						ra = w0 + 1;
						len = min(wl, windows - 1) - ra + 1;
					}
				}
			}
			else
			{
				cells = true;
				ra = w0 + 1;
				len = max(min(wl - 1, windows - 1) - ra + 1, 0);
				if (wl < windows)
				{
					const float sl = fminf(t, (float)(wl + 1) * tick) - (float)wl * tick;
					const uint64_t q = (uint64_t)rintf(sl * qscale);
					if (q)
						merge(wl, q, has1, ow1, oo1, oq1);
					const uint64_t qe = (uint64_t)rintf((H * sl) * qscale);
					if (qe)
						p.ent(wl, qe, fe);
				}
			}
		}
	}
	emit(has0, ow0, oo0, oq0);
	emit(has1, ow1, oo1, oq1);
	if (__ballot(len > 0) == 0ull)
		return;

	// the windows in between, 64 at a time: item j of the wave's concatenated
	// ranges goes to lane j % 64; own[] (64 bytes of LDS for the wave) maps
	// the items of a pass to the lanes whose ranges hold them
	const int lane = __lane_id();
	const int incl = wave_incl_scan(len);
	const int excl = incl - len;
	const int total = __builtin_amdgcn_readlane(incl, 63);
	// window of item j: base + j with base = ra - excl (may be negative); bit 0: cells
	const int rx = (ra - excl) * 2 + (cells ? 1 : 0);
	for (int base = 0; base < total; base += 64)
	{
		const int hi = min(incl, base + 64);
		for (int i = max(excl, base); i < hi; ++i)
			own[i - base] = (uint8_t)lane;
		MBG_WAVE_SYNC();
		const int j = base + lane;
		const bool valid = j < total;
		const int src = valid ? (int)own[lane] : 0;
		MBG_WAVE_SYNC();
		const int rxs = __shfl(rx, src);
		const float ts = __shfl(t, src), Hs = __shfl(H, src);
		const uint64_t os = __shfl(obs, src);
		const int w = (rxs - (rxs & 1)) / 2 + j;
		const float sl = fminf(ts, (float)(w + 1) * tick) - (float)w * tick;
		const uint64_t q = (valid && (rxs & 1)) ? (uint64_t)rintf(sl * qscale) : 0ull;
		emit(q != 0ull, w, os, q);
		if (valid)
		{
			const uint64_t qe = (uint64_t)rintf((Hs * sl) * qscale);
			if (qe)
				fe(w, qe);
		}
	}
}

} // namespace mbg_dev
