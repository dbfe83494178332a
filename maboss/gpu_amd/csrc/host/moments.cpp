// ErrorTH / ErrorProba: the second moments of a run's window statistics from
// a replay of its trajectories (mbg_moments_kernel.h), and the standard
// errors computed from them.  The reference prints both columns as 0
// (src/statistics/window_average_small.cpp:235, :244-245).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <tuple>

#include "../device/mbg_kernel_args.h"
#include "engine.h"
#include "generator.h"
#include "timer.h"

namespace mbg
{

namespace
{

constexpr uint64_t k_max_moment_cells = 1ull << 24;

size_t up16(size_t n) { return (n + 15) & ~(size_t)15; }

uint64_t next_pow2(uint64_t v)
{
	uint64_t c = 1;
	while (c < v)
		c <<= 1;
	return c;
}

void fill_words(program& p, hipStream_t stream, void* ptr, uint64_t words)
{
	if (!words)
		return;
	p.load_util();
	uint64_t value = 0;
	void* args[] = { &ptr, &words, &value };
	const unsigned grid = (unsigned)std::min<uint64_t>(1024, (words + 255) / 256);
	hip_check(hipModuleLaunchKernel(p.fn_fill_u64, grid, 1, 1, 256, 1, 1, 0, stream, args, nullptr),
			  "hipModuleLaunchKernel(mbg_fill_u64)");
}

unsigned __int128 u128_of(const unsigned long long* w) { return ((unsigned __int128)w[1] << 64) | w[0]; }

} // namespace

// Standard error of the mean over n trajectories from exact sums in
// fixed-point units of `unit`: sqrt((E[x^2] - E[x]^2) / (n - 1)); 0 for
// n < 2 or a variance that rounds below 0.
double standard_error(uint64_t s1, unsigned __int128 s2, double n, double unit)
{
	if (n < 2.0)
		return 0.0;
	const double mean = (double)s1 / (n * unit);
	const double m2 = ((double)(uint64_t)(s2 >> 64) * 18446744073709551616.0 + (double)(uint64_t)s2) / (n * unit * unit);
	const double var = (m2 - mean * mean) / (n - 1.0);
	return var > 0.0 ? std::sqrt(var) : 0.0;
}

double result::error_th(int w) const
{
	if (!has_moments || w < 0 || w >= windows)
		return 0.0;
	return standard_error(ent[w], mesq[w], (double)n_traj, 4294967296.0);
}

double result::error_prob(int w, uint64_t obs, uint64_t raw) const
{
	if (!has_moments || w < 0 || w >= windows)
		return 0.0;
	return standard_error(raw, msq[((uint64_t)w << n_obs) | obs], (double)n_traj, discrete ? 1.0 : 4294967296.0);
}

void run_moments(program& p, const sim_params& sp, hipStream_t stream, result& r)
{
	scope_timer tm("moments> replay");
	const model& m = *p.mdl;
	if (r.kernel != 0 || m.n_words() > max_lane_words)
		throw error(MBG_E_LIMIT, "error columns are computed for lane-kernel runs (up to 512 nodes)");
	const int n_obs = m.n_observed(), windows = m.windows(), N = m.n_nodes(), W = m.n_words();
	if (n_obs > 24 || ((uint64_t)windows << n_obs) > k_max_moment_cells)
		throw error(MBG_E_LIMIT, "error columns need windows x 2^n_obs <= 2^24 window cells");
	const uint64_t cells = (uint64_t)windows << n_obs;
	const uint64_t n_traj = sp.n_traj < 0 ? (uint64_t)m.sample_count() : (uint64_t)sp.n_traj;
	const uint64_t seed = sp.seed < 0 ? m.seed() : (uint64_t)sp.seed;
	const int np = (int)p.gen.param_names.size();
	hipFunction_t fn = p.moments_function(sp.moment_states);

	// rates (the run's or the model's) and initial-state probabilities
	std::vector<float> rates = sp.rates;
	if (rates.empty())
		for (int k = 0; k < np; k++)
			rates.push_back(m.variables.at(p.gen.param_names[k]));
	if ((int)rates.size() != np)
		throw error(MBG_E_ARG, "rates must have one value per rate parameter");

	// workspace: counters | view | rates | rate*log2(rate) | masks | istate | esq | sq | spill
	const size_t o_view = 256, o_rate = o_view + up16(sizeof(mbg_param_view)), o_lr = o_rate + up16(4 * (size_t)np + 4),
				 o_en = o_lr + up16(4 * (size_t)np + 4), o_ist = o_en + up16(8 * (size_t)W),
				 o_esq = o_ist + up16(4 * (size_t)N), o_sq = o_esq + 16 * (size_t)windows, o_spill = o_sq + 16 * cells;
	uint64_t spill_cap = std::min<uint64_t>(1ull << 22, std::max<uint64_t>(1ull << 12, n_traj));
	p.order_after_last_run(stream);
	std::shared_ptr<char> up = p.pool->get(o_esq - o_rate);
	std::memset(up.get(), 0, o_esq - o_rate);
	std::memcpy(up.get(), rates.data(), 4 * (size_t)np);
	for (int i = 0; i < N; i++)
		std::memcpy(up.get() + (o_ist - o_rate) + 4 * (size_t)i, &m.nodes[i].istate, 4);

	for (int attempt = 0;; attempt++)
	{
		char* ws = (char*)p.b_mom.ensure(o_spill + 32 * spill_cap);
		hip_check(hipMemcpyAsync(ws + o_rate, up.get(), o_esq - o_rate, hipMemcpyHostToDevice, stream), "hipMemcpyAsync");
		fill_words(p, stream, ws, MBG_MOM_NCOUNTERS);
		fill_words(p, stream, ws + o_esq, 2 * (uint64_t)windows + 2 * cells);
		{
			void* view = ws + o_view;
			const float* rt = (const float*)(ws + o_rate);
			float* lr = (float*)(ws + o_lr);
			uint32_t* en = (uint32_t*)(ws + o_en);
			int nw = W, npar = np;
			void* args[] = { &view, &rt, &lr, &en, &npar, &nw };
			const unsigned g = (unsigned)std::max(1, (std::max(np, 2 * W) + 255) / 256);
			hip_check(hipModuleLaunchKernel(p.fn_rate_view, g, 1, 1, 256, 1, 1, 0, stream, args, nullptr),
					  "hipModuleLaunchKernel(mbg_rate_view)");
		}
		// the view is built on the device; the kernel takes it by value
		mbg_param_view view;
		hip_check(hipMemcpyAsync(&view, ws + o_view, sizeof(view), hipMemcpyDeviceToHost, stream), "hipMemcpyAsync");
		hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize");

		mbg_moments_args a;
		std::memset(&a, 0, sizeof(a));
		a.n_traj = n_traj;
		a.id_base = sp.first_id;
		a.seed_lo = (uint32_t)seed;
		a.seed_hi = (uint32_t)(seed >> 32);
		a.variant = sp.variant;
		a.max_steps = sp.max_steps ? sp.max_steps : (1u << 30);
		a.n_nodes = N;
		a.n_obs = n_obs;
		a.windows = windows;
		a.discrete = m.discrete_time();
		a.max_time = m.max_time();
		a.tick = m.time_tick();
		a.qscale = a.discrete ? 0x1p32f : 0x1p32f / a.tick;
		a.refill_lanes = default_refill_lanes(N);
		a.istate = (const float*)(ws + o_ist);
		a.pv = view;
		a.sq = (unsigned long long*)(ws + o_sq);
		a.esq = (unsigned long long*)(ws + o_esq);
		a.spill = (unsigned long long*)(ws + o_spill);
		a.spill_cap = spill_cap;
		a.counters = (unsigned long long*)ws;

		int per_cu = 0;
		hip_check(hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, 256, 0),
				  "hipModuleOccupancyMaxActiveBlocksPerMultiprocessor");
		const uint64_t need = (n_traj + 255) / 256;
		const unsigned grid =
			(unsigned)std::max<uint64_t>(1, std::min<uint64_t>(need, (uint64_t)std::max(per_cu, 1) * p.num_cus));
		if (n_traj)
		{
			void* args[] = { &a };
			hip_check(hipModuleLaunchKernel(fn, grid, 1, 1, 256, 1, 1, 0, stream, args, nullptr),
					  "hipModuleLaunchKernel(simulate_moments)");
		}
		// counters, esq and sq in one copy; the spill records after the count is known
		const size_t back = o_spill - o_esq;
		std::shared_ptr<char> host = p.pool->get(back + 256);
		hip_check(hipMemcpyAsync(host.get(), ws, 256, hipMemcpyDeviceToHost, stream), "hipMemcpyAsync");
		hip_check(hipMemcpyAsync(host.get() + 256, ws + o_esq, back, hipMemcpyDeviceToHost, stream), "hipMemcpyAsync");
		hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize");
		const uint64_t spilled = ((const unsigned long long*)host.get())[MBG_MOM_CTR_SPILL];
		if (spilled > spill_cap)
		{
			if (attempt >= 4)
				throw error(MBG_E_LIMIT, "second-moment spill list overflow");
			spill_cap = next_pow2(spilled);
			continue;
		}
		std::vector<unsigned long long> rec(4 * spilled);
		if (spilled)
		{
			hip_check(hipMemcpyAsync(rec.data(), ws + o_spill, 32 * spilled, hipMemcpyDeviceToHost, stream),
					  "hipMemcpyAsync");
			hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize");
		}
		const auto* e = (const unsigned long long*)(host.get() + 256);
		r.mesq.resize(windows);
		for (int w = 0; w < windows; w++)
			r.mesq[w] = u128_of(e + 2 * w);
		const auto* q = e + 2 * windows;
		r.msq.resize(cells);
		for (uint64_t i = 0; i < cells; i++)
			r.msq[i] = u128_of(q + 2 * i);
		// spilled partial times: one trajectory's total per (id, window, state), then squared
		std::vector<uint64_t> ord(spilled);
		for (uint64_t i = 0; i < spilled; i++)
			ord[i] = i;
		auto key = [&](uint64_t i) { return std::make_tuple(rec[4 * i], rec[4 * i + 1], rec[4 * i + 2]); };
		std::sort(ord.begin(), ord.end(), [&](uint64_t x, uint64_t y) { return key(x) < key(y); });
		for (uint64_t i = 0; i < spilled;)
		{
			uint64_t j = i, tot = 0;
			while (j < spilled && key(ord[j]) == key(ord[i]))
				tot += rec[4 * ord[j++] + 3];
			const uint64_t cell = (rec[4 * ord[i] + 1] << n_obs) | rec[4 * ord[i] + 2];
			r.msq[cell] += (unsigned __int128)tot * tot;
			i = j;
		}
		r.moment_spills = spilled;
		r.has_moments = true;
		break;
	}
	p.mark_run_done(stream);
}

} // namespace mbg
