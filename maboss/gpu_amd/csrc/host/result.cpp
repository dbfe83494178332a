// Results: normalisation and the reference's three output forms.
//
//   * _probtraj.csv: src/statistics/window_average_small.cpp:181-250
//   * _fp.csv:       src/statistics/fixed_states.cu:154-187
//   * stdout report: src/statistics/final_states.cpp:53-66,
//                    src/statistics/fixed_states.cu:140-152,
//                    src/statistics/window_average_small.cpp:142-179
//
// Window cells arrive as exact unsigned fixed-point sums (dwell * 2^32 / tick
// per slice, or step counts in discrete time); probabilities are formed here
// in double and printed as float, like the reference's float results.
#include <algorithm>
#include <cmath>
#include <fstream>
#include <iostream>

#include "engine.h"

namespace mbg
{

double result::prob(uint64_t raw) const
{
	// continuous: sum(dwell)/(n*tick) = raw / (n * 2^32); discrete: count / n
	return discrete ? (double)raw / norm : (double)raw / (norm * 4294967296.0);
}

double result::th(int w) const { return (double)ent[w] / (norm * 4294967296.0); }

double result::h(int w) const
{
	float e = 0.f;
	for (uint64_t i = woff[w]; i < woff[w + 1]; i++)
	{
		const float p = (float)prob(win.v(i));
		if (p == 0.f)
			continue;
		e += -std::log2(p) * p;
	}
	return e;
}

std::string result::state_name(uint64_t obs) const
{
	std::string s;
	for (size_t j = 0; j < observed.size(); j++)
		if ((obs >> j) & 1ull)
		{
			if (!s.empty())
				s += " -- ";
			s += names[observed[j]];
		}
	return s.empty() ? "<nil>" : s;
}

std::string result::full_state_name(const uint32_t* st) const
{
	std::string s;
	for (size_t i = 0; i < names.size(); i++)
		if ((st[i / 32] >> (i % 32)) & 1u)
		{
			if (!s.empty())
				s += " -- ";
			s += names[i];
		}
	return s.empty() ? "<nil>" : s;
}

void result::index_cells()
{
	// cells sort by (window << n_obs) | state, so window w starts at the
	// first key >= w << n_obs; keys at or past windows << n_obs (the sorted
	// mode's padding) belong to no window
	woff.assign((size_t)windows + 1, 0);
	for (int w = 0; w <= windows; w++)
		woff[w] = win.lower_bound((uint64_t)w << n_obs);
}

void result::sort_fixed()
{
	const size_t W = (size_t)n_words, n = fp_count.size();
	std::vector<size_t> ord(n);
	for (size_t i = 0; i < n; i++)
		ord[i] = i;
	// reference order: most significant word first (src/state.h:43-49)
	std::sort(ord.begin(), ord.end(), [&](size_t x, size_t y) {
		const uint32_t* a = fp_words.data() + x * W;
		const uint32_t* b = fp_words.data() + y * W;
		for (size_t w = W; w-- > 0;)
			if (a[w] != b[w])
				return a[w] < b[w];
		return false;
	});
	std::vector<uint32_t> words(n * W);
	std::vector<uint64_t> count(n);
	for (size_t i = 0; i < n; i++)
	{
		std::copy(fp_words.begin() + (std::ptrdiff_t)(ord[i] * W), fp_words.begin() + (std::ptrdiff_t)((ord[i] + 1) * W),
				  words.begin() + (std::ptrdiff_t)(i * W));
		count[i] = fp_count[ord[i]];
	}
	fp_words.swap(words);
	fp_count.swap(count);
}

void result::wait() const
{
	if (on_ready)
	{
		std::function<void()> f;
		f.swap(on_ready);
		f();
	}
}

result::~result()
{
	// a copy still in flight writes into memory this result holds
	try
	{
		wait();
	}
	catch (...)
	{
	}
}

void result::write_csv(const std::string& prefix, bool with_finals) const
{
	{
		std::ofstream ofs(prefix + "_probtraj.csv");
		if (!ofs)
			throw error(MBG_E_IO, "cannot write " + prefix + "_probtraj.csv");
		size_t max_states = 0;
		for (int w = 0; w < windows; w++)
		{
			size_t c = 0;
			for (uint64_t i = woff[w]; i < woff[w + 1]; i++)
				c += (float)prob(win.v(i)) != 0.f;
			max_states = std::max(max_states, c);
		}
		ofs << "Time\tTH\tErrorTH\tH\tHD=0";
		for (size_t i = 0; i < max_states; i++)
			ofs << "\tState\tProba\tErrorProba";
		ofs << "\n";
		const uint64_t om = obs_mask();
		for (int w = 0; w < windows; w++)
		{
			// the error columns are 0 without second moments, as the reference prints them
			ofs << (float)w * tick << "\t" << (float)th(w) << "\t" << (float)error_th(w) << "\t" << (float)h(w) << "\t"
				<< 0.f;
			for (uint64_t i = woff[w]; i < woff[w + 1]; i++)
			{
				const float p = (float)prob(win.v(i));
				if (p == 0.f)
					continue;
				ofs << "\t" << state_name(win.k(i) & om) << "\t" << p << "\t"
					<< (float)error_prob(w, win.k(i) & om, win.v(i));
			}
			ofs << "\n";
		}
	}
	{
		std::ofstream ofs(prefix + "_fp.csv");
		if (!ofs)
			throw error(MBG_E_IO, "cannot write " + prefix + "_fp.csv");
		ofs << "Fixed Points (" << n_fixed() << ")\n";
		ofs << "FP\tProba\tState";
		for (auto& n : names)
			ofs << "\t" << n;
		ofs << "\n";
		for (uint64_t i = 0; i < n_fixed(); i++)
		{
			const uint32_t* st = fixed_state(i);
			ofs << "#" << i << "\t" << (float)((double)fp_count[i] / norm) << "\t" << full_state_name(st);
			for (size_t k = 0; k < names.size(); k++)
				ofs << "\t" << ((st[k / 32] >> (k % 32)) & 1u);
			ofs << "\n";
		}
	}
	if (with_finals)
	{
		// The reference's final_states_stats::write_csv is empty
		// (src/statistics/final_states.cpp:68); this file is an addition.
		std::ofstream ofs(prefix + "_finalstates.csv");
		if (!ofs)
			throw error(MBG_E_IO, "cannot write " + prefix + "_finalstates.csv");
		ofs << "State\tProba\n";
		for (uint64_t i = 0; i < fin.n; i++)
			ofs << state_name(fin.k(i)) << "\t" << (float)((double)fin.v(i) / norm) << "\n";
	}
}

void result::print(std::ostream& os) const
{
	os << "final points:\n";
	for (uint64_t i = 0; i < fin.n; i++)
		os << (float)((double)fin.v(i) / norm) << " " << state_name(fin.k(i)) << "\n";
	os << "fixed points:\n";
	for (uint64_t i = 0; i < n_fixed(); i++)
		os << (float)((double)fp_count[i] / norm) << " " << full_state_name(fixed_state(i)) << "\n";
	const uint64_t om = obs_mask();
	for (int w = 0; w < windows; w++)
	{
		os << "window (" << (float)w * tick << ", " << (float)(w + 1) * tick << "]\n";
		os << "entropy: " << (float)h(w) << "\n";
		os << "transition entropy: " << (float)th(w) << "\n";
		for (uint64_t i = woff[w]; i < woff[w + 1]; i++)
		{
			const float p = (float)prob(win.v(i));
			if (p == 0.f)
				continue;
			os << p << " " << state_name(win.k(i) & om) << "\n";
		}
	}
	os.flush();
}

void result::write_raw(std::ostream& os) const
{
	os << "MBG_RAW 1\n";
	os << "windows " << windows << " n_obs " << n_obs << " discrete " << discrete << " n " << n_traj << "\n";
	os << "T " << transitions << " " << finished << " " << truncated << "\n";
	for (int w = 0; w < windows; w++)
		if (ent[w])
			os << "E " << w << " " << ent[w] << "\n";
	const uint64_t om = obs_mask();
	for (int w = 0; w < windows; w++)
		for (uint64_t i = woff[w]; i < woff[w + 1]; i++)
			os << "W " << w << " " << (win.k(i) & om) << " " << win.v(i) << "\n";
	for (uint64_t i = 0; i < fin.n; i++)
		os << "F " << fin.k(i) << " " << fin.v(i) << "\n";
	if (has_moments)
	{
		// second moments: R w hi lo (entropy), Q w obs hi lo (cells)
		auto u128s = [](unsigned __int128 v) {
			return std::to_string((uint64_t)(v >> 64)) + " " + std::to_string((uint64_t)v);
		};
		for (int w = 0; w < windows; w++)
			if (mesq[w])
				os << "R " << w << " " << u128s(mesq[w]) << "\n";
		for (uint64_t c = 0; c < msq.size(); c++)
			if (msq[c])
				os << "Q " << (c >> n_obs) << " " << (c & om) << " " << u128s(msq[c]) << "\n";
	}
	for (uint64_t i = 0; i < n_fixed(); i++)
	{
		os << "P";
		const uint32_t* st = fixed_state(i);
		for (int w = 0; w < n_words; w++)
			os << " " << st[w];
		os << " " << fp_count[i] << "\n";
	}
}

} // namespace mbg
