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
	for (auto& c : win[w])
	{
		const float p = (float)prob(c.second);
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

std::string result::full_state_name(const std::vector<uint32_t>& st) const
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
			for (auto& cell : win[w])
				c += (float)prob(cell.second) != 0.f;
			max_states = std::max(max_states, c);
		}
		ofs << "Time\tTH\tErrorTH\tH\tHD=0";
		for (size_t i = 0; i < max_states; i++)
			ofs << "\tState\tProba\tErrorProba";
		ofs << "\n";
		for (int w = 0; w < windows; w++)
		{
			ofs << (float)w * tick << "\t" << (float)th(w) << "\t" << 0.f << "\t" << (float)h(w) << "\t" << 0.f;
			for (auto& cell : win[w])
			{
				const float p = (float)prob(cell.second);
				if (p == 0.f)
					continue;
				ofs << "\t" << state_name(cell.first) << "\t" << p << "\t" << 0.f;
			}
			ofs << "\n";
		}
	}
	{
		std::ofstream ofs(prefix + "_fp.csv");
		if (!ofs)
			throw error(MBG_E_IO, "cannot write " + prefix + "_fp.csv");
		ofs << "Fixed Points (" << fixed.size() << ")\n";
		ofs << "FP\tProba\tState";
		for (auto& n : names)
			ofs << "\t" << n;
		ofs << "\n";
		int i = 0;
		for (auto& fp : fixed)
		{
			ofs << "#" << i++ << "\t" << (float)((double)fp.second / norm) << "\t" << full_state_name(fp.first);
			for (size_t k = 0; k < names.size(); k++)
				ofs << "\t" << ((fp.first[k / 32] >> (k % 32)) & 1u);
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
		for (auto& f : fin)
			ofs << state_name(f.first) << "\t" << (float)((double)f.second / norm) << "\n";
	}
}

void result::print(std::ostream& os) const
{
	os << "final points:\n";
	for (auto& f : fin)
		os << (float)((double)f.second / norm) << " " << state_name(f.first) << "\n";
	os << "fixed points:\n";
	for (auto& fp : fixed)
		os << (float)((double)fp.second / norm) << " " << full_state_name(fp.first) << "\n";
	for (int w = 0; w < windows; w++)
	{
		os << "window (" << (float)w * tick << ", " << (float)(w + 1) * tick << "]\n";
		os << "entropy: " << (float)h(w) << "\n";
		os << "transition entropy: " << (float)th(w) << "\n";
		for (auto& cell : win[w])
		{
			const float p = (float)prob(cell.second);
			if (p == 0.f)
				continue;
			os << p << " " << state_name(cell.first) << "\n";
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
	for (int w = 0; w < windows; w++)
		for (auto& c : win[w])
			os << "W " << w << " " << c.first << " " << c.second << "\n";
	for (auto& f : fin)
		os << "F " << f.first << " " << f.second << "\n";
	for (auto& fp : fixed)
	{
		os << "P";
		for (uint32_t x : fp.first)
			os << " " << x;
		os << " " << fp.second << "\n";
	}
}

} // namespace mbg
