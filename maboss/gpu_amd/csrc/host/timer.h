// Scope timers printed to stderr when MBG_DIAGS=1 (the reference compiles its
// timer_stats out with a constant, src/timer.cpp:5, src/timer.h:26-41).
#pragma once

#include <chrono>
#include <cstdio>
#include <cstdlib>

namespace mbg
{

inline double wall_seconds()
{
	return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

inline bool diags_enabled()
{
	static const bool on = [] {
		const char* e = std::getenv("MBG_DIAGS");
		return e && e[0] && e[0] != '0';
	}();
	return on;
}

struct scope_timer
{
	const char* name;
	double t0;
	explicit scope_timer(const char* n) : name(n), t0(wall_seconds()) {}
	~scope_timer()
	{
		if (diags_enabled())
			std::fprintf(stderr, "%s: %.3f ms\n", name, (wall_seconds() - t0) * 1e3);
	}
};

} // namespace mbg
