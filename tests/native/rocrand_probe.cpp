// Test probe: rocRAND's open-source Philox4x32-10 rounds, as an independent
// implementation for the known-answer cross-check.  Reads "c0 c1 c2 c3 k0 k1"
// (hex) lines, prints the 4 output words.
#include <cstdio>
#include <iostream>

#include <rocrand/rocrand_philox4x32_10.h>

struct probe : rocrand_device::philox4x32_10_engine
{
	uint4 run(uint4 c, uint2 k) { return ten_rounds(c, k); }
};

int main()
{
	probe p;
	unsigned c0, c1, c2, c3, k0, k1;
	while (std::cin >> std::hex >> c0 >> c1 >> c2 >> c3 >> k0 >> k1)
	{
		uint4 r = p.run(uint4 { c0, c1, c2, c3 }, uint2 { k0, k1 });
		std::printf("%08x %08x %08x %08x\n", r.x, r.y, r.z, r.w);
	}
	return 0;
}
