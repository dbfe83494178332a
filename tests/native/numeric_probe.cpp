// Test probe: the engine's numeric primitives (maboss/gpu_amd/csrc/device/
// mbg_numeric.h) compiled for the host, same stdin protocol as the oracle's
// --numeric-probe: "P c0 c1 c2 c3 k0 k1", "L bits", "U x" (hex).
#include <cstdio>
#include <iostream>

#include "mbg_numeric.h"

int main()
{
	char op;
	while (std::cin >> op)
	{
		if (op == 'P')
		{
			mbg_u32x4 c;
			uint32_t k0, k1;
			std::cin >> std::hex >> c.x >> c.y >> c.z >> c.w >> k0 >> k1 >> std::dec;
			mbg_u32x4 r = mbg_philox4x32_10(c, k0, k1);
			std::printf("%08x %08x %08x %08x\n", r.x, r.y, r.z, r.w);
		}
		else if (op == 'L')
		{
			uint32_t b;
			std::cin >> std::hex >> b >> std::dec;
			std::printf("%08x\n", mbg_f2u(mbg_log2f(mbg_u2f(b))));
		}
		else if (op == 'U')
		{
			uint32_t x;
			std::cin >> std::hex >> x >> std::dec;
			std::printf("%08x\n", mbg_f2u(mbg_u01(x)));
		}
	}
	return 0;
}
