// Test harness: runs a model's generated HIP source on the host.  The file is
// compiled with -DGEN_FILE="path" and -I to the device headers; __device__ and
// __shared__ are defined away so mbg_eval / mbg_obs_index become plain host
// functions, and mbg_eval_setup fills its tables as a one-thread block.
// stdin: lines "w0 .. w{W-1} u2bits" (hex); stdout: "total H flip obs" with
// floats as hex bits -- the same format as the oracle's --eval-states.  Exits
// with status 3 if mbg_obs_flip(i) is not the observed-index change of
// flipping node i in any input state.
#define __device__
#define __shared__
#include <cstdio>
#include <iostream>
#include <vector>

#include "mbg_abi.h"

static const struct
{
	int x = 0;
} threadIdx;
static const struct
{
	int x = 1;
} blockDim;

#include GEN_FILE

int main()
{
	std::vector<uint32_t> s(MBG_W);
	uint32_t u2b;
	mbg_eval_setup();
	while (true)
	{
		for (int w = 0; w < MBG_W; w++)
			if (!(std::cin >> std::hex >> s[w]))
				return 0;
		std::cin >> std::hex >> u2b;
		mbg_eval_t e;
		mbg_eval(s.data(), mbg_u2f(u2b), &e);
		const uint64_t obs = mbg_obs_index(s.data());
		std::printf("%08x %08x %d %llx\n", mbg_f2u(e.total), mbg_f2u(e.H), e.flip, (unsigned long long)obs);
		// flipping node i toggles exactly mbg_obs_flip(i) in the observed index
		for (int i = 0; i < MBG_N; i++)
		{
			std::vector<uint32_t> f(s);
			f[i / 32] ^= 1u << (i % 32);
			if (mbg_obs_index(f.data()) != (obs ^ mbg_obs_flip(i)))
			{
				std::fprintf(stderr, "mbg_obs_flip(%d) disagrees with mbg_obs_index\n", i);
				return 3;
			}
		}
	}
}
