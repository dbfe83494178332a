// Test harness: runs a model's generated HIP source on the host.  The file is
// compiled with -DGEN_FILE="path" and -I to the device headers; __device__ and
// __shared__ are defined away so mbg_eval / mbg_obs_index become plain host
// functions, and mbg_eval_setup fills its tables as a one-thread block.
// stdin: lines "w0 .. w{W-1} u2bits" (hex); stdout: "total H flip obs" with
// floats as hex bits -- the same format as the oracle's --eval-states.  Exits
// with status 3 if mbg_obs_flip(i) is not the observed-index change of
// flipping node i in any input state.
//
// With -DGEN_PARAM the rates come from a run-time table (mbg_eval_p): stdin
// starts with "np", np rate bit patterns, W up_en words and W dn_en words
// (hex), then the states as above; each output line then ends with the W
// words of enabled bits mbg_eval_p returns.
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
#ifdef GEN_PARAM
	int np = 0;
	std::cin >> np;
	std::vector<float> rate(np + 1), lr(np + 1);
	std::vector<uint32_t> up(MBG_W), dn(MBG_W);
	for (int k = 0; k < np; k++)
	{
		uint32_t b;
		std::cin >> std::hex >> b;
		rate[k] = mbg_u2f(b);
		lr[k] = mbg_rlog2r(rate[k]);
	}
	for (int w = 0; w < MBG_W; w++)
		std::cin >> std::hex >> up[w];
	for (int w = 0; w < MBG_W; w++)
		std::cin >> std::hex >> dn[w];
	const mbg_param_view pv { rate.data(), lr.data(), up.data(), dn.data() };
	mbg_eval_p_setup(&pv);
#else
	mbg_eval_setup();
#endif
	while (true)
	{
		for (int w = 0; w < MBG_W; w++)
			if (!(std::cin >> std::hex >> s[w]))
				return 0;
		std::cin >> std::hex >> u2b;
		mbg_eval_t e;
#ifdef GEN_PARAM
		std::vector<uint32_t> en(MBG_W);
		mbg_eval_p(s.data(), mbg_u2f(u2b), &pv, &e, en.data());
#else
		mbg_eval(s.data(), mbg_u2f(u2b), &e);
#endif
		const uint64_t obs = mbg_obs_index(s.data());
		std::printf("%08x %08x %d %llx", mbg_f2u(e.total), mbg_f2u(e.H), e.flip, (unsigned long long)obs);
#ifdef GEN_PARAM
		// the enabled words, after the oracle's columns
		for (int w = 0; w < MBG_W; w++)
			std::printf(" %x", en[w]);
#endif
		std::printf("\n");
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
