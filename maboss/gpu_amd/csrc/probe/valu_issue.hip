// VALU issue-rate probe for the roofline of the simulate kernels.
//
// Each thread runs 8 independent dependency chains of one VALU instruction
// (inline asm, so the instruction is exactly the named one), 4 x 8
// instructions per loop iteration; the grid puts 8 waves on every SIMD (and
// then 1 and 2 waves per SIMD for the latency-bound case).  Prints one JSON
// line per case with lane-instructions per second: on gfx950 the 2-operand
// 32-bit integer ops and v_fma_f32 issue about twice as fast as v_bfe_u32,
// v_and_or_b32 and v_mul_hi_u32, which is why the roofline's VALU peak is
// the published FP32 vector rate (one lane-op per lane every 2 cycles).
//
// Every thread writes one word of its own: out has exactly grid x 256 words.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>

#define CHK(x)                                                                    \
	do                                                                            \
	{                                                                             \
		hipError_t e_ = (x);                                                      \
		if (e_ != hipSuccess)                                                     \
		{                                                                         \
			std::fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_));          \
			return 1;                                                             \
		}                                                                         \
	} while (0)

#define CHAINS(I) I(a0, a1) I(a1, a2) I(a2, a3) I(a3, a4) I(a4, a5) I(a5, a6) I(a6, a7) I(a7, a0)
#define OP_XOR(x, y) asm volatile("v_xor_b32 %0, %0, %1" : "+v"(x) : "v"(y));
#define OP_ADD(x, y) asm volatile("v_add_u32 %0, %0, %1" : "+v"(x) : "v"(y));
#define OP_FMA(x, y) asm volatile("v_fma_f32 %0, %0, %1, %1" : "+v"(x) : "v"(y));
#define OP_BFE(x, y) asm volatile("v_bfe_u32 %0, %0, %1, 5" : "+v"(x) : "v"(y));
#define OP_ANDOR(x, y) asm volatile("v_and_or_b32 %0, %0, %1, %1" : "+v"(x) : "v"(y));
#define OP_MULHI(x, y) asm volatile("v_mul_hi_u32 %0, %0, %1" : "+v"(x) : "v"(y));

template <int OP>
__global__ void __launch_bounds__(256) valu_chains(uint32_t* out, uint32_t iters)
{
	uint32_t a0 = threadIdx.x, a1 = a0 * 3u + 1u, a2 = a0 * 5u + 2u, a3 = a0 * 7u + 3u;
	uint32_t a4 = a0 * 11u + 4u, a5 = a0 * 13u + 5u, a6 = a0 * 17u + 6u, a7 = a0 * 19u + 7u;
	for (uint32_t i = 0; i < iters; ++i)
	{
#pragma unroll
		for (int r = 0; r < 4; ++r)
		{
			if constexpr (OP == 0) { CHAINS(OP_XOR) }
			if constexpr (OP == 1) { CHAINS(OP_ADD) }
			if constexpr (OP == 2) { CHAINS(OP_FMA) }
			if constexpr (OP == 3) { CHAINS(OP_BFE) }
			if constexpr (OP == 4) { CHAINS(OP_ANDOR) }
			if constexpr (OP == 5) { CHAINS(OP_MULHI) }
		}
	}
	out[blockIdx.x * 256 + threadIdx.x] = a0 ^ a1 ^ a2 ^ a3 ^ a4 ^ a5 ^ a6 ^ a7;
}

template <int OP>
static int run(const char* name, uint32_t* out, int blocks, int waves_per_simd, uint32_t iters)
{
	hipEvent_t e0, e1;
	CHK(hipEventCreate(&e0));
	CHK(hipEventCreate(&e1));
	valu_chains<OP><<<blocks, 256>>>(out, iters);
	CHK(hipGetLastError());
	CHK(hipDeviceSynchronize());
	const int reps = 5;
	CHK(hipEventRecord(e0));
	for (int r = 0; r < reps; ++r)
		valu_chains<OP><<<blocks, 256>>>(out, iters);
	CHK(hipEventRecord(e1));
	CHK(hipEventSynchronize(e1));
	float ms = 0.0f;
	CHK(hipEventElapsedTime(&ms, e0, e1));
	CHK(hipEventDestroy(e0));
	CHK(hipEventDestroy(e1));
	const double lane_instr = (double)reps * blocks * 256.0 * iters * 32.0;
	std::printf("{\"op\": \"%s\", \"waves_per_simd\": %d, \"ms_per_launch\": %.4f, \"lane_instr_per_s\": %.4e}\n", name,
				waves_per_simd, ms / reps, lane_instr / (ms * 1e-3));
	return 0;
}

int main()
{
	hipDeviceProp_t prop;
	CHK(hipGetDeviceProperties(&prop, 0));
	std::printf("{\"device\": \"%s\", \"cus\": %d, \"clock_khz\": %d}\n", prop.gcnArchName, prop.multiProcessorCount,
				prop.clockRate);
	const int cus = prop.multiProcessorCount;
	// 256-thread blocks are 4 waves, one per SIMD: k blocks per CU = k waves per SIMD
	const int max_blocks = cus * 8;
	uint32_t* out = nullptr;
	CHK(hipMalloc(&out, (size_t)max_blocks * 256 * sizeof(uint32_t)));
	const uint32_t iters = 20000;
	int rc = run<0>("v_xor_b32", out, cus * 8, 8, iters) || run<1>("v_add_u32", out, cus * 8, 8, iters)
			 || run<2>("v_fma_f32", out, cus * 8, 8, iters) || run<3>("v_bfe_u32", out, cus * 8, 8, iters)
			 || run<4>("v_and_or_b32", out, cus * 8, 8, iters) || run<5>("v_mul_hi_u32", out, cus * 8, 8, iters)
			 || run<0>("v_xor_b32", out, cus, 1, iters) || run<0>("v_xor_b32", out, cus * 2, 2, iters);
	CHK(hipFree(out));
	return rc;
}
