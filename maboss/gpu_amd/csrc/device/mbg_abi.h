// Device-side ABI between the per-model generated code (emitted by the host
// generator) and the hand-written lane kernels: hipRTC compiles both in one
// translation unit, so every call below is inlined into the kernel.
//
// The generated functions replace the reference's generated
// compute_transition_rates / compute_transition_entropy /
// get_non_internal_index (src/generator.cpp:124-175, :207-235) and the
// flip selection of src/jit_kernels/simulation.cu:9-21.
#pragma once

#include "mbg_numeric.h"

// Result of evaluating every node's rate in one state.
struct mbg_eval_t
{
	float total; // sum of all node rates, sequential in node order
	float H;     // transition entropy of the observed (non-internal) nodes
	int flip;    // node to flip for the given u2 (first i with prefix_i >= u2*total)
	int pad;
};

// Tables a launch only reads (rates, masks, variant descriptors) are read
// through the constant address space on the device: a load from a
// wave-uniform address there is a scalar load whose value lands in an SGPR,
// where a plain global pointer gives a vector load per lane (the compiler
// cannot prove the table unchanged across the kernel's atomics).  The
// pointers have the same 64-bit layout on the host.
#if defined(__HIP_DEVICE_COMPILE__)
#define MBG_CONST __attribute__((address_space(4)))
#else
#define MBG_CONST
#endif

// Address space of the rate tables a view points to: constant, so the
// wave-uniform rates are scalar loads.  (Vector loads, which the kernels
// once used up to 40 nodes, hoist every rate into a VGPR: on the 30-node
// gradient workload the run-time-rates simulate took 2.30 ms with them
// against 2.21 ms with scalar loads, both at their best occupancy.)
#define MBG_RATE_AS MBG_CONST

// Per-variant runtime tables for the batched and differentiated forms.
struct mbg_param_view
{
	const MBG_RATE_AS float* rate;      // [n_params] value of each $variable used by a rate
	const MBG_RATE_AS float* rate_lr;   // [n_params] rate * log2(rate), precomputed
	const MBG_RATE_AS uint32_t* up_en;  // [W] bit i clear: node i can never switch on
	const MBG_RATE_AS uint32_t* dn_en;  // [W] bit i clear: node i can never switch off
};

// Hides a value from the optimiser (no instruction is emitted), so that
// expressions computed from it are not merged with identical ones computed
// from the original value.
#if defined(__HIP_DEVICE_COMPILE__)
#define MBG_OPAQUE(x) asm volatile("" : "+v"(x))
#else
#define MBG_OPAQUE(x) asm volatile("" : "+r"(x))
#endif

// Hides that a value is a 0 / all-ones mask, so that a select written as
// (u ^ ((u ^ d) & m)) & e stays two bitwise instructions on the mask instead
// of becoming a compare and a conditional move with both constants loaded.
// Unlike MBG_OPAQUE it does not order the code around it.
#if defined(__HIP_DEVICE_COMPILE__)
#define MBG_MASK(x) asm("" : "+v"(x))
#else
#define MBG_MASK(x) ((void)0)
#endif

// Any bitwise function of three words, given by its truth table t: bit
// (4a + 2b + c) of t is the result for input bits a, b, c (gfx950's
// v_bitop3_b32).  The host form sums the table's minterms.
#if defined(__HIP_DEVICE_COMPILE__)
#define MBG_BITOP3(a, b, c, t) __builtin_amdgcn_bitop3_b32((a), (b), (c), (t))
#else
static inline uint32_t mbg_bitop3_host(uint32_t a, uint32_t b, uint32_t c, unsigned t)
{
	uint32_t r = 0u;
	for (unsigned i = 0; i < 8; i++)
		if (t >> i & 1u)
			r |= (i & 4u ? a : ~a) & (i & 2u ? b : ~b) & (i & 1u ? c : ~c);
	return r;
}
#define MBG_BITOP3(a, b, c, t) mbg_bitop3_host((a), (b), (c), (t))
#endif

// Pins running sums at a point of the instruction stream (no instruction is
// emitted): the generated rate code places one after each node, so each
// node's terms are added before the next node's are computed, rather than
// all of them being kept in registers for sums the optimiser sinks to the
// end.
#if defined(__HIP_DEVICE_COMPILE__)
#define MBG_PIN2(a, b) asm volatile("" : "+v"(a), "+v"(b))
#define MBG_PIN4(a, b, c, d) asm volatile("" : "+v"(a), "+v"(b), "+v"(c), "+v"(d))
#else
#define MBG_PIN2(a, b) ((void)0)
#define MBG_PIN4(a, b, c, d) ((void)0)
#endif

// Makes one lane's LDS stores visible to the other lanes of its wave.
#if defined(__HIP_DEVICE_COMPILE__)
#define MBG_WAVE_SYNC()                                                                                                \
	do                                                                                                                 \
	{                                                                                                                  \
		__builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");                                                         \
		__builtin_amdgcn_wave_barrier();                                                                               \
		__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");                                                         \
	} while (0)
#else
#define MBG_WAVE_SYNC() ((void)0)
#endif

#if defined(__HIPCC__) || defined(__HIPCC_RTC__)
// Block-shared tables of mbg_eval: every thread of the block calls it before
// the kernel's first barrier.
extern "C" __device__ void mbg_eval_setup();
// Rates with the model's $variables baked in as exact float constants.
extern "C" __device__ void mbg_eval(const uint32_t* __restrict__ s, float u2, mbg_eval_t* __restrict__ out);
// Per-wave tables of mbg_eval_p for the parameter view pv: every lane of the
// wave calls it (converged) before the wave's first mbg_eval_p with pv.
extern "C" __device__ void mbg_eval_p_setup(const mbg_param_view* __restrict__ pv);
// Rates read from runtime tables (batched variants, gradients).
// en (may be null) receives the W words of enabled bits: bit i set when node
// i's logic differs from its value and up_en / dn_en allow that direction.
extern "C" __device__ void mbg_eval_p(const uint32_t* __restrict__ s, float u2, const mbg_param_view* __restrict__ pv,
									  mbg_eval_t* __restrict__ out, uint32_t* __restrict__ en);
// Parallel bit-extract of the observed nodes, in node order (bit j of the
// result is the j-th non-internal node).
extern "C" __device__ uint64_t mbg_obs_index(const uint32_t* __restrict__ s);
// The bit of node i in the observed index (0 for an internal node): flipping
// node i toggles it.
extern "C" __device__ uint64_t mbg_obs_flip(int i);
#endif
