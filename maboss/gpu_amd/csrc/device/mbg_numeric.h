// Numeric primitives shared by the host library and the HIP kernels.
//
// Everything here is plain IEEE-754 float / uint32 arithmetic that gives the
// same bits on the x86 host and on gfx950, provided both sides are compiled
// with -ffp-contract=off and without fast-math.  The random stream replaces
// the reference's cuRAND XORWOW state (src/jit_kernels/simulation.cu:12,31,53,
// 129), which cannot be reproduced; it is a counter-based Philox4x32-10
// (Salmon, Moraes, Dror, Shaw, "Parallel Random Numbers: As Easy as 1, 2, 3",
// SC'11), so every draw is a pure function of (seed, variant, trajectory id,
// draw index) and no generator state is stored or shuffled between batches.
//
// Stream layout of one trajectory (key = seed, counter = {block, id_lo,
// id_hi, variant}; each block yields four 32-bit words):
//   * initial state: node i uses word (i & 3) of block (i >> 2);
//   * steps start at block B0 = ceil(N / 4);
//   * continuous time, step k: u1 (dwell) = word 2(k&1), u2 (flip) =
//     word 2(k&1)+1 of block B0 + (k >> 1);
//   * discrete time, step k: u2 (flip) = word (k & 3) of block B0 + (k >> 2).
#pragma once

#include <stdint.h>

#if defined(__HIPCC__) || defined(__HIPCC_RTC__)
#define MBG_HD __host__ __device__ __forceinline__
#else
#define MBG_HD static inline
#endif

#define MBG_PHILOX_M0 0xD2511F53u
#define MBG_PHILOX_M1 0xCD9E8D57u
#define MBG_PHILOX_W0 0x9E3779B9u
#define MBG_PHILOX_W1 0xBB67AE85u

struct mbg_u32x4
{
	uint32_t x, y, z, w;
};

// Philox4x32 with 10 rounds.  Known-answer vectors from the Random123
// distribution are checked in the test suite.
MBG_HD mbg_u32x4 mbg_philox4x32_10(mbg_u32x4 c, uint32_t k0, uint32_t k1)
{
#if defined(__HIP_DEVICE_COMPILE__)
#pragma unroll
#endif
	for (int r = 0; r < 10; ++r)
	{
		// one 32 x 32 -> 64-bit product per multiplier (a single v_mad_u64_u32)
		const uint64_t p0 = (uint64_t)MBG_PHILOX_M0 * c.x, p1 = (uint64_t)MBG_PHILOX_M1 * c.z;
		const uint32_t lo0 = (uint32_t)p0, hi0 = (uint32_t)(p0 >> 32);
		const uint32_t lo1 = (uint32_t)p1, hi1 = (uint32_t)(p1 >> 32);
		mbg_u32x4 n;
		n.x = hi1 ^ c.y ^ k0;
		n.y = lo1;
		n.z = hi0 ^ c.w ^ k1;
		n.w = lo0;
		c = n;
		k0 += MBG_PHILOX_W0;
		k1 += MBG_PHILOX_W1;
	}
	return c;
}

// Uniform on (0, 1] with 24-bit resolution: ((x >> 8) + 1) * 2^-24.  Exact in
// float, never 0, and 1 is reachable, like the reference's curand_uniform.
MBG_HD float mbg_u01(uint32_t x) { return (float)((x >> 8) + 1u) * 0x1p-24f; }

MBG_HD uint32_t mbg_f2u(float f)
{
	union
	{
		float f;
		uint32_t u;
	} v;
	v.f = f;
	return v.u;
}

MBG_HD float mbg_u2f(uint32_t u)
{
	union
	{
		float f;
		uint32_t u;
	} v;
	v.u = u;
	return v.f;
}

// log2 for positive finite floats (subnormals included); log2(0) = -inf.
// Method: x = 2^e * m with m in [sqrt(1/2), sqrt(2)), f = m - 1 (exact by
// Sterbenz), log2(1 + f) = f * P(f) with P a degree-9 polynomial obtained by a
// Chebyshev least-squares fit of log2(1+f)/f on that interval, evaluated by
// Horner's rule in float.  Max error about 1 ulp of the result.
MBG_HD float mbg_log2f(float x)
{
	uint32_t b = mbg_f2u(x);
	if (b >= 0x7f800000u)
		return x; // +inf, NaN or negative: not produced by the simulator
	int e = 0;
	if (b < 0x00800000u)
	{
		if (b == 0u)
			return -__builtin_inff();
		b = mbg_f2u(x * 0x1p23f);
		e = -23;
	}
	e += (int)(b >> 23) - 127;
	float m = mbg_u2f((b & 0x007fffffu) | 0x3f800000u);
	const bool hi = m > 1.41421354f;
	m = hi ? m * 0.5f : m;
	e += hi ? 1 : 0;
	const float f = m - 1.0f;
	float p = -0x1.c362bcp-4f;
	p = p * f + 0x1.7d9132p-3f;
	p = p * f + -0x1.87381cp-3f;
	p = p * f + 0x1.a2f86p-3f;
	p = p * f + -0x1.eabd64p-3f;
	p = p * f + 0x1.277e9ap-2f;
	p = p * f + -0x1.715a76p-2f;
	p = p * f + 0x1.ec7094p-2f;
	p = p * f + -0x1.71547p-1f;
	p = p * f + 0x1.715476p+0f;
	return (float)e + f * p;
}

// Natural log built on mbg_log2f: ln(x) = log2(x) * ln(2).
MBG_HD float mbg_logf(float x) { return mbg_log2f(x) * 0.693147182f; }

// r * log2(r), with 0 * log2(0) = 0.  Used by the transition entropy.
MBG_HD float mbg_rlog2r(float r) { return r == 0.0f ? 0.0f : r * mbg_log2f(r); }
