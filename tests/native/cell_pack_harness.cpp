// Round trip of the packed window-cell format through the host decoder.
//
// Window cells reach the host as 12-byte records: record i is the 96-bit
// integer key | field << kbits (three 32-bit words, least significant first).
// With esc_min = 0 (kbits <= 31) the field is the value; otherwise the field
// has 96 - kbits bits and esc_min = 2^(95 - kbits): values below esc_min are
// the field, larger ones sit in a side list at position p with field
// esc_min + p.  This harness encodes random cells into that format with
// 128-bit integer arithmetic (independent of the device kernel's split
// shifts), decodes them with the engine's cell_view (k, v, lower_bound,
// slice) and prints "ok <kbits> <cells> <escaped>" per key width, or fails.
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "engine.h"

namespace mbg
{
// cell_view makes no HIP call; the library's definitions of these are not
// linked here (hostmem.cpp's other members refer to them)
void hip_check(hipError_t, const char*) {}
void device_buffer::release() {}
} // namespace mbg

using mbg::cell_view;

static int fail(const char* what, unsigned kbits, uint64_t i, uint64_t got, uint64_t want)
{
	std::fprintf(stderr, "%s: kbits %u cell %llu: got %llu want %llu\n", what, kbits, (unsigned long long)i,
				 (unsigned long long)got, (unsigned long long)want);
	return 1;
}

static int check(unsigned kbits, uint64_t n, std::mt19937_64& rng)
{
	const uint64_t esc_min = kbits <= 31 ? 0ull : 1ull << (95 - kbits);
	const uint64_t kmask = kbits >= 64 ? ~0ull : (1ull << kbits) - 1ull;
	// sorted distinct keys below 2^kbits (n of them, or every key when fewer exist)
	std::vector<uint64_t> keys;
	if (kbits <= 12)
	{
		for (uint64_t k = 0; k <= kmask && keys.size() < n; k += 1 + (rng() % 3))
			keys.push_back(k);
	}
	else
	{
		uint64_t k = rng() % 97;
		const uint64_t step = (kmask >> 1) / n;
		for (uint64_t i = 0; i < n && k <= kmask; i++, k += 1 + rng() % std::max<uint64_t>(step, 1))
			keys.push_back(k);
		keys.back() = kmask; // the largest key the width holds
	}
	n = keys.size();
	// values: the edges of the field and of the escape, and random widths
	std::vector<uint64_t> vals(n);
	const uint64_t edges[] = { 1, 2, esc_min ? esc_min - 1 : 3, esc_min, esc_min + 1, ~0ull, 1ull << 63, 4294967296ull };
	for (uint64_t i = 0; i < n; i++)
		vals[i] = i < 8 ? edges[i] : (rng() >> (rng() % 64)) | 1ull;
	std::vector<uint32_t> rec(3 * n);
	std::vector<uint64_t> side;
	for (uint64_t i = 0; i < n; i++)
	{
		uint64_t f = vals[i];
		if (esc_min && vals[i] >= esc_min)
		{
			f = esc_min + side.size();
			side.push_back(vals[i]);
		}
		const unsigned __int128 r = (unsigned __int128)keys[i] | ((unsigned __int128)f << kbits);
		if (r >> 96)
			return fail("encoding overflows 96 bits", kbits, i, (uint64_t)(r >> 96), 0);
		rec[3 * i] = (uint32_t)r;
		rec[3 * i + 1] = (uint32_t)(r >> 32);
		rec[3 * i + 2] = (uint32_t)(r >> 64);
	}
	cell_view c;
	c.rec = rec.data();
	c.kbits = kbits;
	c.esc_min = esc_min;
	c.xval = side.data();
	c.nx = side.size();
	c.n = n;
	for (uint64_t i = 0; i < n; i++)
	{
		if (c.k(i) != keys[i])
			return fail("key", kbits, i, c.k(i), keys[i]);
		if (c.v(i) != vals[i])
			return fail("value", kbits, i, c.v(i), vals[i]);
	}
	for (int t = 0; t < 200; t++)
	{
		const uint64_t i = rng() % n;
		if (c.lower_bound(keys[i]) != i)
			return fail("lower_bound", kbits, i, c.lower_bound(keys[i]), i);
		if (keys[i] && (i == 0 || keys[i - 1] != keys[i] - 1) && c.lower_bound(keys[i] - 1) != i)
			return fail("lower_bound below", kbits, i, c.lower_bound(keys[i] - 1), i);
	}
	// a slice keeps decoding against the whole record array and side list
	const uint64_t lo = n / 3, hi = n - n / 5;
	const cell_view s = c.slice(lo, hi, keys[lo]);
	for (uint64_t i = 0; i < s.n; i++)
	{
		if (s.k(i) != keys[lo + i] - keys[lo])
			return fail("slice key", kbits, i, s.k(i), keys[lo + i] - keys[lo]);
		if (s.v(i) != vals[lo + i])
			return fail("slice value", kbits, i, s.v(i), vals[lo + i]);
	}
	// a field past the side list is an error, not a read out of bounds
	if (esc_min)
	{
		cell_view bad = c;
		bad.nx = side.size() ? side.size() - 1 : 0;
		bool threw = false;
		try
		{
			for (uint64_t i = 0; i < n; i++)
				(void)bad.v(i);
		}
		catch (const mbg::error& e)
		{
			threw = e.code == mbg::MBG_E_INTERNAL;
		}
		if (!threw)
			return fail("short side list not detected", kbits, 0, 0, 1);
	}
	std::printf("ok %u %llu %llu\n", kbits, (unsigned long long)n, (unsigned long long)side.size());
	return 0;
}

int main()
{
	std::mt19937_64 rng(12345);
	const unsigned widths[] = { 1, 7, 12, 24, 31, 32, 33, 45, 57, 59, 60 };
	for (unsigned kb : widths)
		if (check(kb, 5000, rng))
			return 1;
	return 0;
}
