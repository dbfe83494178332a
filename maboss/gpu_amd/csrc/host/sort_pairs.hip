// Device-side ordering of the statistics before they leave the device.
//
//   * sort_reduce_pairs: the sorted window mode.  The simulate kernel appends
//     (cell key, fixed-point dwell) pairs for the window histogram instead of
//     inserting them into a hash table; this step sorts the pairs by key
//     (rocPRIM radix sort over the key's significant bits) and sums the values
//     of equal keys (rocPRIM reduce-by-key).  The sums are unsigned integers,
//     so the result does not depend on the order in which the pairs were
//     appended.
//   * sort_pairs: the same radix sort for cells compacted out of a hash table
//     (keys already distinct).
//   * compact_dense: the nonzero cells of a dense histogram in index order
//     (rocPRIM select, order preserving), so the host receives only occupied
//     cells, already sorted.
// The reference has no counterpart: its window histograms are dense float
// arrays indexed by at most 20 observed bits, copied to the host whole
// (src/main.cpp:142; src/statistics/window_average_small.cpp:91-107).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>

#include <rocprim/rocprim.hpp>

#include "engine.h"

namespace mbg
{

void sort_reduce_pairs(device_buffer& temp, unsigned long long* keys, unsigned long long* vals,
					   unsigned long long* keys_alt, unsigned long long* vals_alt, uint64_t n, int key_bits,
					   unsigned long long* unique_count, hipStream_t stream, unsigned long long** out_keys,
					   unsigned long long** out_vals)
{
	rocprim::double_buffer<unsigned long long> kb(keys, keys_alt), vb(vals, vals_alt);
	size_t sort_bytes = 0, red_bytes = 0;
	hip_check(rocprim::radix_sort_pairs(nullptr, sort_bytes, kb, vb, (size_t)n, 0u, (unsigned)key_bits, stream),
			  "rocprim::radix_sort_pairs (size)");
	// the reduction reads the sorted pair of buffers and writes the other one
	hip_check(rocprim::reduce_by_key(nullptr, red_bytes, keys, vals, (size_t)n, keys_alt, vals_alt, unique_count,
									 rocprim::plus<unsigned long long>(), rocprim::equal_to<unsigned long long>(),
									 stream),
			  "rocprim::reduce_by_key (size)");
	void* t = temp.ensure(std::max<size_t>(std::max(sort_bytes, red_bytes), 256));
	size_t tb = temp.bytes;
	hip_check(rocprim::radix_sort_pairs(t, tb, kb, vb, (size_t)n, 0u, (unsigned)key_bits, stream),
			  "rocprim::radix_sort_pairs");
	unsigned long long* sk = kb.current();
	unsigned long long* sv = vb.current();
	unsigned long long* ok = sk == keys ? keys_alt : keys;
	unsigned long long* ov = sv == vals ? vals_alt : vals;
	tb = temp.bytes;
	hip_check(rocprim::reduce_by_key(t, tb, sk, sv, (size_t)n, ok, ov, unique_count,
									 rocprim::plus<unsigned long long>(), rocprim::equal_to<unsigned long long>(),
									 stream),
			  "rocprim::reduce_by_key");
	*out_keys = ok;
	*out_vals = ov;
}

void sort_pairs(device_buffer& temp, unsigned long long* keys, unsigned long long* vals, unsigned long long* keys_alt,
				unsigned long long* vals_alt, uint64_t n, int key_bits, hipStream_t stream,
				unsigned long long** out_keys, unsigned long long** out_vals)
{
	rocprim::double_buffer<unsigned long long> kb(keys, keys_alt), vb(vals, vals_alt);
	size_t sort_bytes = 0;
	hip_check(rocprim::radix_sort_pairs(nullptr, sort_bytes, kb, vb, (size_t)n, 0u, (unsigned)key_bits, stream),
			  "rocprim::radix_sort_pairs (size)");
	void* t = temp.ensure(std::max<size_t>(sort_bytes, 256));
	size_t tb = temp.bytes;
	hip_check(rocprim::radix_sort_pairs(t, tb, kb, vb, (size_t)n, 0u, (unsigned)key_bits, stream),
			  "rocprim::radix_sort_pairs");
	*out_keys = kb.current();
	*out_vals = vb.current();
}

namespace
{
struct nonzero
{
	__host__ __device__ bool operator()(unsigned long long v) const { return v != 0ull; }
};
} // namespace

void compact_dense(device_buffer& temp, const unsigned long long* dense, uint64_t n, unsigned long long* out_keys,
				   unsigned long long* out_vals, unsigned long long* count, hipStream_t stream)
{
	auto flags = rocprim::make_transform_iterator(dense, nonzero {});
	rocprim::counting_iterator<unsigned long long> idx(0ull);
	size_t b1 = 0, b2 = 0;
	hip_check(rocprim::select(nullptr, b1, idx, flags, out_keys, count, (size_t)n, stream), "rocprim::select (size)");
	hip_check(rocprim::select(nullptr, b2, dense, flags, out_vals, count, (size_t)n, stream), "rocprim::select (size)");
	void* t = temp.ensure(std::max<size_t>(std::max(b1, b2), 256));
	size_t tb = temp.bytes;
	hip_check(rocprim::select(t, tb, idx, flags, out_keys, count, (size_t)n, stream), "rocprim::select");
	tb = temp.bytes;
	hip_check(rocprim::select(t, tb, dense, flags, out_vals, count, (size_t)n, stream), "rocprim::select");
}

} // namespace mbg
