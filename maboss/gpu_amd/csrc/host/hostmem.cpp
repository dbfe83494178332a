// Host memory of the results: the pinned block pool, sorted cell views, and
// the copy slots a program's runs leave the device through.
#include <algorithm>

#include "engine.h"

namespace mbg
{

namespace
{
// blocks are rounded up to this size, so runs of similar size share them
constexpr size_t k_pin_granule = 2u << 20;
// a free block is taken for a request of at least this fraction of its size
constexpr size_t k_pin_waste = 4;
// the pool keeps at most this much pinned memory unused
constexpr size_t k_pin_retain = 32ull << 30;
} // namespace

pinned_pool::~pinned_pool()
{
	for (auto& b : free_)
		(void)hipHostFree(b.second);
}

std::shared_ptr<char> pinned_pool::get(size_t bytes)
{
	const size_t want = std::max<size_t>((bytes + k_pin_granule - 1) / k_pin_granule, 1) * k_pin_granule;
	char* p = nullptr;
	size_t have = 0;
	{
		std::lock_guard<std::mutex> lk(mu_);
		size_t best = free_.size();
		for (size_t i = 0; i < free_.size(); i++)
			if (free_[i].first >= want && free_[i].first <= k_pin_waste * want
				&& (best == free_.size() || free_[i].first < free_[best].first))
				best = i;
		if (best < free_.size())
		{
			have = free_[best].first;
			p = free_[best].second;
			retained_ -= have;
			free_.erase(free_.begin() + (std::ptrdiff_t)best);
		}
	}
	if (!p)
	{
		void* q = nullptr;
		hip_check(hipHostMalloc(&q, want, hipHostMallocDefault), "hipHostMalloc");
		p = (char*)q;
		have = want;
	}
	std::shared_ptr<pinned_pool> self = shared_from_this();
	return std::shared_ptr<char>(p, [self, have](char* q) { self->put(q, have); });
}

void pinned_pool::put(char* p, size_t bytes)
{
	{
		std::lock_guard<std::mutex> lk(mu_);
		if (retained_ + bytes <= k_pin_retain)
		{
			free_.emplace_back(bytes, p);
			retained_ += bytes;
			return;
		}
	}
	(void)hipHostFree(p);
}

uint64_t cell_view::lower_bound(uint64_t kk) const
{
	uint64_t lo = 0, hi = n;
	while (lo < hi)
	{
		const uint64_t mid = lo + (hi - lo) / 2;
		if (k(mid) < kk)
			lo = mid + 1;
		else
			hi = mid;
	}
	return lo;
}

uint64_t cell_view::packed_val(uint64_t r) const
{
	uint64_t f = (packed_lo(r) >> kbits) | ((uint64_t)rec[3 * r + 2] << (64 - kbits));
	if (!esc_min)
		return f;
	const unsigned vb = 96u - kbits;
	if (vb < 64)
		f &= (1ull << vb) - 1ull;
	if (f < esc_min)
		return f;
	if (f - esc_min >= nx)
		throw error(MBG_E_INTERNAL, "packed window cell points past the side list");
	return xval[f - esc_min];
}

cell_view cell_view::slice(uint64_t lo, uint64_t hi, uint64_t s) const
{
	cell_view c = *this;
	if (rec)
		c.first = first + lo;
	else
	{
		c.key = key + lo;
		c.val = val + lo;
	}
	c.n = hi - lo;
	c.sub = s;
	return c;
}

cell_view cell_view::from(std::vector<uint64_t> keys, std::vector<uint64_t> vals)
{
	auto own = std::make_shared<std::pair<std::vector<uint64_t>, std::vector<uint64_t>>>(std::move(keys), std::move(vals));
	cell_view c;
	c.key = own->first.data();
	c.val = own->second.data();
	c.n = own->first.size();
	c.hold = own;
	return c;
}

out_slot::~out_slot()
{
	if (copied)
		(void)hipEventDestroy(copied);
}

void program::order_after_last_run(hipStream_t stream)
{
	if (run_done)
		hip_check(hipStreamWaitEvent(stream, run_done, 0), "hipStreamWaitEvent");
}

void program::mark_run_done(hipStream_t stream)
{
	if (!run_done)
		hip_check(hipEventCreateWithFlags(&run_done, hipEventDisableTiming), "hipEventCreateWithFlags");
	hip_check(hipEventRecord(run_done, stream), "hipEventRecord");
}

void program::drain_copies()
{
	for (auto& s : slots)
		if (s.pending && s.copied)
		{
			(void)hipEventSynchronize(s.copied);
			s.pending = false;
		}
}

} // namespace mbg
