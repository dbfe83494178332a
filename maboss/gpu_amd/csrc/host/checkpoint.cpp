// Checkpoint and resume at trajectory-id granularity.  The reference's only
// resumable state is the epoch boundary of its runner, which keeps nothing
// across processes (src/simulation_runner.cu:123-161); here the state is a
// result: exact integer statistics plus the run's identity, saved to a
// file.  A run is resumed by simulating the ids after the saved range and
// adding the two results, which gives the single run's statistics bit for
// bit (the sums are integers and trajectory id i always draws the same
// stream).
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sys/stat.h>
#include <unistd.h>

#include "engine.h"
#include "timer.h"

namespace mbg
{

namespace
{

constexpr char k_magic[8] = { 'M', 'B', 'G', 'C', 'K', 'P', 'T', '1' };

struct fnv_hash
{
	uint64_t h = 1469598103934665603ull;
	void bytes(const void* d, size_t n)
	{
		const unsigned char* p = (const unsigned char*)d;
		for (size_t i = 0; i < n; i++)
		{
			h ^= p[i];
			h *= 1099511628211ull;
		}
	}
	template <typename T>
	void pod(const T& v)
	{
		bytes(&v, sizeof(v));
	}
};

// A flat buffer written or read field by field; the file ends with the
// FNV-1a hash of everything before it, so a truncated or damaged file is
// refused.
struct writer
{
	std::vector<char> buf;
	void raw(const void* d, size_t n) { buf.insert(buf.end(), (const char*)d, (const char*)d + n); }
	template <typename T>
	void pod(const T& v)
	{
		raw(&v, sizeof(v));
	}
	template <typename T>
	void vec(const T* d, uint64_t n)
	{
		pod(n);
		raw(d, n * sizeof(T));
	}
};

struct reader
{
	const std::vector<char>& buf;
	size_t at = 0;
	void raw(void* d, size_t n)
	{
		if (n > buf.size() - at)
			throw error(MBG_E_IO, "checkpoint file is truncated");
		std::memcpy(d, buf.data() + at, n);
		at += n;
	}
	template <typename T>
	T pod()
	{
		T v;
		raw(&v, sizeof(v));
		return v;
	}
	template <typename T>
	std::vector<T> vec(uint64_t limit)
	{
		const uint64_t n = pod<uint64_t>();
		if (n > limit || n * sizeof(T) > buf.size() - at)
			throw error(MBG_E_IO, "checkpoint file is damaged");
		std::vector<T> v(n);
		raw(v.data(), n * sizeof(T));
		return v;
	}
};

void check_same_run(const result& a, const result& b)
{
	if (!a.fingerprint || !b.fingerprint)
		throw error(MBG_E_ARG, "results of batched launches are not merged");
	if (a.fingerprint != b.fingerprint || a.seed != b.seed || a.variant != b.variant)
		throw error(MBG_E_ARG, "the results come from different models, rates, seeds or stream tags");
	if (a.windows != b.windows || a.n_obs != b.n_obs || a.n_words != b.n_words || a.discrete != b.discrete)
		throw error(MBG_E_ARG, "the results have different layouts");
	if (a.user_norm != b.user_norm || (a.user_norm && a.norm != b.norm))
		throw error(MBG_E_ARG, "the results have different normalizations");
	if (a.has_moments != b.has_moments)
		throw error(MBG_E_ARG, "one result has second moments and the other has none");
	if (a.kernel != b.kernel)
		throw error(MBG_E_ARG, "the results come from different kernels (their trajectories differ)");
}

// (key, value) cells of two key-sorted views, equal keys summed
cell_view merge_cells(const cell_view& a, const cell_view& b)
{
	std::vector<uint64_t> k, v;
	k.reserve(a.n + b.n);
	v.reserve(a.n + b.n);
	uint64_t i = 0, j = 0;
	while (i < a.n || j < b.n)
	{
		if (j >= b.n || (i < a.n && a.k(i) < b.k(j)))
		{
			k.push_back(a.k(i));
			v.push_back(a.v(i++));
		}
		else if (i >= a.n || b.k(j) < a.k(i))
		{
			k.push_back(b.k(j));
			v.push_back(b.v(j++));
		}
		else
		{
			k.push_back(a.k(i));
			v.push_back(a.v(i++) + b.v(j++));
		}
	}
	return cell_view::from(std::move(k), std::move(v));
}

} // namespace

uint64_t run_fingerprint(const program& p, const sim_params& sp)
{
	const model& m = *p.mdl;
	fnv_hash f;
	f.bytes(p.gen.source.data(), p.gen.source.size());
	f.pod((int)p.opts.entropy);
	f.pod(m.max_time());
	f.pod(m.time_tick());
	f.pod(m.windows());
	f.pod(m.n_observed());
	f.pod(m.discrete_time());
	for (const auto& n : m.nodes)
	{
		f.bytes(n.name.data(), n.name.size());
		f.pod(n.istate);
	}
	for (float r : sp.rates)
		f.pod(r);
	f.pod(sp.max_steps);
	return f.h | 1ull; // never 0 (0 marks a result that is not merged)
}

void save_result(const result& r, const std::string& path)
{
	r.wait();
	writer w;
	w.raw(k_magic, sizeof(k_magic));
	w.pod(r.fingerprint);
	w.pod(r.seed);
	w.pod(r.variant);
	w.pod(r.first_id);
	w.pod(r.n_traj);
	w.pod(r.windows);
	w.pod(r.n_obs);
	w.pod(r.n_words);
	w.pod(r.discrete);
	w.pod(r.tick);
	w.pod(r.max_time);
	w.pod(r.norm);
	w.pod((uint32_t)r.user_norm);
	w.pod(r.kernel);
	w.pod(r.transitions);
	w.pod(r.finished);
	w.pod(r.truncated);
	std::vector<uint64_t> k(r.win.n), v(r.win.n);
	for (uint64_t i = 0; i < r.win.n; i++)
	{
		k[i] = r.win.k(i);
		v[i] = r.win.v(i);
	}
	w.vec(k.data(), k.size());
	w.vec(v.data(), v.size());
	w.vec(r.ent.data(), r.ent.size());
	k.assign(r.fin.n, 0);
	v.assign(r.fin.n, 0);
	for (uint64_t i = 0; i < r.fin.n; i++)
	{
		k[i] = r.fin.k(i);
		v[i] = r.fin.v(i);
	}
	w.vec(k.data(), k.size());
	w.vec(v.data(), v.size());
	w.vec(r.fp_words.data(), r.fp_words.size());
	w.vec(r.fp_count.data(), r.fp_count.size());
	w.pod((uint32_t)r.has_moments);
	if (r.has_moments)
	{
		w.vec(r.msq.data(), r.msq.size());
		w.vec(r.mesq.data(), r.mesq.size());
	}
	fnv_hash f;
	f.bytes(w.buf.data(), w.buf.size());
	w.pod(f.h);
	// replaced atomically: a crash mid-write leaves the previous checkpoint
	const std::string tmp = path + ".tmp" + std::to_string((unsigned long long)::getpid());
	{
		std::ofstream o(tmp, std::ios::binary);
		o.write(w.buf.data(), (std::streamsize)w.buf.size());
		o.flush();
		if (!o)
			throw error(MBG_E_IO, "cannot write " + tmp);
	}
	if (std::rename(tmp.c_str(), path.c_str()) != 0)
		throw error(MBG_E_IO, "cannot replace " + path);
}

std::unique_ptr<result> load_result(const program& p, const std::string& path)
{
	std::ifstream in(path, std::ios::binary);
	if (!in)
		throw error(MBG_E_IO, "cannot open " + path);
	std::vector<char> buf((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
	if (buf.size() < sizeof(k_magic) + 8 || std::memcmp(buf.data(), k_magic, sizeof(k_magic)) != 0)
		throw error(MBG_E_IO, path + " is not a checkpoint");
	fnv_hash f;
	f.bytes(buf.data(), buf.size() - 8);
	uint64_t tail;
	std::memcpy(&tail, buf.data() + buf.size() - 8, 8);
	if (tail != f.h)
		throw error(MBG_E_IO, "checkpoint file " + path + " is damaged or truncated");
	buf.resize(buf.size() - 8);
	reader rd { buf, sizeof(k_magic) };
	const model& m = *p.mdl;
	auto r = std::make_unique<result>();
	r->fingerprint = rd.pod<uint64_t>();
	r->seed = rd.pod<uint64_t>();
	r->variant = rd.pod<uint32_t>();
	r->first_id = rd.pod<uint64_t>();
	r->n_traj = rd.pod<uint64_t>();
	r->windows = rd.pod<int>();
	r->n_obs = rd.pod<int>();
	r->n_words = rd.pod<int>();
	r->discrete = rd.pod<int>();
	r->tick = rd.pod<float>();
	r->max_time = rd.pod<float>();
	r->norm = rd.pod<double>();
	r->user_norm = rd.pod<uint32_t>() != 0;
	r->kernel = rd.pod<int>();
	r->transitions = rd.pod<uint64_t>();
	r->finished = rd.pod<uint64_t>();
	r->truncated = rd.pod<uint64_t>();
	if (r->windows != m.windows() || r->n_obs != m.n_observed() || r->n_words != m.n_words())
		throw error(MBG_E_ARG, "checkpoint " + path + " belongs to another model");
	const uint64_t lim = (uint64_t)buf.size();
	auto wk = rd.vec<uint64_t>(lim), wv = rd.vec<uint64_t>(lim);
	r->ent = rd.vec<uint64_t>(lim);
	auto fk = rd.vec<uint64_t>(lim), fv = rd.vec<uint64_t>(lim);
	r->fp_words = rd.vec<uint32_t>(lim);
	r->fp_count = rd.vec<uint64_t>(lim);
	if (wk.size() != wv.size() || fk.size() != fv.size() || r->ent.size() != (size_t)r->windows
		|| r->fp_words.size() != r->fp_count.size() * (size_t)r->n_words)
		throw error(MBG_E_IO, "checkpoint file " + path + " is damaged");
	r->win = cell_view::from(std::move(wk), std::move(wv));
	r->fin = cell_view::from(std::move(fk), std::move(fv));
	r->has_moments = rd.pod<uint32_t>() != 0;
	if (r->has_moments)
	{
		r->msq = rd.vec<unsigned __int128>(lim);
		r->mesq = rd.vec<unsigned __int128>(lim);
		if (r->msq.size() != ((size_t)r->windows << r->n_obs) || r->mesq.size() != (size_t)r->windows)
			throw error(MBG_E_IO, "checkpoint file " + path + " is damaged");
	}
	for (const auto& n : m.nodes)
		r->names.push_back(n.name);
	r->observed = m.observed_nodes();
	r->index_cells();
	return r;
}

void merge_into(result& a, const result& b)
{
	a.wait();
	b.wait();
	check_same_run(a, b);
	if (b.first_id != a.first_id + a.n_traj && a.first_id != b.first_id + b.n_traj)
		throw error(MBG_E_ARG, "the results' trajectory ranges are not adjacent");
	a.win = merge_cells(a.win, b.win);
	a.fin = merge_cells(a.fin, b.fin);
	for (int w = 0; w < a.windows; w++)
		a.ent[w] += b.ent[w];
	// fixed points: both sorted most significant word first
	const size_t W = (size_t)a.n_words;
	auto less = [&](const uint32_t* x, const uint32_t* y) {
		for (size_t k = W; k-- > 0;)
			if (x[k] != y[k])
				return x[k] < y[k];
		return false;
	};
	std::vector<uint32_t> words;
	std::vector<uint64_t> count;
	size_t i = 0, j = 0;
	while (i < a.fp_count.size() || j < b.fp_count.size())
	{
		const uint32_t* x = i < a.fp_count.size() ? a.fixed_state(i) : nullptr;
		const uint32_t* y = j < b.fp_count.size() ? b.fixed_state(j) : nullptr;
		if (y == nullptr || (x != nullptr && less(x, y)))
		{
			words.insert(words.end(), x, x + W);
			count.push_back(a.fp_count[i++]);
		}
		else if (x == nullptr || less(y, x))
		{
			words.insert(words.end(), y, y + W);
			count.push_back(b.fp_count[j++]);
		}
		else
		{
			words.insert(words.end(), x, x + W);
			count.push_back(a.fp_count[i++] + b.fp_count[j++]);
		}
	}
	a.fp_words.swap(words);
	a.fp_count.swap(count);
	if (a.has_moments)
	{
		for (size_t c = 0; c < a.msq.size(); c++)
			a.msq[c] += b.msq[c];
		for (int w = 0; w < a.windows; w++)
			a.mesq[w] += b.mesq[w];
		a.moment_spills += b.moment_spills;
	}
	a.first_id = std::min(a.first_id, b.first_id);
	a.n_traj += b.n_traj;
	if (!a.user_norm)
		a.norm = (double)a.n_traj;
	a.transitions += b.transitions;
	a.finished += b.finished;
	a.truncated += b.truncated;
	a.kernel_ms += b.kernel_ms;
	a.simulate_ms += b.simulate_ms;
	a.total_ms += b.total_ms;
	a.index_cells();
}

std::unique_ptr<result> simulate_checkpointed(program& p, const sim_params& sp, uint64_t chunk,
											  const std::string& path, hipStream_t stream)
{
	scope_timer tm("runner> checkpointed");
	const model& m = *p.mdl;
	if (sp.device_only)
		throw error(MBG_E_ARG, "a checkpointed run keeps its statistics on the host");
	const uint64_t n_total = sp.n_traj < 0 ? (uint64_t)m.sample_count() : (uint64_t)sp.n_traj;
	if (!chunk)
		chunk = n_total ? n_total : 1;
	std::unique_ptr<result> acc;
	struct stat st;
	if (::stat(path.c_str(), &st) == 0)
	{
		acc = load_result(p, path);
		const uint64_t seed = sp.seed < 0 ? m.seed() : (uint64_t)sp.seed;
		if (acc->fingerprint != run_fingerprint(p, sp) || acc->seed != seed || acc->variant != sp.variant)
			throw error(MBG_E_ARG, "checkpoint " + path + " belongs to another model, rate vector, seed or stream tag");
		if (acc->first_id != sp.first_id || acc->n_traj > n_total)
			throw error(MBG_E_ARG, "checkpoint " + path + " covers other trajectory ids");
		if (acc->has_moments != sp.errors)
			throw error(MBG_E_ARG, "checkpoint " + path + (sp.errors ? " has no" : " has") + " second moments");
	}
	uint64_t done = acc ? acc->n_traj : 0;
	while (done < n_total || !acc)
	{
		sim_params c = sp;
		c.first_id = sp.first_id + done;
		c.n_traj = (int64_t)std::min<uint64_t>(chunk, n_total - done);
		std::unique_ptr<result> r = simulate(p, c, stream);
		if (sp.errors)
			run_moments(p, c, stream, *r);
		if (acc)
			merge_into(*acc, *r);
		else
			acc = std::move(r);
		done += (uint64_t)c.n_traj;
		save_result(*acc, path);
	}
	return acc;
}

} // namespace mbg
