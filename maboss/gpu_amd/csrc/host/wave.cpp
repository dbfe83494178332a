// Node tables of the wave-per-trajectory kernel.
//
// For each node of a fast-form model (rate_up = @logic ? U : 0, rate_down =
// @logic ? 0 : D with U, D free of node references; src/generator.cpp:57-107)
// the logic formula is tabulated over its distinct parent nodes by the host
// interpreter: bit j of the truth table is the logic value when the parents'
// bits spell j.  The child lists (nodes whose logic reads a node) drive the
// incremental re-evaluation after a flip.
#include <algorithm>
#include <set>

#include "engine.h"

namespace mbg
{

namespace
{

void collect_nodes(const model& m, const node_t& cur, const expr& e, std::set<int>& out, int depth)
{
	if (depth > 64)
		throw error(MBG_E_MODEL, "alias nesting too deep in node " + cur.name);
	if (e.kind == op::node)
		out.insert(m.node_index(e.name));
	if (e.kind == op::alias)
	{
		if (const expr* t = cur.attr(e.name.substr(1)))
			collect_nodes(m, cur, *t, out, depth + 1);
		return;
	}
	for (const expr* c : { e.a.get(), e.b.get(), e.c.get() })
		if (c)
			collect_nodes(m, cur, *c, out, depth);
}

bool fast_rate(const model& m, const node_t& n, const expr* e, bool up, float& val)
{
	if (!e || e->kind != op::ternary || e->a->kind != op::alias || e->a->name != "@logic")
		return false;
	const expr* zero = up ? e->c.get() : e->b.get();
	const expr* v = up ? e->b.get() : e->c.get();
	try
	{
		if (m.evaluate(*zero) != 0.0f)
			return false;
		val = m.evaluate(*v);
	}
	catch (const error&)
	{
		return false;
	}
	(void)n;
	return true;
}

} // namespace

wave_tables build_wave_tables(const model& m)
{
	wave_tables t;
	const int N = m.n_nodes();
	if (N > 64 * 32)
		throw error(MBG_E_LIMIT, "the wave kernel supports up to 2048 nodes");
	t.npl = (N + 63) / 64;
	while (t.rows < t.npl)
		t.rows <<= 1;
	std::vector<std::vector<int>> children(N);
	t.nodes.resize(N);
	const auto obs = m.observed_nodes();
	if (obs.size() > 63)
		throw error(MBG_E_LIMIT, "more than 63 observed nodes");
	std::vector<int> obs_pos(N, -1);
	for (size_t j = 0; j < obs.size(); j++)
		obs_pos[obs[j]] = (int)j;

	for (int i = 0; i < N; i++)
	{
		const node_t& n = m.nodes[i];
		mbg_wnode& w = t.nodes[i];
		float up = 0, down = 0;
		const expr* logic = n.attr("logic");
		if (!logic || !fast_rate(m, n, n.attr("rate_up"), true, up) || !fast_rate(m, n, n.attr("rate_down"), false, down))
			throw error(MBG_E_LIMIT, "node " + n.name
										 + " is not of the form rate_up = @logic ? u : 0, rate_down = @logic ? 0 : d "
										   "(required by the wave kernel)");
		std::set<int> par;
		collect_nodes(m, n, *logic, par, 0);
		if (par.size() > MBG_WAVE_MAXPAR)
			throw error(MBG_E_LIMIT, "node " + n.name + " has " + std::to_string(par.size())
										 + " inputs; the wave kernel supports up to 6");
		std::vector<int> pv(par.begin(), par.end());
		w.npar = (unsigned char)pv.size();
		for (size_t j = 0; j < pv.size(); j++)
			w.par[j] = (unsigned short)pv[j];
		w.obs_pos = obs_pos[i] >= 0 ? (unsigned short)obs_pos[i] : (unsigned short)MBG_WAVE_NOOBS;
		w.up = up;
		w.down = down;
		std::vector<uint32_t> st((size_t)m.n_words(), 0u);
		for (uint32_t a = 0; a < (1u << pv.size()); a++)
		{
			std::fill(st.begin(), st.end(), 0u);
			for (size_t j = 0; j < pv.size(); j++)
				if ((a >> j) & 1u)
					st[pv[j] / 32] |= 1u << (pv[j] % 32);
			if (m.eval_in_state(n, *logic, st.data()) != 0.0f)
				w.tt |= 1ull << a;
		}
		for (int p : pv)
			if (p != i)
				children[p].push_back(i);
	}
	t.child_off.resize(N + 1);
	for (int i = 0; i < N; i++)
	{
		t.child_off[i] = (unsigned)t.child.size();
		for (int c : children[i])
			t.child.push_back((unsigned short)c);
	}
	t.child_off[N] = (unsigned)t.child.size();
	for (int o : obs)
		t.obs.push_back((unsigned short)o);

	// one contiguous image, staged into LDS as is by the kernel
	auto put = [&](const void* p, size_t n) {
		const char* c = (const char*)p;
		t.image.insert(t.image.end(), c, c + n);
	};
	put(t.nodes.data(), t.nodes.size() * sizeof(mbg_wnode));
	put(t.child_off.data(), t.child_off.size() * 4);
	put(t.child.data(), t.child.size() * 2);
	put(t.obs.data(), t.obs.size() * 2);
	// LDS rate slot of every node, then of every observed node (slot of node
	// i: lane i / npl, row i % npl of the wave's [npl][64] rate array)
	std::vector<unsigned short> ra(N), oba(obs.size());
	for (int i = 0; i < N; i++)
		ra[i] = (unsigned short)((i % t.npl) * 64 + i / t.npl);
	for (size_t j = 0; j < obs.size(); j++)
		oba[j] = ra[obs[j]];
	put(ra.data(), ra.size() * 2);
	put(oba.data(), oba.size() * 2);
	while (t.image.size() % 16)
		t.image.push_back(0);
	// per observed node {up, down, up*log2(up), down*log2(down)}: the entropy
	// term of a nonzero rate without a log per step
	t.obs_tab_off = (int)t.image.size();
	for (int o : obs)
	{
		const mbg_wnode& w = t.nodes[o];
		const float v[4] = { w.up, w.down, mbg_rlog2r(w.up), mbg_rlog2r(w.down) };
		put(v, sizeof(v));
	}
	return t;
}

} // namespace mbg
