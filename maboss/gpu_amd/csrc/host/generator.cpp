// Per-model HIP source generator.
//
// Follows src/generator.cpp: one rate per node (generate_node_transitions
// :84-122, including the "fast form" of is_logic_ternary_expression :57-82),
// the sequential total (generate_aggregate_function :124-144), the transition
// entropy of the observed nodes (:146-175) and the observed-state index
// (generate_non_internal_index :207-235).  Differences, all deliberate:
//   * constants are emitted as exact hexadecimal floats instead of 6
//     significant digits (src/expressions.cpp:185, :213; src/generator.cpp:23);
//   * expressions are typed: booleans stay booleans, arithmetic is float, so
//     every operator has the float semantics of the reference's evaluate()
//     (src/parser/expressions.cpp:10-81) rather than C++ int/double promotion;
//   * XOR is supported (bool inequality);
//   * the flip selection (src/jit_kernels/simulation.cu:9-21) is emitted with
//     the rates, reusing the same prefix sums as the total;
//   * a second flavour reads the $variables from a runtime table and applies
//     per-variant mutant masks (batched and differentiated forms).
#include "generator.h"

#include <cmath>
#include <cstdio>
#include <map>
#include <set>
#include <sstream>

#include "../device/mbg_numeric.h"

namespace mbg
{

float rate_log2_rate(float r) { return mbg_rlog2r(r); }

std::string float_literal(float v)
{
	if (v == 0.0f)
		return std::signbit(v) ? "(-0.0f)" : "0.0f";
	char buf[64];
	std::snprintf(buf, sizeof(buf), "%af", (double)v);
	std::string s(buf);
	if (s[0] == '-')
		return "(" + s + ")";
	return s;
}

namespace
{

enum class ty
{
	b, // C++ bool expression
	f, // float
	u  // uint32_t 0 or 1 (node values and the logic built from them stay integers)
};

struct code
{
	std::string s;
	ty t;
};

const expr* strip_parens(const expr* e)
{
	while (e && e->kind == op::paren)
		e = e->a.get();
	return e;
}

bool node_free(const model& m, const node_t& cur, const expr& e, int depth = 0)
{
	if (depth > 64)
		return false;
	switch (e.kind)
	{
		case op::node: return false;
		case op::alias: {
			const expr* t = cur.attr(e.name.substr(1));
			return t && node_free(m, cur, *t, depth + 1);
		}
		default: break;
	}
	for (const expr* c : { e.a.get(), e.b.get(), e.c.get() })
		if (c && !node_free(m, cur, *c, depth))
			return false;
	return true;
}

class emitter
{
public:
	// Node values are read as bits of the state words named <words>0, <words>1, ...
	emitter(const model& m, const node_t& cur, bool param, const std::map<std::string, int>& pidx,
			const std::string& words = "sw")
		: m_(m), cur_(cur), param_(param), pidx_(pidx), words_(words)
	{}

	int literals = 0;

	code nat(const expr& e)
	{
		switch (e.kind)
		{
			case op::lit: return { float_literal(e.value), ty::f };
			case op::node: {
				int i = m_.node_index(e.name);
				if (i < 0)
					throw error(MBG_E_MODEL, "unknown node name: " + e.name);
				literals++;
				return { bit(i), ty::u };
			}
			case op::var: {
				auto it = m_.variables.find(e.name);
				if (it == m_.variables.end())
					throw error(MBG_E_MODEL, "undefined variable " + e.name);
				if (param_)
					return { "pv->rate[" + std::to_string(pidx_.at(e.name)) + "]", ty::f };
				return { float_literal(it->second), ty::f };
			}
			case op::alias: {
				std::string an = e.name.substr(1);
				const expr* t = cur_.attr(an);
				if (!t)
					throw error(MBG_E_MODEL, "Node " + cur_.name + " does not have attribute " + an);
				if (stack_.count(an))
					throw error(MBG_E_MODEL, "cyclic alias @" + an + " in node " + cur_.name);
				stack_.insert(an);
				code c = nat(*t);
				stack_.erase(an);
				return { "(" + c.s + ")", c.t };
			}
			case op::paren: {
				code c = nat(*e.a);
				return { "(" + c.s + ")", c.t };
			}
			case op::neg: return { "(-" + as_f(*e.a) + ")", ty::f };
			case op::pos: return { as_f(*e.a), ty::f };
			case op::lnot: return { "(" + as_u(*e.a) + " ^ 1u)", ty::u };
			case op::land: return { "(" + as_u(*e.a) + " & " + as_u(*e.b) + ")", ty::u };
			case op::lor: return { "(" + as_u(*e.a) + " | " + as_u(*e.b) + ")", ty::u };
			case op::lxor: return { "(" + as_u(*e.a) + " ^ " + as_u(*e.b) + ")", ty::u };
			case op::eq:
			case op::ne: {
				code l = nat(*e.a), r = nat(*e.b);
				const char* o = e.kind == op::eq ? " == " : " != ";
				if (l.t != ty::f && r.t != ty::f)
					return { "(" + to_u(l) + o + to_u(r) + ")", ty::b };
				return { "(" + to_f(l) + o + to_f(r) + ")", ty::b };
			}
			case op::lt: return cmp(e, " < ");
			case op::le: return cmp(e, " <= ");
			case op::gt: return cmp(e, " > ");
			case op::ge: return cmp(e, " >= ");
			case op::add: return arith(e, " + ");
			case op::sub: return arith(e, " - ");
			case op::mul: return arith(e, " * ");
			case op::div: return arith(e, " / ");
			case op::ternary: {
				std::string c = as_b(*e.a);
				code l = nat(*e.b), r = nat(*e.c);
				if (l.t != ty::f && r.t != ty::f)
					return { "(" + c + " ? " + to_u(l) + " : " + to_u(r) + ")", ty::u };
				return { "(" + c + " ? " + to_f(l) + " : " + to_f(r) + ")", ty::f };
			}
		}
		throw error(MBG_E_INTERNAL, "bad expression");
	}

	std::string as_b(const expr& e) { return to_b(nat(e)); }
	std::string as_f(const expr& e) { return to_f(nat(e)); }
	std::string as_u(const expr& e) { return to_u(nat(e)); }

	// The value of a boolean expression at bit t of a 32-bit word (other
	// bits unspecified): node values are shifted from their own bit straight
	// to bit t and combined with word-wide logic operations.  Each node's
	// logic then uses shifts of its own (they depend on its bit position), so
	// the compiler does not share extracted bits between nodes and keep them
	// live across the whole rate evaluation.
	std::string as_aligned(const expr& e, int t)
	{
		switch (e.kind)
		{
			case op::node: {
				const int i = m_.node_index(e.name);
				if (i < 0)
					throw error(MBG_E_MODEL, "unknown node name: " + e.name);
				literals++;
				const int b = i % 32;
				const std::string w = words_ + std::to_string(i / 32);
				if (b > t)
					return "(" + w + " >> " + std::to_string(b - t) + "u)";
				if (b < t)
					return "(" + w + " << " + std::to_string(t - b) + "u)";
				return w;
			}
			case op::alias: {
				std::string an = e.name.substr(1);
				const expr* d = cur_.attr(an);
				if (!d)
					throw error(MBG_E_MODEL, "Node " + cur_.name + " does not have attribute " + an);
				if (stack_.count(an))
					throw error(MBG_E_MODEL, "cyclic alias @" + an + " in node " + cur_.name);
				stack_.insert(an);
				std::string c = as_aligned(*d, t);
				stack_.erase(an);
				return c;
			}
			case op::paren: return as_aligned(*e.a, t);
			case op::lnot: return "(~" + as_aligned(*e.a, t) + ")";
			case op::land: return "(" + as_aligned(*e.a, t) + " & " + as_aligned(*e.b, t) + ")";
			case op::lor: return "(" + as_aligned(*e.a, t) + " | " + as_aligned(*e.b, t) + ")";
			case op::lxor: return "(" + as_aligned(*e.a, t) + " ^ " + as_aligned(*e.b, t) + ")";
			default: {
				const std::string u = as_u(e);
				return t ? "(" + u + " << " + std::to_string(t) + "u)" : u;
			}
		}
	}
	// value (0 or 1, uint32_t) of node i
	std::string bit(int i) const
	{
		return "((" + words_ + std::to_string(i / 32) + " >> " + std::to_string(i % 32) + "u) & 1u)";
	}

private:
	static std::string to_b(const code& c)
	{
		return c.t == ty::b ? c.s : c.t == ty::u ? "(" + c.s + " != 0u)" : "(" + c.s + " != 0.0f)";
	}
	static std::string to_f(const code& c)
	{
		return c.t == ty::f ? c.s : c.t == ty::u ? "((float)" + c.s + ")" : "(" + c.s + " ? 1.0f : 0.0f)";
	}
	static std::string to_u(const code& c)
	{
		return c.t == ty::u ? c.s : c.t == ty::b ? "(" + c.s + " ? 1u : 0u)" : "((" + c.s + " != 0.0f) ? 1u : 0u)";
	}

	code cmp(const expr& e, const char* o) { return { "(" + as_f(*e.a) + o + as_f(*e.b) + ")", ty::b }; }
	code arith(const expr& e, const char* o) { return { "(" + as_f(*e.a) + o + as_f(*e.b) + ")", ty::f }; }

	const model& m_;
	const node_t& cur_;
	bool param_;
	const std::map<std::string, int>& pidx_;
	std::string words_;
	std::set<std::string> stack_;
};

// src/generator.cpp:57-82: `@logic ? V : 0` (up) or `@logic ? 0 : V` (down).
bool logic_ternary(const model& m, const node_t& n, const expr* e, bool up, const expr*& val)
{
	if (!e || e->kind != op::ternary)
		return false;
	const expr* c = e->a.get();
	if (c->kind != op::alias || c->name != "@logic" || !n.attr("logic"))
		return false;
	const expr* zero = up ? e->c.get() : e->b.get();
	val = up ? e->b.get() : e->c.get();
	if (!node_free(m, n, *zero) || !node_free(m, n, *val))
		return false;
	try
	{
		if (m.evaluate(*zero) != 0.0f)
			return false;
	}
	catch (const error&)
	{
		return false;
	}
	return true;
}

int bare_param(const expr* e, const std::map<std::string, int>& pidx)
{
	e = strip_parens(e);
	if (e && e->kind == op::var)
	{
		auto it = pidx.find(e->name);
		if (it != pidx.end())
			return it->second;
	}
	return -1;
}

void collect_vars(const model& m, const node_t& n, const expr& e, std::set<std::string>& out, int depth = 0)
{
	if (depth > 64)
		return;
	if (e.kind == op::var)
		out.insert(e.name);
	if (e.kind == op::alias)
	{
		if (const expr* t = n.attr(e.name.substr(1)))
			collect_vars(m, n, *t, out, depth + 1);
		return;
	}
	for (const expr* c : { e.a.get(), e.b.get(), e.c.get() })
		if (c)
			collect_vars(m, n, *c, out, depth);
}

// Rate evaluation.  Each node's rate is one of 0, its up rate or its down
// rate; which one is decided by its own value x and by whether its logic
// differs from x (the node is "enabled").  The first pass evaluates every
// node's logic once, collects the enabled bits in words E (bit i = node i
// enabled), and sums the rates in node order (the total, and the entropy
// sums of the observed nodes).  The flip selection needs the prefix sums
// again once the total is known: models up to k_one_pass_nodes keep them in
// registers, larger ones recompute them in a second pass from E and x with
// the same float additions in the same order (so the selected node is the
// same), reading the state through an empty asm statement so the compiler
// does not merge the two passes' bit extractions.
//
// Grouped second pass (baked rates, every node in the fast form, log-sum
// entropy): the first pass also keeps the running sum at the end of every
// group of k_group nodes.  The prefix sums are non-decreasing (rates are
// non-negative), so the selected node lies in the first group whose end sum
// reaches u2 * total, and every node before that group counts.  Only that
// group is re-summed, from the kept sum at its start, with the rates read
// from a table of (up, down) bit patterns: the same additions on the same
// values as the full pass, so the same node is selected, for k_group nodes'
// work instead of N.
constexpr int k_one_pass_nodes = 40;
constexpr int k_group = 8;

void emit_eval(std::ostringstream& os, const model& m, const gen_options& opt, bool param,
			   const std::map<std::string, int>& pidx, generated& g)
{
	const int N = m.n_nodes(), W = m.n_words();
	const auto obs = m.observed_nodes();
	std::vector<bool> is_obs(N, false);
	for (int i : obs)
		is_obs[i] = true;
	const bool log_sum = opt.entropy == entropy_formula::log_sum;
	const bool two_pass = N > k_one_pass_nodes;

	struct node_code
	{
		bool fast = false;
		int up_p = -1, dn_p = -1;
		const expr *U = nullptr, *D = nullptr;
	};
	std::vector<node_code> nc(N);
	for (int i = 0; i < N; i++)
	{
		const node_t& n = m.nodes[i];
		nc[i].fast = logic_ternary(m, n, n.attr("rate_up"), true, nc[i].U)
					 && logic_ternary(m, n, n.attr("rate_down"), false, nc[i].D);
		if (nc[i].fast)
		{
			nc[i].up_p = bare_param(nc[i].U, pidx);
			nc[i].dn_p = bare_param(nc[i].D, pidx);
		}
	}

	bool all_fast = true;
	for (int i = 0; i < N; i++)
		all_fast = all_fast && nc[i].fast;
	const bool grouped = two_pass && log_sum && all_fast;
	const int n_groups = (N + k_group - 1) / k_group;
	if (param)
	{
		// The grouped second pass of mbg_eval_p reads the (down << 32 | up)
		// rate bit patterns of the wave's rate table: one table per wave (up
		// to 4 per block), filled by mbg_eval_p_setup from the parameter view
		// before the wave's first mbg_eval_p with that view.
		if (grouped)
		{
			os << "__attribute__((aligned(16))) static __shared__ unsigned long long mbg_udp_lds[4][" << n_groups * k_group
			   << "];\n\n";
			os << "extern \"C\" __device__ void mbg_eval_p_setup(const mbg_param_view* __restrict__ pv)\n{\n";
			os << "\tunsigned long long* t = mbg_udp_lds[(threadIdx.x >> 6) & 3];\n";
			os << "\tif ((threadIdx.x & 63) == 0)\n\t{\n";
			for (int i = 0; i < N; i++)
			{
				emitter em(m, m.nodes[i], param, pidx, "sw");
				std::set<std::string> uv, dv;
				collect_vars(m, m.nodes[i], *nc[i].U, uv);
				collect_vars(m, m.nodes[i], *nc[i].D, dv);
				auto code = [&](const expr* v, int p_i, const std::set<std::string>& vars) -> std::string {
					if (p_i >= 0)
						return "mbg_f2u(pv->rate[" + std::to_string(p_i) + "])";
					if (vars.empty())
					{
						char buf[32];
						std::snprintf(buf, sizeof(buf), "0x%08xu", mbg_f2u(m.evaluate(*v)));
						return buf;
					}
					return "mbg_f2u(" + em.as_f(*v) + ")";
				};
				os << "\t\tt[" << i << "] = (unsigned long long)" << code(nc[i].D, nc[i].dn_p, dv) << " << 32 | "
				   << code(nc[i].U, nc[i].up_p, uv) << ";\n";
			}
			os << "\t}\n\tMBG_WAVE_SYNC();\n}\n\n";
		}
		else
			os << "extern \"C\" __device__ void mbg_eval_p_setup(const mbg_param_view* __restrict__ pv) { (void)pv; }\n\n";
	}
	else
	{
		// The grouped second pass reads (down << 32 | up) rate bit patterns per
		// node, padded to whole groups, from a block-shared copy of this table
		// that mbg_eval_setup fills before the kernel's first barrier.
		if (grouped)
		{
			os << "__attribute__((aligned(16))) static __device__ const unsigned long long mbg_ud[" << n_groups * k_group
			   << "] = {";
			for (int i = 0; i < n_groups * k_group; i++)
			{
				uint64_t v = 0;
				if (i < N)
					v = (uint64_t)mbg_f2u(m.evaluate(*nc[i].D)) << 32 | mbg_f2u(m.evaluate(*nc[i].U));
				char buf[48];
				std::snprintf(buf, sizeof(buf), "%s0x%016llxull", i == 0 ? "\n\t" : i % 4 ? ", " : ",\n\t",
							  (unsigned long long)v);
				os << buf;
			}
			os << "};\n__attribute__((aligned(16))) static __shared__ unsigned long long mbg_ud_lds["
			   << n_groups * k_group << "];\n\n";
			os << "extern \"C\" __device__ void mbg_eval_setup()\n{\n\tfor (int i = threadIdx.x; i < "
			   << n_groups * k_group << "; i += blockDim.x)\n\t\tmbg_ud_lds[i] = mbg_ud[i];\n}\n\n";
		}
		else
			os << "extern \"C\" __device__ void mbg_eval_setup() {}\n\n";
	}

	// bit patterns of a fast-form node's rates (and r*log2(r) of them)
	auto bits_of = [&](emitter& em, const expr* v, int pidx_v, bool lr) -> std::string {
		if (param)
		{
			if (pidx_v >= 0)
				return "mbg_f2u(pv->" + std::string(lr ? "rate_lr[" : "rate[") + std::to_string(pidx_v) + "])";
			return lr ? "mbg_f2u(mbg_rlog2r(" + em.as_f(*v) + "))" : "mbg_f2u(" + em.as_f(*v) + ")";
		}
		const float val = m.evaluate(*v);
		char buf[32];
		std::snprintf(buf, sizeof(buf), "0x%08xu", mbg_f2u(lr ? rate_log2_rate(val) : val));
		return buf;
	};
	// the selected rate r (and l) of a fast-form node from the masks xm (x)
	// and msk (enabled), branch-free: same value as (x == L) ? 0 : (x ? D : U)
	// With run-time rates ub and db are wave-uniform (scalar registers), and a
	// vector instruction reads one of them: the select is written as two
	// three-input ops, ((xm & msk & (ub ^ db)) ^ (msk & ub)), each reading one
	// (truth tables 0x80 and 0x78), which saves the copy of ub into a vector
	// register the compiler's own form needs per node.
	auto select_expr = [&](const char* u, const char* d) {
		if (param)
			return std::string("mbg_u2f(MBG_BITOP3(MBG_BITOP3(xm, msk, ") + u + " ^ " + d + ", 0x80), msk, " + u
				   + ", 0x78))";
		return std::string("mbg_u2f((") + u + " ^ ((" + u + " ^ " + d + ") & xm)) & msk)";
	};
	auto select_rate = [&](emitter& em, int i, bool with_l) {
		os << "\t\tconst uint32_t ub = " << bits_of(em, nc[i].U, nc[i].up_p, false)
		   << ", db = " << bits_of(em, nc[i].D, nc[i].dn_p, false) << ";\n";
		os << "\t\tr = " << select_expr("ub", "db") << ";\n";
		if (with_l)
		{
			os << "\t\tconst uint32_t ul = " << bits_of(em, nc[i].U, nc[i].up_p, true)
			   << ", dl = " << bits_of(em, nc[i].D, nc[i].dn_p, true) << ";\n";
			os << "\t\tl = " << select_expr("ul", "dl") << ";\n";
		}
	};
	// a node outside the fast form: its rate expressions as written
	auto general_rate = [&](emitter& em, int i, bool with_l) {
		const node_t& n = m.nodes[i];
		const int w = i / 32, b = i % 32;
		const std::string rd = em.as_f(*n.attr("rate_down")), ru = em.as_f(*n.attr("rate_up"));
		os << "\t\tconst uint32_t x = " << em.bit(i) << ";\n";
		os << "\t\tr = x ? " << rd << " : " << ru << ";\n";
		if (param)
			os << "\t\tif ((((x ? de" << w << " : ue" << w << ") >> " << b << "u) & 1u) == 0u) r = 0.0f;\n";
		if (with_l)
			os << "\t\tl = mbg_rlog2r(r);\n";
	};

	if (param)
		os << "extern \"C\" __device__ void mbg_eval_p(const uint32_t* __restrict__ s, float u2, "
			  "const mbg_param_view* __restrict__ pv, mbg_eval_t* __restrict__ out, uint32_t* __restrict__ en)\n{\n";
	else
		os << "extern \"C\" __device__ void mbg_eval(const uint32_t* __restrict__ s, float u2, "
			  "mbg_eval_t* __restrict__ out)\n{\n";
	for (int w = 0; w < W; w++)
		os << "\tconst uint32_t sw" << w << " = s[" << w << "];\n\tuint32_t E" << w << " = 0u;\n";
	if (param)
		for (int w = 0; w < W; w++)
			os << "\tconst uint32_t ue" << w << " = pv->up_en[" << w << "], de" << w << " = pv->dn_en[" << w
			   << "];\n\tconst uint32_t al" << w << " = ue" << w << " ^ ((ue" << w << " ^ de" << w << ") & sw" << w
			   << "); // directions allowed in the current state\n";
	// with baked rates and log-sum entropy, the count of observed nodes with a
	// nonzero rate comes from E, x and masks of the nonzero up / down rates
	const bool nnz_masks = log_sum && !param && all_fast;
	os << "\tfloat acc = 0.0f, S = 0.0f, T = 0.0f;\n\tint nnz = 0;\n";
	os << "\t// pass 1: logic, enabled bits, total and entropy sums\n";
	for (int i = 0; i < N; i++)
	{
		const node_t& n = m.nodes[i];
		const int w = i / 32, b = i % 32;
		const bool lw = is_obs[i] && log_sum;
		emitter em(m, n, param, pidx, "sw");
		os << "\t// node " << i << ": " << n.name << (is_obs[i] ? " (observed)" : "") << "\n";
		if (!two_pass)
			os << "\tfloat r" << i << ";\n";
		os << "\t{\n\t\tfloat r" << (lw ? ", l" : "") << ";\n";
		if (nc[i].fast)
		{
			// xv: bit b set when the node is enabled (its logic differs from
			// its value and, with run-time masks, that direction is allowed);
			// one value feeds both the enabled word and the mask, so the
			// compiler does not evaluate the logic twice
			os << "\t\tuint32_t xv = (sw" << w << " ^ " << em.as_aligned(*n.attr("logic"), b) << ")"
			   << (param ? " & al" + std::to_string(w) : std::string()) << ";\n\t\tMBG_MASK(xv);\n";
			os << "\t\tE" << w << " |= xv & " << (1u << b) << "u;\n";
			os << "\t\tuint32_t xm = 0u - ((sw" << w << " >> " << b << "u) & 1u);\n\t\tMBG_MASK(xm);\n";
			os << "\t\tuint32_t msk = 0u - ((xv >> " << b << "u) & 1u);\n";
			select_rate(em, i, lw);
		}
		else
			general_rate(em, i, lw);
		if (!two_pass)
			os << "\t\tr" << i << " = r;\n";
		os << "\t\tacc = acc + r;\n";
		if (lw && nnz_masks)
			os << "\t\tS = S + r; T = T + l;\n";
		else if (lw)
			os << "\t\tS = S + r; T = T + l; nnz += (r != 0.0f) ? 1 : 0;\n";
		else if (is_obs[i])
			os << "\t\tS = S + r;\n";
		// the sums and the enabled word are pinned after every 4th node: each
		// group's terms are added before the next group is computed (no sum
		// sunk to the end with every term kept live), while the scheduler may
		// still interleave the group's nodes (with the selection pass below
		// unpinned: the 50-node batched kernel 8 % faster, the 100-node plain
		// kernel about 1 %)
		if (two_pass && (i % 4 == 3 || i == N - 1))
			os << "\t\tMBG_PIN4(acc, S, T, nnz);\n\t\tMBG_OPAQUE(E" << w << ");\n";
		os << "\t}\n";
		if (!two_pass)
			os << "\tconst float c" << i << " = acc;\n";
		if (grouped && i % k_group == k_group - 1 && i / k_group < n_groups - 1)
			os << "\tconst float g" << i / k_group << " = acc;\n";

		node_info& ni = g.nodes[i];
		ni.fast_form = nc[i].fast;
		ni.up_param = nc[i].up_p;
		ni.down_param = nc[i].dn_p;
		ni.literals = em.literals;
		g.total_literals += em.literals;
	}
	os << "\tconst float total = acc;\n";
	if (param)
	{
		// the enabled bits (directions a mutant forbids excluded), for callers
		// that need them (the gradient kernel)
		os << "\tif (en)\n\t{\n";
		for (int w = 0; w < W; w++)
			os << "\t\ten[" << w << "] = E" << w << ";\n";
		os << "\t}\n";
	}
	if (nnz_masks && !obs.empty())
	{
		// observed, enabled, and the rate in the node's direction nonzero
		// (+-0 count as zero, as r != 0.0f does)
		os << "\tnnz = 0";
		for (int w = 0; w < W; w++)
		{
			uint32_t om = 0, unz = 0, dnz = 0;
			for (int i = w * 32; i < std::min(N, w * 32 + 32); i++)
			{
				const uint32_t bit = 1u << (i % 32);
				if (is_obs[i])
					om |= bit;
				if (mbg_f2u(m.evaluate(*nc[i].U)) & 0x7fffffffu)
					unz |= bit;
				if (mbg_f2u(m.evaluate(*nc[i].D)) & 0x7fffffffu)
					dnz |= bit;
			}
			if (!om)
				continue;
			char buf[160];
			std::snprintf(buf, sizeof(buf), " + __builtin_popcount(E%d & ((sw%d & 0x%08xu) | (~sw%d & 0x%08xu)) & 0x%08xu)", w, w,
						  dnz & om, w, unz & om, om);
			os << buf;
		}
		os << ";\n";
	}
	// flip selection (src/jit_kernels/simulation.cu:9-21): the number of
	// prefix sums strictly below u2 * total
	os << "\tconst float rr = u2 * total;\n\tint flip = 0;\n";
	if (!log_sum && !obs.empty())
		os << "\tfloat H = 0.0f;\n";
	if (!two_pass)
	{
		for (int i = 0; i < N; i++)
			os << "\tflip += (c" << i << " < rr) ? 1 : 0;\n";
		if (!log_sum && !obs.empty())
		{
			os << "\tif (S != 0.0f)\n\t{\n\t\tfloat p;\n";
			for (int i : obs)
				os << "\t\tp = r" << i << " / S; H = H - ((p == 0.0f) ? 0.0f : mbg_log2f(p) * p);\n";
			os << "\t}\n";
		}
	}
	else if (grouped)
	{
		os << "\t// pass 2 (grouped): the group holding the selected node, and the sum before it\n";
		os << "\tint gsel = 0;\n\tfloat acc2 = 0.0f;\n";
		for (int k = 0; k < n_groups - 1; k++)
			os << "\tif (g" << k << " < rr) { gsel = " << k + 1 << "; acc2 = g" << k << "; }\n";
		os << "\tMBG_PIN2(acc2, gsel);\n";
		// a group lies inside one state word
		static_assert(32 % k_group == 0, "groups must not straddle state words");
		os << "\tconst int gw = gsel / " << 32 / k_group << ", gb = (gsel % " << 32 / k_group << ") * " << k_group
		   << ";\n";
		// (with run-time masks E already excludes the directions a mutant forbids)
		os << "\tuint32_t ge = E0, gx = sw0;\n";
		for (int w = 1; w < W; w++)
			os << "\tge = (gw == " << w << ") ? E" << w << " : ge;\n\tgx = (gw == " << w << ") ? sw" << w << " : gx;\n";
		os << "\tge >>= gb;\n\tgx >>= gb;\n";
		if (param)
			os << "\tconst unsigned long long* gud = mbg_udp_lds[(threadIdx.x >> 6) & 3] + gsel * " << k_group << ";\n";
		else
			os << "\tconst unsigned long long* gud = mbg_ud_lds + gsel * " << k_group << ";\n";
		os << "\tint cnt = 0;\n";
		os << "#pragma unroll\n\tfor (int j = 0; j < " << k_group << "; ++j)\n\t{\n";
		os << "\t\tconst unsigned long long ud = gud[j];\n";
		os << "\t\tconst uint32_t ub = (uint32_t)ud, db = (uint32_t)(ud >> 32);\n";
		os << "\t\tconst uint32_t xm = 0u - ((gx >> j) & 1u), msk = 0u - ((ge >> j) & 1u);\n";
		os << "\t\tacc2 = acc2 + mbg_u2f((ub ^ ((ub ^ db) & xm)) & msk);\n";
		// the compiler schedules the group freely (no pins)
		os << "\t\tcnt += (acc2 < rr) ? 1 : 0;\n\t}\n";
		os << "\tflip = gsel * " << k_group << " + cnt;\n";
	}
	else
	{
		os << "\t// pass 2: the same prefix sums from E and the state\n";
		for (int w = 0; w < W; w++)
			os << "\tuint32_t sv" << w << " = s[" << w << "];\n\tMBG_OPAQUE(sv" << w << ");\n";
		os << "\tfloat acc2 = 0.0f;\n";
		for (int i = 0; i < N; i++)
		{
			const node_t& n = m.nodes[i];
			const int w = i / 32, b = i % 32;
			emitter em(m, n, param, pidx, "sv");
			os << "\t{\n\t\tfloat r;\n";
			if (nc[i].fast)
			{
				os << "\t\tconst uint32_t xm = 0u - ((sv" << w << " >> " << b << "u) & 1u);\n";
				os << "\t\tuint32_t msk = 0u - ((E" << w << " >> " << b << "u) & 1u);\n";
				select_rate(em, i, false);
			}
			else
				general_rate(em, i, false);
			os << "\t\tacc2 = acc2 + r;\n\t\tflip += (acc2 < rr) ? 1 : 0;\n";
			if (!log_sum && is_obs[i])
				os << "\t\tif (S != 0.0f) { const float p = r / S; H = H - ((p == 0.0f) ? 0.0f : mbg_log2f(p) * p); }\n";
			os << "\t\tMBG_PIN2(acc2, flip);\n\t}\n";
		}
	}

	// transition entropy of the observed nodes (src/generator.cpp:146-175)
	if (obs.empty())
		os << "\tconst float Hout = 0.0f;\n";
	else if (log_sum)
		os << "\tconst float h = mbg_log2f(S) - T / S;\n\tconst float Hout = (nnz > 1 && h > 0.0f) ? h : 0.0f;\n";
	else
		os << "\tconst float Hout = H;\n";
	os << "\t(void)T;\n\t(void)nnz;\n";
	for (int w = 0; w < W; w++)
		os << "\t(void)E" << w << ";\n";
	os << "\tout->total = total;\n\tout->H = Hout;\n\tout->flip = flip;\n\tout->pad = 0;\n}\n\n";
}

void emit_obs_index(std::ostringstream& os, const model& m)
{
	const auto obs = m.observed_nodes();
	os << "extern \"C\" __device__ uint64_t mbg_obs_index(const uint32_t* __restrict__ s)\n{\n\tuint64_t v = 0;\n";
	// Runs of consecutive observed nodes inside one word are extracted with a
	// single mask and shift (a parallel bit-extract in node order).
	size_t j = 0;
	while (j < obs.size())
	{
		int i0 = obs[j];
		size_t k = j + 1;
		while (k < obs.size() && obs[k] == obs[k - 1] + 1 && obs[k] / 32 == i0 / 32)
			k++;
		int len = (int)(k - j);
		uint32_t mask = len == 32 ? 0xffffffffu : ((1u << len) - 1u);
		os << "\tv |= (uint64_t)((s[" << i0 / 32 << "] >> " << i0 % 32 << ") & " << mask << "u) << " << j << ";\n";
		j = k;
	}
	os << "\treturn v;\n}\n\n";
	// the observed-index bit of node i: word masks of the observed nodes and
	// the number of observed nodes in the words before
	const int W = m.n_words(), N = m.n_nodes();
	std::vector<bool> is_obs(N, false);
	for (int i : obs)
		is_obs[i] = true;
	os << "extern \"C\" __device__ uint64_t mbg_obs_flip(int i)\n{\n";
	os << "\tconst int w = i >> 5, b = i & 31;\n";
	int before = 0;
	for (int w = 0; w < W; w++)
	{
		uint32_t om = 0;
		for (int i = w * 32; i < std::min(N, w * 32 + 32); i++)
			if (is_obs[i])
				om |= 1u << (i % 32);
		char buf[128];
		if (w == 0)
			std::snprintf(buf, sizeof(buf), "\tuint32_t om = 0x%08xu;\n\tint base = 0;\n", om);
		else
			std::snprintf(buf, sizeof(buf), "\tom = (w == %d) ? 0x%08xu : om;\n\tbase = (w == %d) ? %d : base;\n", w, om,
						  w, before);
		os << buf;
		before += __builtin_popcount(om);
	}
	os << "\treturn ((om >> b) & 1u) ? 1ull << (base + __builtin_popcount(om & ((1u << b) - 1u))) : 0ull;\n}\n";
}

} // namespace

generated generate(const model& m, const gen_options& opt)
{
	generated g;
	const int N = m.n_nodes();
	g.nodes.resize(N);

	std::set<std::string> used;
	for (auto& n : m.nodes)
		for (const char* a : { "rate_up", "rate_down" })
			if (const expr* e = n.attr(a))
				collect_vars(m, n, *e, used);
	std::map<std::string, int> pidx;
	for (auto& v : used)
	{
		pidx[v] = (int)g.param_names.size();
		g.param_names.push_back(v);
	}

	std::ostringstream os;
	os << "// Generated by the MaBoSS MI355X generator: " << N << " nodes, " << m.n_words() << " state words, "
	   << m.n_observed() << " observed nodes.\n";
	os << "#include \"mbg_abi.h\"\n\n";
	os << "#define MBG_N " << N << "\n#define MBG_W " << m.n_words() << "\n#define MBG_NOBS " << m.n_observed()
	   << "\n\n";
	emit_eval(os, m, opt, false, pidx, g);
	// the parametric flavour does not recount literals
	generated tmp;
	tmp.nodes.resize(N);
	emit_eval(os, m, opt, true, pidx, tmp);
	for (int i = 0; i < N; i++)
	{
		g.nodes[i].up_param = tmp.nodes[i].up_param;
		g.nodes[i].down_param = tmp.nodes[i].down_param;
	}
	emit_obs_index(os, m);
	g.source = os.str();
	return g;
}

// d total / d theta_j in a state is the number of nodes whose current rate
// is theta_j: node i in the fast form has rate theta_u (up_param) when off
// with logic on, theta_d (down_param) when on with logic off, else 0.  The
// emitted mbg_grad_dwell adds the record's dwell to the per-parameter double
// accumulators (T) of the nodes where that indicator is set; mbg_grad_on
// counts, for one parameter, the node directions whose indicator is set.
grad_code generate_grad(const model& m, const generated& g)
{
	const int N = m.n_nodes();
	std::map<std::string, int> pidx;
	for (size_t k = 0; k < g.param_names.size(); k++)
		pidx[g.param_names[k]] = (int)k;
	grad_code gc;
	gc.n_params = (int)g.param_names.size();
	if (gc.n_params == 0)
		throw error(MBG_E_LIMIT, "the model's rates use no $variable to differentiate");
	if (gc.n_params > 128)
		throw error(MBG_E_LIMIT, "the gradient form keeps one register pair per parameter: at most 128 parameters");
	if (m.discrete_time())
		throw error(MBG_E_LIMIT, "the gradient form is defined for continuous time");
	gc.jparam.assign(2 * (size_t)N, 0xffff);

	std::ostringstream os;
	os << "\n#define MBG_GRAD_NP " << gc.n_params << "\n\n";
	os << "__device__ __forceinline__ void mbg_grad_dwell(const uint32_t* __restrict__ s, const uint32_t* __restrict__ en, "
		  "double dt, double (&T)[MBG_GRAD_NP])\n{\n";
	for (int w = 0; w < m.n_words(); w++)
		os << "\tconst uint32_t sw" << w << " = s[" << w << "], ew" << w << " = en[" << w << "];\n";
	// per parameter and state word: nodes whose up (um) / down (dm) rate it is
	std::vector<std::vector<uint32_t>> um(gc.n_params, std::vector<uint32_t>(m.n_words(), 0u)), dm = um;
	for (int i = 0; i < N; i++)
	{
		const node_t& n = m.nodes[i];
		const node_info& ni = g.nodes[i];
		auto uses = [&](const char* attr) {
			std::set<std::string> v;
			if (const expr* e = n.attr(attr))
				collect_vars(m, n, *e, v);
			return !v.empty();
		};
		const bool up_uses = uses("rate_up"), down_uses = uses("rate_down");
		if (!up_uses && !down_uses)
			continue;
		if (!ni.fast_form || (up_uses && ni.up_param < 0) || (down_uses && ni.down_param < 0))
			throw error(MBG_E_LIMIT, "the gradient form needs node " + n.name
										 + "'s rates to be `@logic ? $u : 0` / `@logic ? 0 : $d` or parameter-free");
		// the node's rate is its up parameter when off and enabled (logic on),
		// its down parameter when on and enabled (logic off)
		const int w = i / 32, b = i % 32;
		// c is 0 or 1, so fma(c, dt, T) is T + dt or T exactly
		os << "\t{ // node " << i << ": " << n.name << "\n";
		os << "\t\tconst uint32_t x = (sw" << w << " >> " << b << "u) & 1u, e = (ew" << w << " >> " << b
		   << "u) & 1u;\n";
		if (ni.up_param >= 0)
		{
			gc.jparam[2 * (size_t)i] = (unsigned short)ni.up_param;
			um[ni.up_param][w] |= 1u << b;
			os << "\t\t{ const double c = (double)(e & (x ^ 1u)); T[" << ni.up_param << "] = fma(c, dt, T["
			   << ni.up_param << "]); }\n";
		}
		if (ni.down_param >= 0)
		{
			gc.jparam[2 * (size_t)i + 1] = (unsigned short)ni.down_param;
			dm[ni.down_param][w] |= 1u << b;
			os << "\t\t{ const double c = (double)(e & x); T[" << ni.down_param << "] = fma(c, dt, T["
			   << ni.down_param << "]); }\n";
		}
		os << "\t}\n";
	}
	os << "}\n\n";
	// c_j = d total / d theta_j in a state, from the words of enabled up (cu)
	// and down (cd) directions: the number of enabled node directions whose
	// rate is theta_j.  The masks are a constant table, so the kernel's
	// unrolled loop over j reads them as immediates.
	const int W = m.n_words();
	for (int d = 0; d < 2; d++)
	{
		os << "static __device__ const uint32_t mbg_grad_" << (d ? "dm" : "um") << "[" << gc.n_params * W << "] = {";
		for (int j = 0; j < gc.n_params; j++)
			for (int w = 0; w < W; w++)
			{
				char buf[24];
				std::snprintf(buf, sizeof(buf), "%s0x%08xu", j + w == 0 ? "" : ", ", d ? dm[j][w] : um[j][w]);
				os << buf;
			}
		os << "};\n";
	}
	os << "__device__ __forceinline__ uint32_t mbg_grad_on(int j, const uint32_t* __restrict__ cu, "
		  "const uint32_t* __restrict__ cd)\n{\n\tuint32_t c = 0u;\n";
	for (int w = 0; w < W; w++)
		os << "\tc += __builtin_popcount(cu[" << w << "] & mbg_grad_um[j * " << W << " + " << w
		   << "]) + __builtin_popcount(cd[" << w << "] & mbg_grad_dm[j * " << W << " + " << w << "]);\n";
	os << "\treturn c;\n}\n";
	gc.source = os.str();
	return gc;
}

} // namespace mbg
