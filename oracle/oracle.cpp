// Host oracle of the MaBoSS simulate/statistics path.
//
// TEST INFRASTRUCTURE ONLY.  This program is the checker the GPU engine is
// compared against (and, timed, the CPU baseline); the engine never calls
// it.  It is a separate restatement of the reference's algorithm, written
// independently of the engine's sources: its own .bnd/.cfg parser, its own
// expression interpreter (optionally compiled to C++ with g++ for speed), its
// own Philox and log2, one trajectory at a time in plain loops.
//
// Reference semantics restated here (file:line in MaBoSS.GPU):
//   parser / evaluate:   src/parser/scanner.ll:20-93, src/parser/parser.yy:71-159,
//                        src/parser/expressions.cpp:10-231, src/parser/driver.cpp:9-99
//   node rates:          src/generator.cpp:84-144 (rate = on ? rate_down : rate_up)
//   transition entropy:  src/generator.cpp:146-175
//   observed index:      src/generator.cpp:207-235
//   initial state:       src/jit_kernels/simulation.cu:34-71
//   SSA step:            src/jit_kernels/simulation.cu:106-163, flip :9-21
//   window statistics:   src/jit_kernels/window_average_small.cu:50-136
//   final states:        src/jit_kernels/final_states.cu:9-21
//   fixed points:        src/statistics/fixed_states.cu:84-125, order src/state.h:43-49
//   normalisation/CSV:   src/statistics/window_average_small.cpp:126-250,
//                        src/statistics/fixed_states.cu:154-187,
//                        src/statistics/final_states.cpp:53-66
// The random stream is the engine's Philox4x32-10 layout (the reference's
// cuRAND XORWOW cannot be reproduced); window sums are unsigned fixed point
// (dwell * 2^32 / tick, round to nearest) so results are exact integers.
//
// Usage: maboss_oracle [options] model.bnd model.cfg
//   --n N            trajectories (default: sample_count)
//   --first-id F     id of the first trajectory (default 0)
//   --seed S         seed (default: seed_pseudorandom)
//   --variant V      stream tag (default 0)
//   --threads T      OpenMP threads (default: all)
//   --compiled       compile the formulas with g++ instead of interpreting
//   --entropy E      log_sum (default) or reference
//   --order O        seq (default: lane kernel order) or wave (wave kernel order)
//   --max-steps K    step cap per trajectory (default 2^30)
//   --raw PATH       exact integer dump (same format as the engine's)
//   -o PREFIX        write PREFIX_probtraj.csv and PREFIX_fp.csv
//   --print PATH     write the stdout report to PATH
//   --errors         exact per-trajectory second moments: R / Q lines in the
//                    raw dump, ErrorTH / ErrorProba in the CSV (0 without)
//   --time           print a JSON timing line to stderr
//   --eval-states    read lines "w0 .. w{W-1} u2bits" on stdin, print
//                    "total H flip obs" (float bits in hex) per line
//   --grad-in PATH   likelihood-ratio gradient mode: PATH holds the upstream
//                    gradient dL/dP as windows * 2^n_obs doubles; prints
//                    "transitions N" and "name dL/dtheta" per parameter
//   --grad-out PATH  where to write those lines (default stdout)
// maboss_oracle --numeric-probe reads "P c0 c1 c2 c3 k0 k1" (Philox),
// "L bits" (log2), "U x" (uniform) lines on stdin and prints the results.
#include <algorithm>
#include <cerrno>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <dlfcn.h>
#include <fstream>
#include <iostream>
#include <map>
#include <memory>
#include <set>
#include <sstream>
#include <string>
#include <unistd.h>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace orc
{

using u128 = unsigned __int128;

// Standard error of a mean over n trajectories from exact sums: s1 = sum of
// the per-trajectory values, s2 = sum of their squares, both in fixed-point
// units of `unit`: sqrt((E[x^2] - E[x]^2) / (n - 1)), 0 for n < 2 or a
// negative rounding of the variance.
double std_error(uint64_t s1, u128 s2, double n, double unit)
{
	if (n < 2.0)
		return 0.0;
	const double mean = (double)s1 / (n * unit);
	const double m2 = ((double)(uint64_t)(s2 >> 64) * 18446744073709551616.0 + (double)(uint64_t)s2) / (n * unit * unit);
	const double var = (m2 - mean * mean) / (n - 1.0);
	return var > 0.0 ? std::sqrt(var) : 0.0;
}

[[noreturn]] void die(const std::string& m)
{
	std::fprintf(stderr, "oracle: %s\n", m.c_str());
	std::exit(2);
}

// ----------------------------------------------------------------- numerics

// Philox4x32-10 (Salmon et al., SC'11).
struct u4
{
	uint32_t v[4];
};

u4 philox(u4 c, uint32_t k0, uint32_t k1)
{
	for (int r = 0; r < 10; r++)
	{
		uint64_t p0 = (uint64_t)0xD2511F53u * c.v[0];
		uint64_t p1 = (uint64_t)0xCD9E8D57u * c.v[2];
		u4 n;
		n.v[0] = (uint32_t)(p1 >> 32) ^ c.v[1] ^ k0;
		n.v[1] = (uint32_t)p1;
		n.v[2] = (uint32_t)(p0 >> 32) ^ c.v[3] ^ k1;
		n.v[3] = (uint32_t)p0;
		c = n;
		k0 += 0x9E3779B9u;
		k1 += 0xBB67AE85u;
	}
	return c;
}

inline float u01(uint32_t x) { return (float)((x >> 8) + 1u) * 5.9604644775390625e-08f; }

inline uint32_t fbits(float f)
{
	uint32_t u;
	std::memcpy(&u, &f, 4);
	return u;
}
inline float bitsf(uint32_t u)
{
	float f;
	std::memcpy(&f, &u, 4);
	return f;
}

// log2 by exponent split + degree-9 polynomial in f = m - 1, m in
// [sqrt(1/2), sqrt(2)); coefficients of a Chebyshev fit of log2(1+f)/f.
float log2_poly(float x)
{
	uint32_t b = fbits(x);
	if (b >= 0x7f800000u)
		return x;
	int e = 0;
	if (b < 0x00800000u)
	{
		if (b == 0)
			return -INFINITY;
		b = fbits(x * 8388608.0f);
		e = -23;
	}
	e += (int)(b >> 23) - 127;
	float m = bitsf((b & 0x007fffffu) | 0x3f800000u);
	if (m > 1.41421354f)
	{
		m = m * 0.5f;
		e += 1;
	}
	const float f = m - 1.0f;
	static const float C[10] = { 0x1.715476p+0f,  -0x1.71547p-1f, 0x1.ec7094p-2f,  -0x1.715a76p-2f, 0x1.277e9ap-2f,
								 -0x1.eabd64p-3f, 0x1.a2f86p-3f,  -0x1.87381cp-3f, 0x1.7d9132p-3f,  -0x1.c362bcp-4f };
	float p = C[9];
	for (int i = 8; i >= 0; i--)
		p = p * f + C[i];
	return (float)e + f * p;
}

inline float ln_poly(float x) { return log2_poly(x) * 0.693147182f; }

// ------------------------------------------------------------------- parser

enum K
{
	LIT,
	NODE,
	VAR,
	ALIAS,
	PAREN,
	NEG,
	POS,
	NOT,
	ADD,
	SUB,
	MUL,
	DIV,
	AND,
	OR,
	XOR,
	EQ,
	NE,
	LT,
	LE,
	GT,
	GE,
	TERN
};

struct Ex
{
	K k;
	float val = 0;
	double exact = 0;
	std::string name;
	int idx = -1; // resolved node index
	std::shared_ptr<Ex> a, b, c;
};
using ExP = std::shared_ptr<Ex>;

struct Node
{
	std::string name;
	std::vector<std::pair<std::string, ExP>> attrs;
	float istate = 0.5f;
	ExP attr(const std::string& n) const
	{
		for (auto& a : attrs)
			if (a.first == n)
				return a.second;
		return nullptr;
	}
};

struct Model
{
	std::vector<Node> nodes;
	std::map<std::string, float> vars;
	std::map<std::string, float> consts { { "discrete_time", 0.f },
										  { "sample_count", 1e6f },
										  { "max_time", 10.f },
										  { "time_tick", 0.1f },
										  { "seed_pseudorandom", 0.f } };
	std::map<std::string, double> consts_exact { { "sample_count", 1e6 }, { "seed_pseudorandom", 0 } };
	// `X.mutant = ON | OFF | 1 | 0;` of the .cfg, node index -> +1 on / -1 off
	std::map<int, int> mutants;
	int find(const std::string& n) const
	{
		for (size_t i = 0; i < nodes.size(); i++)
			if (nodes[i].name == n)
				return (int)i;
		return -1;
	}
};

struct Tok
{
	int t = 0; // 0 end, 'n' number, 'f' float, 'i' ident, '$' var, '@' alias, or operator code
	std::string s;
	double num = 0;
};

// operator codes beyond single characters
enum
{
	T_AND = 256,
	T_OR,
	T_XOR,
	T_NOT,
	T_EQ,
	T_NE,
	T_LE,
	T_GE
};

std::vector<Tok> lex(const std::string& src, const std::string& fn)
{
	std::vector<Tok> out;
	size_t i = 0, n = src.size();
	auto is_id0 = [](char c) { return std::isalpha((unsigned char)c); };
	auto is_id = [](char c) { return std::isalnum((unsigned char)c) || c == '_'; };
	while (i < n)
	{
		char c = src[i];
		if (std::isspace((unsigned char)c))
		{
			i++;
			continue;
		}
		if (c == '/' && i + 1 < n && src[i + 1] == '/')
		{
			while (i < n && src[i] != '\n')
				i++;
			continue;
		}
		if (c == '/' && i + 1 < n && src[i + 1] == '*')
		{
			size_t e = src.find("*/", i + 2);
			if (e == std::string::npos)
				die(fn + ": unterminated comment");
			i = e + 2;
			continue;
		}
		Tok t;
		std::string two = src.substr(i, 2);
		if (two == "&&" || two == "||" || two == "==" || two == "!=" || two == "<=" || two == ">=")
		{
			t.t = two == "&&" ? T_AND : two == "||" ? T_OR : two == "==" ? T_EQ : two == "!=" ? T_NE : two == "<=" ? T_LE : T_GE;
			i += 2;
			out.push_back(t);
			continue;
		}
		if (std::strchr("-+*/()=?:{};[],<>", c))
		{
			t.t = c;
			i++;
			out.push_back(t);
			continue;
		}
		if (c == '&' || c == '|' || c == '^' || c == '!')
		{
			t.t = c == '&' ? T_AND : c == '|' ? T_OR : c == '^' ? T_XOR : T_NOT;
			i++;
			out.push_back(t);
			continue;
		}
		if (std::isdigit((unsigned char)c) || (c == '.' && i + 1 < n && std::isdigit((unsigned char)src[i + 1])))
		{
			size_t j = i;
			bool fl = false;
			while (j < n && std::isdigit((unsigned char)src[j]))
				j++;
			if (j < n && src[j] == '.')
			{
				fl = true;
				j++;
				while (j < n && std::isdigit((unsigned char)src[j]))
					j++;
			}
			if (j < n && (src[j] == 'e' || src[j] == 'E'))
			{
				size_t k = j + 1;
				if (k < n && (src[k] == '+' || src[k] == '-'))
					k++;
				if (k < n && std::isdigit((unsigned char)src[k]))
				{
					while (k < n && std::isdigit((unsigned char)src[k]))
						k++;
					j = k;
					fl = true;
				}
			}
			t.s = src.substr(i, j - i);
			t.t = fl ? 'f' : 'n';
			t.num = std::strtod(t.s.c_str(), nullptr);
			i = j;
			out.push_back(t);
			continue;
		}
		if (c == '.')
		{
			t.t = '.';
			i++;
			out.push_back(t);
			continue;
		}
		if (c == '$' || c == '@')
		{
			size_t j = i + 1;
			if (j >= n || !is_id0(src[j]))
				die(fn + ": invalid character " + std::string(1, c));
			while (j < n && is_id(src[j]))
				j++;
			t.t = c;
			t.s = src.substr(i, j - i);
			i = j;
			out.push_back(t);
			continue;
		}
		if (is_id0(c))
		{
			size_t j = i;
			while (j < n && is_id(src[j]))
				j++;
			t.s = src.substr(i, j - i);
			std::string l = t.s;
			for (auto& ch : l)
				ch = (char)std::tolower((unsigned char)ch);
			if (l == "and")
				t.t = T_AND;
			else if (l == "or")
				t.t = T_OR;
			else if (l == "xor")
				t.t = T_XOR;
			else if (l == "not")
				t.t = T_NOT;
			else if (l == "true" || l == "false")
			{
				t.t = 'n';
				t.num = l == "true" ? 1 : 0;
			}
			else
				t.t = 'i';
			i = j;
			out.push_back(t);
			continue;
		}
		if ((unsigned char)c == 0xEF)
		{
			i += 3;
			continue;
		}
		die(fn + ": invalid character " + std::string(1, c));
	}
	out.push_back(Tok {});
	return out;
}

struct Parser
{
	std::vector<Tok> t;
	size_t p = 0;
	std::string fn;
	Model& m;
	Parser(const std::string& src, const std::string& f, Model& mm) : t(lex(src, f)), fn(f), m(mm) {}

	const Tok& cur() const { return t[p]; }
	Tok take() { return t[p++]; }
	Tok need(int k)
	{
		if (cur().t != k)
			die(fn + ": syntax error near token " + std::to_string(p));
		return take();
	}

	// bison precedence (parser.yy:126-134): ?: < XOR < OR < AND < == != < rel < + - < * / < unary
	static int prec(int k)
	{
		switch (k)
		{
			case '?': return 1;
			case T_XOR: return 2;
			case T_OR: return 3;
			case T_AND: return 4;
			case T_EQ:
			case T_NE: return 5;
			case '<':
			case '>':
			case T_LE:
			case T_GE: return 6;
			case '+':
			case '-': return 7;
			case '*':
			case '/': return 8;
		}
		return 0;
	}

	ExP mk(K k)
	{
		auto e = std::make_shared<Ex>();
		e->k = k;
		return e;
	}

	ExP unary()
	{
		Tok x = take();
		ExP e;
		switch (x.t)
		{
			case 'n':
			case 'f':
				e = mk(LIT);
				e->val = (float)x.num;
				e->exact = x.num;
				return e;
			case 'i':
				e = mk(NODE);
				e->name = x.s;
				return e;
			case '$':
				e = mk(VAR);
				e->name = x.s;
				return e;
			case '@':
				e = mk(ALIAS);
				e->name = x.s.substr(1);
				return e;
			case '(':
				e = mk(PAREN);
				e->a = expr(0);
				need(')');
				return e;
			case '-':
			case '+':
			case T_NOT:
				e = mk(x.t == '-' ? NEG : x.t == '+' ? POS : NOT);
				e->a = expr(9);
				return e;
		}
		die(fn + ": syntax error (unexpected token)");
	}

	ExP expr(int minp)
	{
		ExP l = unary();
		for (;;)
		{
			int k = cur().t, pr = prec(k);
			if (!pr || pr <= minp)
				return l;
			take();
			if (k == '?')
			{
				ExP e = mk(TERN);
				e->a = l;
				e->b = expr(0);
				need(':');
				e->c = expr(pr - 1);
				l = e;
				continue;
			}
			K kk = k == T_XOR ? XOR : k == T_OR ? OR : k == T_AND ? AND : k == T_EQ ? EQ : k == T_NE ? NE
				: k == '<' ? LT : k == '>' ? GT : k == T_LE ? LE : k == T_GE ? GE : k == '+' ? ADD
				: k == '-' ? SUB : k == '*' ? MUL : DIV;
			ExP e = mk(kk);
			e->a = l;
			e->b = expr(pr);
			l = e;
		}
	}

	float eval(const ExP& e);

	void bnd()
	{
		while (cur().t)
		{
			Tok kw = need('i');
			std::string l = kw.s;
			for (auto& ch : l)
				ch = (char)std::tolower((unsigned char)ch);
			if (l != "node")
				die(fn + ": expected 'node' keyword");
			Node nd;
			nd.name = need('i').s;
			need('{');
			do
			{
				std::string an = need('i').s;
				need('=');
				ExP e = expr(0);
				need(';');
				nd.attrs.emplace_back(an, e);
			} while (cur().t != '}');
			need('}');
			if (m.find(nd.name) >= 0)
				die("Node " + nd.name + " already exists");
			if (!nd.attr("rate_up") || !nd.attr("rate_down"))
				die("Node " + nd.name + " lacks rate_up or rate_down");
			m.nodes.push_back(nd);
		}
	}

	void cfg()
	{
		while (cur().t)
		{
			if (cur().t == '$')
			{
				std::string v = take().s;
				need('=');
				ExP e = expr(0);
				need(';');
				m.vars[v] = eval(e);
			}
			else if (cur().t == '[')
			{
				take();
				std::string nn = need('i').s;
				need(']');
				need('.');
				if (need('i').s != "istate")
					die("expected istate");
				need('=');
				ExP l = expr(0);
				need('[');
				double vl = need('n').num;
				need(']');
				need(',');
				ExP r = expr(0);
				need('[');
				double vr = need('n').num;
				need(']');
				need(';');
				if (!((vl == 0 && vr == 1) || (vl == 1 && vr == 0)))
					die("numbers in [] must be 0 and 1");
				int i = m.find(nn);
				if (i < 0)
					die("Unknown node " + nn);
				float lv = eval(l), rv = eval(r);
				m.nodes[i].istate = vl == 1 ? lv / (lv + rv) : rv / (lv + rv);
			}
			else
			{
				std::string id = need('i').s;
				if (cur().t == '.')
				{
					take();
					std::string an = need('i').s;
					need('=');
					ExP e = expr(0);
					need(';');
					int i = m.find(id);
					if (i < 0)
						die("Unknown node " + id);
					if (an == "istate")
						m.nodes[i].istate = eval(e);
					else if (an == "mutant")
					{
						std::string v = e->k == NODE ? e->name : "";
						for (char& ch : v)
							ch = (char)toupper((unsigned char)ch);
						const float f = e->k == NODE ? -1.0f : eval(e);
						if (v == "ON" || f == 1.0f)
							m.mutants[i] = 1;
						else if (v == "OFF" || f == 0.0f)
							m.mutants[i] = -1;
						else
							die("mutant must be ON or OFF");
					}
					else
						m.nodes[i].attrs.emplace_back(an, e);
				}
				else
				{
					need('=');
					ExP e = expr(0);
					need(';');
					m.consts[id] = eval(e);
					const Ex* x = e.get();
					double sign = 1;
					while (x->k == PAREN || x->k == POS || x->k == NEG)
					{
						if (x->k == NEG)
							sign = -sign;
						x = x->a.get();
					}
					m.consts_exact[id] = x->k == LIT ? sign * x->exact : (double)m.consts[id];
				}
			}
		}
	}
};

// Constant evaluation with float semantics (expressions.cpp:10-81).
float Parser::eval(const ExP& e)
{
	switch (e->k)
	{
		case LIT: return e->val;
		case VAR: {
			auto it = m.vars.find(e->name);
			if (it == m.vars.end())
				die("undefined variable " + e->name);
			return it->second;
		}
		case NODE:
		case ALIAS: die("node or alias in a constant expression");
		case PAREN:
		case POS: return eval(e->a);
		case NEG: return -eval(e->a);
		case NOT: return !eval(e->a);
		case ADD: return eval(e->a) + eval(e->b);
		case SUB: return eval(e->a) - eval(e->b);
		case MUL: return eval(e->a) * eval(e->b);
		case DIV: return eval(e->a) / eval(e->b);
		case AND: return eval(e->a) && eval(e->b);
		case OR: return eval(e->a) || eval(e->b);
		case XOR: return (eval(e->a) != 0) != (eval(e->b) != 0);
		case EQ: return eval(e->a) == eval(e->b);
		case NE: return eval(e->a) != eval(e->b);
		case LT: return eval(e->a) < eval(e->b);
		case LE: return eval(e->a) <= eval(e->b);
		case GT: return eval(e->a) > eval(e->b);
		case GE: return eval(e->a) >= eval(e->b);
		case TERN: return eval(e->a) ? eval(e->b) : eval(e->c);
	}
	return 0;
}

std::string slurp(const std::string& p)
{
	std::ifstream f(p, std::ios::binary);
	if (!f)
		die("cannot open " + p);
	std::stringstream s;
	s << f.rdbuf();
	return s.str();
}

// --------------------------------------------------------------- evaluation

// Resolved rate program of one node: the rate_up and rate_down trees with
// aliases inlined and node names turned into indices.
ExP resolve(const Model& m, const Node& nd, const ExP& e, int depth)
{
	if (depth > 64)
		die("alias nesting too deep in " + nd.name);
	auto c = std::make_shared<Ex>(*e);
	if (e->k == ALIAS)
	{
		ExP t = nd.attr(e->name);
		if (!t)
			die("Node " + nd.name + " does not have attribute " + e->name);
		auto p = std::make_shared<Ex>();
		p->k = PAREN;
		p->a = resolve(m, nd, t, depth + 1);
		return p;
	}
	if (e->k == NODE)
	{
		c->idx = m.find(e->name);
		if (c->idx < 0)
			die("unknown node name: " + e->name);
	}
	if (e->k == VAR)
	{
		auto it = m.vars.find(e->name);
		if (it == m.vars.end())
			die("undefined variable " + e->name);
		c->val = it->second;
	}
	if (e->a)
		c->a = resolve(m, nd, e->a, depth);
	if (e->b)
		c->b = resolve(m, nd, e->b, depth);
	if (e->c)
		c->c = resolve(m, nd, e->c, depth);
	return c;
}

float run(const Ex* e, const uint32_t* s)
{
	switch (e->k)
	{
		case LIT:
		case VAR: return e->val;
		case NODE: return (float)((s[e->idx >> 5] >> (e->idx & 31)) & 1u);
		case ALIAS: return 0; // resolved away
		case PAREN:
		case POS: return run(e->a.get(), s);
		case NEG: return -run(e->a.get(), s);
		case NOT: return !run(e->a.get(), s);
		case ADD: return run(e->a.get(), s) + run(e->b.get(), s);
		case SUB: return run(e->a.get(), s) - run(e->b.get(), s);
		case MUL: return run(e->a.get(), s) * run(e->b.get(), s);
		case DIV: return run(e->a.get(), s) / run(e->b.get(), s);
		case AND: return run(e->a.get(), s) && run(e->b.get(), s);
		case OR: return run(e->a.get(), s) || run(e->b.get(), s);
		case XOR: return (run(e->a.get(), s) != 0) != (run(e->b.get(), s) != 0);
		case EQ: return run(e->a.get(), s) == run(e->b.get(), s);
		case NE: return run(e->a.get(), s) != run(e->b.get(), s);
		case LT: return run(e->a.get(), s) < run(e->b.get(), s);
		case LE: return run(e->a.get(), s) <= run(e->b.get(), s);
		case GT: return run(e->a.get(), s) > run(e->b.get(), s);
		case GE: return run(e->a.get(), s) >= run(e->b.get(), s);
		case TERN: return run(e->a.get(), s) ? run(e->b.get(), s) : run(e->c.get(), s);
	}
	return 0;
}

// C++ text of a resolved tree with the same float semantics as run().
std::string emit(const Ex* e)
{
	auto f = [](float v) {
		char b[64];
		std::snprintf(b, sizeof b, "%af", (double)v);
		return "(" + std::string(b) + ")";
	};
	auto A = [&] { return emit(e->a.get()); };
	auto B = [&] { return emit(e->b.get()); };
	auto bin = [&](const char* o) { return "(" + A() + o + B() + ")"; };
	auto cmp = [&](const char* o) { return "((" + A() + o + B() + ") ? 1.0f : 0.0f)"; };
	switch (e->k)
	{
		case LIT:
		case VAR: return f(e->val);
		case NODE:
			return "((float)((s[" + std::to_string(e->idx >> 5) + "] >> " + std::to_string(e->idx & 31) + ") & 1u))";
		case ALIAS: return "0.0f";
		case PAREN:
		case POS: return A();
		case NEG: return "(-" + A() + ")";
		case NOT: return "((" + A() + " != 0.0f) ? 0.0f : 1.0f)";
		case ADD: return bin(" + ");
		case SUB: return bin(" - ");
		case MUL: return bin(" * ");
		case DIV: return bin(" / ");
		case AND: return "(((" + A() + " != 0.0f) && (" + B() + " != 0.0f)) ? 1.0f : 0.0f)";
		case OR: return "(((" + A() + " != 0.0f) || (" + B() + " != 0.0f)) ? 1.0f : 0.0f)";
		case XOR: return "(((" + A() + " != 0.0f) != (" + B() + " != 0.0f)) ? 1.0f : 0.0f)";
		case EQ: return cmp(" == ");
		case NE: return cmp(" != ");
		case LT: return cmp(" < ");
		case LE: return cmp(" <= ");
		case GT: return cmp(" > ");
		case GE: return cmp(" >= ");
		case TERN: return "((" + A() + " != 0.0f) ? " + B() + " : " + emit(e->c.get()) + ")";
	}
	return "0.0f";
}

typedef void (*rates_fn)(const uint32_t* s, float* r);

struct Sim
{
	int N = 0, W = 0, nobs = 0, windows = 0;
	bool discrete = false;
	float max_time = 0, tick = 0;
	std::vector<float> istate;
	std::vector<int> obs;
	std::vector<ExP> up, down;
	rates_fn compiled = nullptr;
	bool reference_entropy = false;
	// Wave order: the summation order of the wave-per-trajectory kernel.
	// Node i belongs to lane i / npl (64 lanes, npl = ceil(N / 64) nodes
	// each); lane sums are sequential over the lane's slice and combined by
	// the inclusive scan of lane_scan.
	bool wave = false;
	int npl = 1;
	uint32_t seed_lo = 0, seed_hi = 0, variant = 0, max_steps = 1u << 30;

	// The wave kernel's scan over 64 lanes in four rows of 16: a Kogge-Stone
	// scan inside each row (offsets 1, 2, 4, 8; lanes with no partner add 0),
	// then rows 1 and 3 add the last lane of the row before, then rows 2 and 3
	// add lane 31 (every other lane adds 0 in those two steps).
	static void lane_scan(float* v)
	{
		float nv[64];
		for (int off = 1; off < 16; off <<= 1)
		{
			for (int l = 0; l < 64; l++)
				nv[l] = v[l] + ((l % 16) >= off ? v[l - off] : 0.0f);
			std::memcpy(v, nv, sizeof(nv));
		}
		for (int l = 0; l < 64; l++)
			nv[l] = v[l] + ((l / 16) % 2 == 1 ? v[16 * (l / 16) - 1] : 0.0f);
		std::memcpy(v, nv, sizeof(nv));
		for (int l = 0; l < 64; l++)
			nv[l] = v[l] + (l >= 32 ? v[31] : 0.0f);
		std::memcpy(v, nv, sizeof(nv));
	}

	// Total rate; in wave order also the lane sums L and scan P.
	float total_of(const float* r, float* L, float* P) const
	{
		if (!wave)
		{
			float total = 0.0f;
			for (int i = 0; i < N; i++)
				total = total + r[i];
			return total;
		}
		for (int l = 0; l < 64; l++)
		{
			float acc = 0.0f;
			for (int k = 0; k < npl; k++)
			{
				const int i = l * npl + k;
				acc = acc + (i < N ? r[i] : 0.0f);
			}
			L[l] = P[l] = acc;
		}
		lane_scan(P);
		return P[63];
	}

	// Flip selection for rr = u2 * total.
	int select_flip(const float* r, float total, float rr, const float* L, const float* P) const
	{
		if (!wave)
		{
			// number of prefix sums strictly below rr (simulation.cu:9-21)
			float acc = 0.0f;
			int flip = 0;
			for (int i = 0; i < N; i++)
			{
				acc = acc + r[i];
				flip += (acc < rr) ? 1 : 0;
			}
			return flip;
		}
		(void)total;
		// first lane whose scan reaches rr and whose slice has weight; else
		// the last lane with weight
		int ls = -1;
		for (int l = 0; l < 64 && ls < 0; l++)
			if (P[l] >= rr && L[l] > 0.0f)
				ls = l;
		if (ls < 0)
			for (int l = 63; l >= 0 && ls < 0; l--)
				if (L[l] > 0.0f)
					ls = l;
		const float rem = rr - (ls == 0 ? 0.0f : P[ls - 1]);
		// first slot whose running sum reaches rem with a nonzero rate; else
		// the last slot with a nonzero rate
		float q = 0.0f;
		int ks = -1, last = -1;
		for (int k = 0; k < npl; k++)
		{
			const int i = ls * npl + k;
			const float v = i < N ? r[i] : 0.0f;
			q = q + v;
			if (v > 0.0f)
			{
				last = k;
				if (ks < 0 && q >= rem)
					ks = k;
			}
		}
		return ls * npl + (ks >= 0 ? ks : last);
	}

	void rates(const uint32_t* s, float* r) const
	{
		if (compiled)
		{
			compiled(s, r);
			return;
		}
		for (int i = 0; i < N; i++)
		{
			const bool on = (s[i >> 5] >> (i & 31)) & 1u;
			r[i] = run(on ? down[i].get() : up[i].get(), s);
		}
	}

	uint64_t obs_index(const uint32_t* s) const
	{
		uint64_t v = 0;
		for (int j = 0; j < nobs; j++)
			v |= (uint64_t)((s[obs[j] >> 5] >> (obs[j] & 31)) & 1u) << j;
		return v;
	}

	float entropy(const float* r) const
	{
		if (nobs == 0)
			return 0.0f;
		if (reference_entropy)
		{
			// src/generator.cpp:146-175
			float S = 0.0f;
			for (int i : obs)
				S = S + r[i];
			float H = 0.0f;
			if (S != 0.0f)
				for (int i : obs)
				{
					const float p = r[i] / S;
					H = H - ((p == 0.0f) ? 0.0f : log2_poly(p) * p);
				}
			return H;
		}
		// H = log2(S) - T/S, T = sum r log2 r; 0 if at most one rate is nonzero
		float S = 0.0f, T = 0.0f;
		int nnz = 0;
		if (wave)
		{
			// observed node k contributes to lane k % 64 (sequentially), then
			// lane_scan of the lane values
			float Sl[64], Tl[64];
			for (int l = 0; l < 64; l++)
			{
				Sl[l] = 0.0f;
				Tl[l] = 0.0f;
				for (int k = l; k < nobs; k += 64)
				{
					const float v = r[obs[k]];
					Sl[l] = Sl[l] + v;
					Tl[l] = Tl[l] + (v == 0.0f ? 0.0f : v * log2_poly(v));
					nnz += v != 0.0f;
				}
			}
			lane_scan(Sl);
			lane_scan(Tl);
			S = Sl[63];
			T = Tl[63];
			const float h = log2_poly(S) - T / S;
			return (nnz > 1 && h > 0.0f) ? h : 0.0f;
		}
		for (int i : obs)
		{
			S = S + r[i];
			T = T + (r[i] == 0.0f ? 0.0f : r[i] * log2_poly(r[i]));
			nnz += r[i] != 0.0f;
		}
		const float h = log2_poly(S) - T / S;
		return (nnz > 1 && h > 0.0f) ? h : 0.0f;
	}
};

// Open-addressing map from 64-bit cell keys to counts (linear probing,
// Fibonacci hashing, grown at half load); the key ~0 marks an empty slot (no
// cell key reaches it: window and observed bits fit in 63 bits).
struct CellMap
{
	std::vector<uint64_t> keys, vals;
	size_t used = 0;
	int shift = 64;

	static constexpr uint64_t k_empty = ~0ull;
	void rehash(size_t cap)
	{
		std::vector<uint64_t> ok, ov;
		ok.swap(keys);
		ov.swap(vals);
		keys.assign(cap, k_empty);
		vals.assign(cap, 0);
		shift = 64 - __builtin_ctzll(cap);
		used = 0;
		for (size_t i = 0; i < ok.size(); i++)
			if (ok[i] != k_empty)
				add(ok[i], ov[i]);
	}
	void add(uint64_t key, uint64_t q)
	{
		if (2 * (used + 1) > keys.size())
			rehash(keys.empty() ? 1024 : 2 * keys.size());
		const size_t mask = keys.size() - 1;
		size_t i = (size_t)((key * 0x9E3779B97F4A7C15ull) >> shift);
		while (keys[i] != k_empty && keys[i] != key)
			i = (i + 1) & mask;
		if (keys[i] == k_empty)
		{
			keys[i] = key;
			used++;
		}
		vals[i] += q;
	}
	template <typename F>
	void each(F&& f) const
	{
		for (size_t i = 0; i < keys.size(); i++)
			if (keys[i] != k_empty)
				f(keys[i], vals[i]);
	}
};

struct Stats
{
	int nobs = 0, windows = 0;
	bool dense = false;
	std::vector<uint64_t> dwin;                          // dense (w << nobs) | obs
	CellMap hwin;                                        // sparse
	std::vector<uint64_t> ent;
	CellMap fin;
	std::map<std::vector<uint32_t>, uint64_t> fixedp;   // key: words reversed (MSW first)
	uint64_t transitions = 0, finished = 0, truncated = 0;
	// --errors: per-trajectory second moments, exact.  sq[(w << nobs) | obs]
	// = sum over trajectories of P^2, P the trajectory's fixed-point time in
	// obs during window w; esq[w] = sum of T^2, T its entropy sum in window w.
	bool moments = false;
	std::unordered_map<uint64_t, u128> sq;
	std::vector<u128> esq;

	void init(int nobs_, int windows_)
	{
		nobs = nobs_;
		windows = windows_;
		dense = nobs <= 40 && ((uint64_t)windows << nobs) <= (1ull << 20);
		if (dense)
			dwin.assign((size_t)windows << nobs, 0);
		ent.assign(windows, 0);
		esq.assign(windows, 0);
	}
	void add(int w, uint64_t o, uint64_t q)
	{
		const uint64_t key = ((uint64_t)w << nobs) | o;
		if (dense)
			dwin[key] += q;
		else
			hwin.add(key, q);
	}
	void merge(const Stats& o)
	{
		if (dense)
			for (size_t i = 0; i < dwin.size(); i++)
				dwin[i] += o.dwin[i];
		else
			o.hwin.each([&](uint64_t k, uint64_t v) { hwin.add(k, v); });
		for (int w = 0; w < windows; w++)
			ent[w] += o.ent[w];
		o.fin.each([&](uint64_t k, uint64_t v) { fin.add(k, v); });
		for (auto& kv : o.fixedp)
			fixedp[kv.first] += kv.second;
		transitions += o.transitions;
		finished += o.finished;
		truncated += o.truncated;
		for (auto& kv : o.sq)
			sq[kv.first] += kv.second;
		for (int w = 0; w < windows; w++)
			esq[w] += o.esq[w];
	}
};

// One trajectory's contributions to the window it is in: its time in each
// observed state and its entropy sum, squared into Stats::sq / esq when the
// trajectory leaves the window (or ends).  A state visited twice in one
// window is one entry: the square is of the trajectory's total.
struct WindowMoments
{
	int w = -1;
	std::vector<std::pair<uint64_t, uint64_t>> cells; // (obs, P)
	uint64_t T = 0;

	void at(int w_, Stats& st)
	{
		if (w_ != w)
			flush(st);
		w = w_;
	}
	void add(uint64_t o, uint64_t q)
	{
		for (auto& c : cells)
			if (c.first == o)
			{
				c.second += q;
				return;
			}
		cells.emplace_back(o, q);
	}
	void flush(Stats& st)
	{
		if (w >= 0)
		{
			for (auto& c : cells)
				st.sq[((uint64_t)w << st.nobs) | c.first] += (u128)c.second * c.second;
			st.esq[w] += (u128)T * T;
		}
		cells.clear();
		T = 0;
		w = -1;
	}
};

// One trajectory, reference SSA loop (simulation.cu:106-163) with fused
// window accumulation (window_average_small.cu:72-91 / :123-129).
void trajectory(const Sim& S, uint64_t id, Stats& st, std::vector<float>& r, std::vector<uint32_t>& s)
{
	WindowMoments wm;
	const uint32_t idlo = (uint32_t)id, idhi = (uint32_t)(id >> 32);
	std::fill(s.begin(), s.end(), 0u);
	for (int i = 0; i < S.N; i += 4)
	{
		u4 x = philox(u4 { { (uint32_t)(i >> 2), idlo, idhi, S.variant } }, S.seed_lo, S.seed_hi);
		for (int j = 0; j < 4 && i + j < S.N; j++)
			if (u01(x.v[j]) <= S.istate[i + j])
				s[(i + j) >> 5] |= 1u << ((i + j) & 31);
	}
	const uint32_t B0 = ((uint32_t)S.N + 3u) >> 2;
	const float qscale = S.discrete ? 4294967296.0f : 4294967296.0f / S.tick;
	float t = 0.0f;
	for (uint32_t k = 0;; )
	{
		float u1 = 1.0f, u2;
		if (S.discrete)
		{
			u4 x = philox(u4 { { B0 + (k >> 2), idlo, idhi, S.variant } }, S.seed_lo, S.seed_hi);
			u2 = u01(x.v[k & 3]);
		}
		else
		{
			u4 x = philox(u4 { { B0 + (k >> 1), idlo, idhi, S.variant } }, S.seed_lo, S.seed_hi);
			u1 = u01(x.v[2 * (k & 1)]);
			u2 = u01(x.v[2 * (k & 1) + 1]);
		}
		S.rates(s.data(), r.data());
		float L[64], P[64];
		const float total = S.total_of(r.data(), L, P);
		const float tp = t;
		bool fixed = false;
		float H = 0.0f;
		if (total == 0.0f)
		{
			fixed = true;
			t = S.max_time;
		}
		else
		{
			if (S.discrete)
			{
				t = t + S.tick;
				t = (S.max_time - S.tick < t) ? S.max_time : t;
			}
			else
				t = t + (-ln_poly(u1)) / total;
			t = std::fmin(t, S.max_time);
			H = S.entropy(r.data());
		}
		st.transitions++;
		const uint64_t o = S.obs_index(s.data());
		if (S.discrete)
		{
			const int w = (int)std::round(tp / S.tick);
			if (w < S.windows)
			{
				st.add(w, o, 1);
				const uint64_t qe = (uint64_t)std::rint(H * qscale);
				st.ent[w] += qe;
				if (st.moments)
				{
					wm.at(w, st);
					wm.add(o, 1);
					wm.T += qe;
				}
			}
		}
		else
		{
			int w = (int)std::floor(tp / S.tick);
			float b = tp;
			while (t > b)
			{
				const float wend = (float)(w + 1) * S.tick;
				const float en = std::fmin(t, wend);
				const float sl = en - b;
				if (w < S.windows)
				{
					const uint64_t q = (uint64_t)std::rint(sl * qscale);
					if (q)
						st.add(w, o, q);
					const uint64_t qe = (uint64_t)std::rint((H * sl) * qscale);
					st.ent[w] += qe;
					if (st.moments)
					{
						wm.at(w, st);
						if (q)
							wm.add(o, q);
						wm.T += qe;
					}
				}
				w++;
				b = en;
			}
		}
		k++;
		if (t >= S.max_time)
		{
			st.finished++;
			st.fin.add(o, 1);
			if (fixed)
			{
				std::vector<uint32_t> key(s.rbegin(), s.rend());
				st.fixedp[key]++;
			}
			wm.flush(st);
			return;
		}
		if (k >= S.max_steps)
		{
			st.truncated++;
			wm.flush(st);
			return;
		}
		const int flip = S.select_flip(r.data(), total, u2 * total, L, P);
		s[flip >> 5] ^= 1u << (flip & 31);
	}
}

// ------------------------------------------------------- likelihood-ratio gradient
//
// No counterpart in the reference.  Window probability P_w(x) = E[(1/tick)
// * time in window w with observed state x]; for a loss with upstream
// gradient G[w][x] = dL/dP_w(x), the score-function estimator is
//   dL/dtheta_j = 1/(n tick) * sum_traj sum_w Gw * S_j(e_w),
//   Gw = sum over the dwell slices of window w of G[w][obs] * slice,
//   S_j(t) = sum_{jumps k, t_k <= t} (d r_{f_k} / d theta_j) / r_{f_k}
//            - integral_0^t sum_i d r_i / d theta_j du,
// e_w the end of window w (or the end of the trajectory).  Computed here
// directly, in double precision, with derivatives of the rate expressions
// taken by differentiating the expression trees.

// d e / d theta for the variable named p, at state s (values by run()).
double dval(const Ex* e, const uint32_t* s, const std::string& p)
{
	switch (e->k)
	{
		case VAR: return e->name == p ? 1.0 : 0.0;
		case PAREN:
		case POS: return dval(e->a.get(), s, p);
		case NEG: return -dval(e->a.get(), s, p);
		case ADD: return dval(e->a.get(), s, p) + dval(e->b.get(), s, p);
		case SUB: return dval(e->a.get(), s, p) - dval(e->b.get(), s, p);
		case MUL:
			return dval(e->a.get(), s, p) * run(e->b.get(), s) + (double)run(e->a.get(), s) * dval(e->b.get(), s, p);
		case DIV:
		{
			const double a = run(e->a.get(), s), b = run(e->b.get(), s);
			return (dval(e->a.get(), s, p) * b - a * dval(e->b.get(), s, p)) / (b * b);
		}
		case TERN: return run(e->a.get(), s) != 0.0f ? dval(e->b.get(), s, p) : dval(e->c.get(), s, p);
		default: return 0.0; // literals, node values and boolean results are locally constant
	}
}

void collect_var_names(const Ex* e, std::set<std::string>& out)
{
	if (!e)
		return;
	if (e->k == VAR)
		out.insert(e->name);
	collect_var_names(e->a.get(), out);
	collect_var_names(e->b.get(), out);
	collect_var_names(e->c.get(), out);
}

struct Grad
{
	std::vector<std::string> names;            // parameters, sorted by name
	std::vector<std::vector<int>> up_p, down_p; // parameters each rate tree uses
	std::vector<double> G;                      // [windows << nobs]

	void init(const Sim& S)
	{
		std::set<std::string> all;
		for (int i = 0; i < S.N; i++)
		{
			collect_var_names(S.up[i].get(), all);
			collect_var_names(S.down[i].get(), all);
		}
		names.assign(all.begin(), all.end());
		auto uses = [&](const Ex* e) {
			std::set<std::string> v;
			collect_var_names(e, v);
			std::vector<int> idx;
			for (auto& nm : v)
				idx.push_back((int)(std::lower_bound(names.begin(), names.end(), nm) - names.begin()));
			return idx;
		};
		for (int i = 0; i < S.N; i++)
		{
			up_p.push_back(uses(S.up[i].get()));
			down_p.push_back(uses(S.down[i].get()));
		}
	}
};

// One trajectory (the same draws and float arithmetic as trajectory()),
// adding sum_w Gw * S(e_w) to acc.
void trajectory_grad(const Sim& S, const Grad& g, uint64_t id, std::vector<double>& acc, std::vector<float>& r,
					 std::vector<uint32_t>& s, uint64_t& transitions)
{
	const size_t P = g.names.size();
	std::vector<double> score(P, 0.0), drate(P);
	const uint32_t idlo = (uint32_t)id, idhi = (uint32_t)(id >> 32);
	std::fill(s.begin(), s.end(), 0u);
	for (int i = 0; i < S.N; i += 4)
	{
		u4 x = philox(u4 { { (uint32_t)(i >> 2), idlo, idhi, S.variant } }, S.seed_lo, S.seed_hi);
		for (int j = 0; j < 4 && i + j < S.N; j++)
			if (u01(x.v[j]) <= S.istate[i + j])
				s[(i + j) >> 5] |= 1u << ((i + j) & 31);
	}
	const uint32_t B0 = ((uint32_t)S.N + 3u) >> 2;
	double gw = 0.0; // G-weighted occupancy of the window in progress
	float t = 0.0f;
	for (uint32_t k = 0;;)
	{
		u4 x = philox(u4 { { B0 + (k >> 1), idlo, idhi, S.variant } }, S.seed_lo, S.seed_hi);
		const float u1 = u01(x.v[2 * (k & 1)]);
		const float u2 = u01(x.v[2 * (k & 1) + 1]);
		S.rates(s.data(), r.data());
		float L[64], P64[64];
		const float total = S.total_of(r.data(), L, P64);
		const float tp = t;
		if (total == 0.0f)
			t = S.max_time;
		else
			t = std::fmin(t + (-ln_poly(u1)) / total, S.max_time);
		transitions++;
		// d total / d theta in this state
		std::fill(drate.begin(), drate.end(), 0.0);
		for (int i = 0; i < S.N; i++)
		{
			const bool on = (s[i >> 5] >> (i & 31)) & 1u;
			const Ex* e = on ? S.down[i].get() : S.up[i].get();
			for (int j : on ? g.down_p[i] : g.up_p[i])
				drate[j] += dval(e, s.data(), g.names[j]);
		}
		const uint64_t o = S.obs_index(s.data());
		int w = (int)std::floor(tp / S.tick);
		float b = tp;
		double elapsed = 0.0; // dwell of this record so far
		while (t > b)
		{
			const float wend = (float)(w + 1) * S.tick;
			const float en = std::fmin(t, wend);
			const float sl = en - b;
			if (w < S.windows)
				gw += g.G[((size_t)w << S.nobs) | o] * (double)sl;
			elapsed += (double)sl;
			if (en == wend)
			{
				// window w ends inside this record: S_j(e_w) = score_j - drate_j * elapsed
				for (size_t j = 0; j < P; j++)
					acc[j] += gw * (score[j] - drate[j] * elapsed);
				gw = 0.0;
			}
			w++;
			b = en;
		}
		for (size_t j = 0; j < P; j++)
			score[j] -= drate[j] * elapsed;
		k++;
		if (t >= S.max_time || k >= S.max_steps)
		{
			for (size_t j = 0; j < P; j++)
				acc[j] += gw * score[j];
			return;
		}
		const int flip = S.select_flip(r.data(), total, u2 * total, L, P64);
		const bool on = (s[flip >> 5] >> (flip & 31)) & 1u;
		const Ex* e = on ? S.down[flip].get() : S.up[flip].get();
		for (int j : on ? g.down_p[flip] : g.up_p[flip])
			score[j] += dval(e, s.data(), g.names[j]) / (double)r[flip];
		s[flip >> 5] ^= 1u << (flip & 31);
	}
}

rates_fn compile_rates(const Sim& S, std::string& so_path)
{
	char dir[] = "/tmp/maboss_oracle_XXXXXX";
	if (!mkdtemp(dir))
		die("mkdtemp failed");
	std::string src = std::string(dir) + "/rates.cpp";
	so_path = std::string(dir) + "/rates.so";
	{
		std::ofstream f(src);
		f << "#include <stdint.h>\nextern \"C\" void rates(const uint32_t* s, float* r) {\n";
		for (int i = 0; i < S.N; i++)
			f << "  r[" << i << "] = ((s[" << (i >> 5) << "] >> " << (i & 31) << ") & 1u) ? " << emit(S.down[i].get())
			  << " : " << emit(S.up[i].get()) << ";\n";
		f << "}\n";
	}
	std::string cmd = "g++ -O3 -march=native -ffp-contract=off -fno-fast-math -shared -fPIC -o " + so_path + " " + src;
	if (std::system(cmd.c_str()) != 0)
		die("compiling the formulas failed: " + cmd);
	void* h = dlopen(so_path.c_str(), RTLD_NOW);
	if (!h)
		die(std::string("dlopen: ") + dlerror());
	return (rates_fn)dlsym(h, "rates");
}

std::string name_of(const Model& m, const std::vector<int>& idx, uint64_t bits)
{
	std::string s;
	for (size_t j = 0; j < idx.size(); j++)
		if ((bits >> j) & 1ull)
		{
			if (!s.empty())
				s += " -- ";
			s += m.nodes[idx[j]].name;
		}
	return s.empty() ? "<nil>" : s;
}

int numeric_probe()
{
	char op;
	while (std::cin >> op)
	{
		if (op == 'P')
		{
			u4 c;
			uint32_t k0, k1;
			std::cin >> std::hex >> c.v[0] >> c.v[1] >> c.v[2] >> c.v[3] >> k0 >> k1 >> std::dec;
			u4 r = philox(c, k0, k1);
			std::printf("%08x %08x %08x %08x\n", r.v[0], r.v[1], r.v[2], r.v[3]);
		}
		else if (op == 'L')
		{
			uint32_t b;
			std::cin >> std::hex >> b >> std::dec;
			std::printf("%08x\n", fbits(log2_poly(bitsf(b))));
		}
		else if (op == 'U')
		{
			uint32_t x;
			std::cin >> std::hex >> x >> std::dec;
			std::printf("%08x\n", fbits(u01(x)));
		}
	}
	return 0;
}

} // namespace orc

int main(int argc, char** argv)
{
	using namespace orc;
	if (argc == 2 && std::string(argv[1]) == "--numeric-probe")
		return numeric_probe();
	std::vector<std::string> pos;
	long long n_override = -1, first = 0, seed_override = -1;
	int threads = 0;
	bool compiled = false, timing = false, eval_states = false, errors = false;
	std::string raw, prefix, print_path, entropy = "log_sum", order = "seq", grad_in, grad_out;
	unsigned variant = 0, max_steps = 1u << 30;
	for (int i = 1; i < argc; i++)
	{
		std::string a = argv[i];
		auto next = [&]() -> std::string {
			if (i + 1 >= argc)
				die("missing value for " + a);
			return argv[++i];
		};
		if (a == "--n")
			n_override = std::stoll(next());
		else if (a == "--first-id")
			first = std::stoll(next());
		else if (a == "--seed")
			seed_override = std::stoll(next());
		else if (a == "--variant")
			variant = (unsigned)std::stoul(next());
		else if (a == "--threads")
			threads = std::stoi(next());
		else if (a == "--compiled")
			compiled = true;
		else if (a == "--order")
			order = next();
		else if (a == "--entropy")
			entropy = next();
		else if (a == "--max-steps")
			max_steps = (unsigned)std::stoul(next());
		else if (a == "--raw")
			raw = next();
		else if (a == "-o")
			prefix = next();
		else if (a == "--print")
			print_path = next();
		else if (a == "--time")
			timing = true;
		else if (a == "--eval-states")
			eval_states = true;
		else if (a == "--errors")
			errors = true;
		else if (a == "--grad-in")
			grad_in = next();
		else if (a == "--grad-out")
			grad_out = next();
		else
			pos.push_back(a);
	}
	if (pos.size() != 2)
		die("usage: maboss_oracle [options] model.bnd model.cfg");

	Model m;
	{
		Parser pb(slurp(pos[0]), pos[0], m);
		pb.bnd();
		Parser pc(slurp(pos[1]), pos[1], m);
		pc.cfg();
	}
	// A forced-on node never switches off and starts on; forced off, the
	// reverse.  The blocked rate is replaced by the constant 0, so the rest
	// of the oracle sees an ordinary model.
	for (auto& [i, dir] : m.mutants)
	{
		auto zero = std::make_shared<Ex>();
		zero->k = LIT;
		Node& nd = m.nodes[i];
		bool found = false;
		for (auto& a : nd.attrs)
			if (a.first == (dir > 0 ? "rate_down" : "rate_up"))
			{
				a.second = zero;
				found = true;
			}
		if (!found)
			die("mutant node without the rate it blocks");
		nd.istate = dir > 0 ? 1.0f : 0.0f;
	}

	Sim S;
	S.N = (int)m.nodes.size();
	S.W = (S.N + 31) / 32;
	for (int i = 0; i < S.N; i++)
	{
		const Node& nd = m.nodes[i];
		ExP isint = nd.attr("is_internal");
		bool internal = false;
		if (isint)
		{
			Parser tmp("", "", m);
			internal = tmp.eval(isint) != 0;
		}
		if (!internal)
			S.obs.push_back(i);
		S.istate.push_back(nd.istate);
		S.up.push_back(resolve(m, nd, nd.attr("rate_up"), 0));
		S.down.push_back(resolve(m, nd, nd.attr("rate_down"), 0));
	}
	S.nobs = (int)S.obs.size();
	S.discrete = m.consts["discrete_time"] != 0;
	S.max_time = m.consts["max_time"];
	S.tick = m.consts["time_tick"];
	S.windows = (int)std::ceil(S.max_time / S.tick);
	S.reference_entropy = entropy == "reference";
	S.wave = order == "wave";
	S.npl = (S.N + 63) / 64;
	if (S.wave && S.reference_entropy)
		die("the wave order uses the log-sum entropy");
	S.max_steps = max_steps;
	S.variant = variant;
	uint64_t seed = seed_override >= 0 ? (uint64_t)seed_override : (uint64_t)m.consts_exact["seed_pseudorandom"];
	S.seed_lo = (uint32_t)seed;
	S.seed_hi = (uint32_t)(seed >> 32);
	const uint64_t n = n_override >= 0 ? (uint64_t)n_override : (uint64_t)m.consts_exact["sample_count"];
	if (S.nobs > 40 && S.windows > (1 << 22))
		die("window key does not fit");

	if (eval_states)
	{
		std::vector<uint32_t> st(S.W);
		std::vector<float> r(S.N);
		uint32_t u2b;
		while (true)
		{
			for (int w = 0; w < S.W; w++)
				if (!(std::cin >> std::hex >> st[w]))
					return 0;
			std::cin >> std::hex >> u2b;
			S.rates(st.data(), r.data());
			float L[64], P[64];
			const float total = S.total_of(r.data(), L, P);
			const int flip = total == 0.0f ? 0 : S.select_flip(r.data(), total, bitsf(u2b) * total, L, P);
			std::printf("%08x %08x %d %llx\n", fbits(total), fbits(S.entropy(r.data())), flip,
						(unsigned long long)S.obs_index(st.data()));
		}
	}

	if (!grad_in.empty())
	{
		if (S.discrete)
			die("the gradient is defined for continuous time");
		if (S.wave)
			die("the gradient uses the lane order");
		Grad g;
		g.init(S);
		const size_t cells = (size_t)S.windows << S.nobs;
		g.G.resize(cells);
		{
			std::ifstream f(grad_in, std::ios::binary);
			if (!f.read((char*)g.G.data(), (std::streamsize)(cells * sizeof(double))))
				die("cannot read " + std::to_string(cells) + " doubles from " + grad_in);
		}
#ifdef _OPENMP
		if (threads > 0)
			omp_set_num_threads(threads);
		const int gth = omp_get_max_threads();
#else
		const int gth = 1;
#endif
		std::vector<std::vector<double>> accs(gth, std::vector<double>(g.names.size(), 0.0));
		std::vector<uint64_t> trans(gth, 0);
#pragma omp parallel
		{
#ifdef _OPENMP
			const int th = omp_get_thread_num();
#else
			const int th = 0;
#endif
			std::vector<float> r(S.N);
			std::vector<uint32_t> st(S.W);
#pragma omp for schedule(dynamic, 64)
			for (long long i = 0; i < (long long)n; i++)
				trajectory_grad(S, g, (uint64_t)first + (uint64_t)i, accs[th], r, st, trans[th]);
		}
		std::vector<double> tot(g.names.size(), 0.0);
		uint64_t ntr = 0;
		for (int th = 0; th < gth; th++)
		{
			for (size_t j = 0; j < tot.size(); j++)
				tot[j] += accs[th][j];
			ntr += trans[th];
		}
		FILE* f = grad_out.empty() ? stdout : std::fopen(grad_out.c_str(), "w");
		if (!f)
			die("cannot write " + grad_out);
		std::fprintf(f, "transitions %llu\n", (unsigned long long)ntr);
		for (size_t j = 0; j < tot.size(); j++)
			std::fprintf(f, "%s %.17g\n", g.names[j].c_str(), tot[j] / ((double)n * (double)S.tick));
		if (f != stdout)
			std::fclose(f);
		return 0;
	}

	std::string so;
	const double tc0 = std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
	if (compiled)
		S.compiled = compile_rates(S, so);
	const double tc1 = std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();

#ifdef _OPENMP
	if (threads > 0)
		omp_set_num_threads(threads);
	int nth = omp_get_max_threads();
#else
	int nth = 1;
#endif
	std::vector<Stats> per(nth);
	for (auto& st : per)
	{
		st.init(S.nobs, S.windows);
		st.moments = errors;
	}
	const auto t0 = std::chrono::steady_clock::now();
#pragma omp parallel
	{
#ifdef _OPENMP
		Stats& st = per[omp_get_thread_num()];
#else
		Stats& st = per[0];
#endif
		std::vector<float> r(S.N);
		std::vector<uint32_t> s(S.W);
#pragma omp for schedule(dynamic, 64)
		for (long long i = 0; i < (long long)n; i++)
			trajectory(S, (uint64_t)first + (uint64_t)i, st, r, s);
	}
	Stats tot = std::move(per[0]);
	for (int i = 1; i < nth; i++)
		tot.merge(per[i]);
	const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();

	// collect windows sorted by (w, obs)
	std::vector<std::vector<std::pair<uint64_t, uint64_t>>> win(S.windows);
	const uint64_t omask = S.nobs >= 64 ? ~0ull : ((1ull << S.nobs) - 1);
	if (tot.dense)
	{
		for (uint64_t key = 0; key < tot.dwin.size(); key++)
			if (tot.dwin[key])
				win[key >> S.nobs].emplace_back(key & omask, tot.dwin[key]);
	}
	else
	{
		tot.hwin.each([&](uint64_t k, uint64_t v) {
			if (v)
				win[k >> S.nobs].emplace_back(k & omask, v);
		});
		for (auto& v : win)
			std::sort(v.begin(), v.end());
	}
	std::vector<std::pair<uint64_t, uint64_t>> fin;
	tot.fin.each([&](uint64_t k, uint64_t v) { fin.emplace_back(k, v); });
	std::sort(fin.begin(), fin.end());

	if (timing)
		std::fprintf(stderr,
					 "{\"trajectories\": %llu, \"transitions\": %llu, \"seconds\": %.6f, \"threads\": %d, "
					 "\"transitions_per_s\": %.6e, \"compile_seconds\": %.3f}\n",
					 (unsigned long long)n, (unsigned long long)tot.transitions, secs, nth,
					 tot.transitions / std::max(secs, 1e-12), tc1 - tc0);

	if (!raw.empty())
	{
		std::ofstream f(raw);
		f << "MBG_RAW 1\n";
		f << "windows " << S.windows << " n_obs " << S.nobs << " discrete " << (int)S.discrete << " n " << n << "\n";
		f << "T " << tot.transitions << " " << tot.finished << " " << tot.truncated << "\n";
		for (int w = 0; w < S.windows; w++)
			if (tot.ent[w])
				f << "E " << w << " " << tot.ent[w] << "\n";
		for (int w = 0; w < S.windows; w++)
			for (auto& c : win[w])
				f << "W " << w << " " << c.first << " " << c.second << "\n";
		for (auto& c : fin)
			f << "F " << c.first << " " << c.second << "\n";
		if (errors)
		{
			auto u128s = [](u128 v) { return std::to_string((uint64_t)(v >> 64)) + " " + std::to_string((uint64_t)v); };
			for (int w = 0; w < S.windows; w++)
				if (tot.esq[w])
					f << "R " << w << " " << u128s(tot.esq[w]) << "\n";
			std::vector<std::pair<uint64_t, u128>> q(tot.sq.begin(), tot.sq.end());
			std::sort(q.begin(), q.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
			for (auto& c : q)
				if (c.second)
					f << "Q " << (c.first >> S.nobs) << " " << (c.first & omask) << " " << u128s(c.second) << "\n";
		}
		for (auto& kv : tot.fixedp)
		{
			f << "P";
			for (auto it = kv.first.rbegin(); it != kv.first.rend(); ++it)
				f << " " << *it;
			f << " " << kv.second << "\n";
		}
	}

	auto prob = [&](uint64_t raw_v) {
		return S.discrete ? (double)raw_v / (double)n : (double)raw_v / ((double)n * 4294967296.0);
	};
	auto Hw = [&](int w) {
		float e = 0.f;
		for (auto& c : win[w])
		{
			float p = (float)prob(c.second);
			if (p != 0.f)
				e += -std::log2(p) * p;
		}
		return e;
	};
	auto THw = [&](int w) { return (float)((double)tot.ent[w] / ((double)n * 4294967296.0)); };
	// error columns: 0 without --errors, as the reference prints
	auto err_th = [&](int w) {
		return errors ? (float)std_error(tot.ent[w], tot.esq[w], (double)n, 4294967296.0) : 0.f;
	};
	auto err_p = [&](int w, uint64_t obs, uint64_t s1) {
		if (!errors)
			return 0.f;
		auto it = tot.sq.find(((uint64_t)w << S.nobs) | obs);
		return (float)std_error(s1, it == tot.sq.end() ? 0 : it->second, (double)n, S.discrete ? 1.0 : 4294967296.0);
	};
	std::vector<int> all(S.N);
	for (int i = 0; i < S.N; i++)
		all[i] = i;
	auto full_name = [&](const std::vector<uint32_t>& rev) {
		uint64_t dummy = 0;
		(void)dummy;
		std::string s;
		for (int i = 0; i < S.N; i++)
		{
			const uint32_t word = rev[rev.size() - 1 - (i >> 5)];
			if ((word >> (i & 31)) & 1u)
			{
				if (!s.empty())
					s += " -- ";
				s += m.nodes[i].name;
			}
		}
		return s.empty() ? std::string("<nil>") : s;
	};
	if (!print_path.empty())
	{
		std::ofstream f(print_path);
		f << "final points:\n";
		for (auto& c : fin)
			f << (float)((double)c.second / n) << " " << name_of(m, S.obs, c.first) << "\n";
		f << "fixed points:\n";
		for (auto& kv : tot.fixedp)
			f << (float)((double)kv.second / n) << " " << full_name(kv.first) << "\n";
		for (int w = 0; w < S.windows; w++)
		{
			f << "window (" << (float)w * S.tick << ", " << (float)(w + 1) * S.tick << "]\n";
			f << "entropy: " << Hw(w) << "\n";
			f << "transition entropy: " << THw(w) << "\n";
			for (auto& c : win[w])
			{
				float p = (float)prob(c.second);
				if (p != 0.f)
					f << p << " " << name_of(m, S.obs, c.first) << "\n";
			}
		}
	}
	if (!prefix.empty())
	{
		std::ofstream f(prefix + "_probtraj.csv");
		size_t maxs = 0;
		for (int w = 0; w < S.windows; w++)
		{
			size_t c = 0;
			for (auto& x : win[w])
				c += (float)prob(x.second) != 0.f;
			maxs = std::max(maxs, c);
		}
		f << "Time\tTH\tErrorTH\tH\tHD=0";
		for (size_t i = 0; i < maxs; i++)
			f << "\tState\tProba\tErrorProba";
		f << "\n";
		for (int w = 0; w < S.windows; w++)
		{
			f << (float)w * S.tick << "\t" << THw(w) << "\t" << err_th(w) << "\t" << Hw(w) << "\t" << 0.f;
			for (auto& c : win[w])
			{
				float p = (float)prob(c.second);
				if (p != 0.f)
					f << "\t" << name_of(m, S.obs, c.first) << "\t" << p << "\t" << err_p(w, c.first, c.second);
			}
			f << "\n";
		}
		std::ofstream g(prefix + "_fp.csv");
		g << "Fixed Points (" << tot.fixedp.size() << ")\nFP\tProba\tState";
		for (auto& nd : m.nodes)
			g << "\t" << nd.name;
		g << "\n";
		int i = 0;
		for (auto& kv : tot.fixedp)
		{
			g << "#" << i++ << "\t" << (float)((double)kv.second / n) << "\t" << full_name(kv.first);
			for (int k = 0; k < S.N; k++)
				g << "\t" << ((kv.first[kv.first.size() - 1 - (k >> 5)] >> (k & 31)) & 1u);
			g << "\n";
		}
	}
	if (!so.empty())
		unlink(so.c_str());
	return 0;
}
