// Lexer, parser and constant evaluation for .bnd/.cfg models.
//
// Token set and precedence follow src/parser/scanner.ll:20-93 and
// src/parser/parser.yy:71-159 (precedence table :126-134).  Two productions
// the reference declares but cannot use are made to work here: XOR
// (src/parser/expressions.cpp:78-79 throws) and the bracket istate form
// `[N].istate = a[1], b[0];` (src/parser/parser.yy:102-108, whose tokens the
// reference lexer lacks).  `/* ... */` comments are accepted as well.
#include "model.h"

#include <algorithm>
#include <cctype>
#include <cerrno>
#include <climits>
#include <cmath>
#include <cstdlib>
#include <fstream>
#include <set>
#include <sstream>

namespace mbg
{

namespace
{

enum class tok
{
	end,
	minus,
	plus,
	star,
	slash,
	lparen,
	rparen,
	assign,
	ternary,
	colon,
	land,
	lor,
	lxor,
	lnot,
	eq,
	ne,
	lt,
	le,
	gt,
	ge,
	lbrace,
	rbrace,
	dot,
	semicolon,
	lbracket,
	rbracket,
	comma,
	number,
	fnumber,
	identifier,
	variable,
	alias
};

struct token
{
	tok kind = tok::end;
	std::string text;
	double num = 0.0;
	int line = 1, col = 1;
};

class lexer
{
public:
	lexer(const std::string& src, const std::string& fname) : s_(src), fname_(fname) {}

	[[noreturn]] void fail(int line, int col, const std::string& msg) const
	{
		throw error(MBG_E_PARSE, fname_ + ":" + std::to_string(line) + "." + std::to_string(col) + ": " + msg);
	}

	token next()
	{
		skip();
		token t;
		t.line = line_;
		t.col = col_;
		if (p_ >= s_.size())
			return t;
		char c = s_[p_];
		auto two = [&](char a, char b) { return c == a && p_ + 1 < s_.size() && s_[p_ + 1] == b; };
		auto emit = [&](tok k, size_t len) {
			t.kind = k;
			t.text = s_.substr(p_, len);
			adv(len);
			return t;
		};
		if (two('&', '&'))
			return emit(tok::land, 2);
		if (two('|', '|'))
			return emit(tok::lor, 2);
		if (two('=', '='))
			return emit(tok::eq, 2);
		if (two('!', '='))
			return emit(tok::ne, 2);
		if (two('<', '='))
			return emit(tok::le, 2);
		if (two('>', '='))
			return emit(tok::ge, 2);
		switch (c)
		{
			case '-': return emit(tok::minus, 1);
			case '+': return emit(tok::plus, 1);
			case '*': return emit(tok::star, 1);
			case '/': return emit(tok::slash, 1);
			case '(': return emit(tok::lparen, 1);
			case ')': return emit(tok::rparen, 1);
			case '=': return emit(tok::assign, 1);
			case '?': return emit(tok::ternary, 1);
			case ':': return emit(tok::colon, 1);
			case '&': return emit(tok::land, 1);
			case '|': return emit(tok::lor, 1);
			case '^': return emit(tok::lxor, 1);
			case '!': return emit(tok::lnot, 1);
			case '<': return emit(tok::lt, 1);
			case '>': return emit(tok::gt, 1);
			case '{': return emit(tok::lbrace, 1);
			case '}': return emit(tok::rbrace, 1);
			case ';': return emit(tok::semicolon, 1);
			case '[': return emit(tok::lbracket, 1);
			case ']': return emit(tok::rbracket, 1);
			case ',': return emit(tok::comma, 1);
			default: break;
		}
		// numbers: {int} | fracconst exppart? | [0-9]+ exppart
		if (std::isdigit((unsigned char)c) || (c == '.' && p_ + 1 < s_.size() && std::isdigit((unsigned char)s_[p_ + 1])))
		{
			size_t q = p_;
			bool is_float = false;
			while (q < s_.size() && std::isdigit((unsigned char)s_[q]))
				q++;
			if (q < s_.size() && s_[q] == '.')
			{
				is_float = true;
				q++;
				while (q < s_.size() && std::isdigit((unsigned char)s_[q]))
					q++;
			}
			if (q < s_.size() && (s_[q] == 'e' || s_[q] == 'E'))
			{
				size_t r = q + 1;
				if (r < s_.size() && (s_[r] == '+' || s_[r] == '-'))
					r++;
				if (r < s_.size() && std::isdigit((unsigned char)s_[r]))
				{
					while (r < s_.size() && std::isdigit((unsigned char)s_[r]))
						r++;
					q = r;
					is_float = true;
				}
			}
			std::string txt = s_.substr(p_, q - p_);
			errno = 0;
			if (is_float)
			{
				double v = std::strtod(txt.c_str(), nullptr);
				if (errno == ERANGE)
					fail(t.line, t.col, "float is out of range: " + txt);
				t.kind = tok::fnumber;
				t.num = v;
			}
			else
			{
				// 64-bit range (the reference stops at INT_MAX, scanner.ll:98-105),
				// so large seeds keep every digit
				long long v = std::strtoll(txt.c_str(), nullptr, 10);
				if (errno == ERANGE)
					fail(t.line, t.col, "integer is out of range: " + txt);
				t.kind = tok::number;
				t.num = (double)v;
			}
			t.text = txt;
			adv(q - p_);
			return t;
		}
		if (c == '.')
			return emit(tok::dot, 1);
		auto ident_len = [&](size_t from) {
			size_t q = from;
			if (q < s_.size() && std::isalpha((unsigned char)s_[q]))
			{
				q++;
				while (q < s_.size() && (std::isalnum((unsigned char)s_[q]) || s_[q] == '_'))
					q++;
			}
			return q - from;
		};
		if (c == '$' || c == '@')
		{
			size_t l = ident_len(p_ + 1);
			if (l == 0)
				fail(t.line, t.col, std::string("invalid character: ") + c);
			return emit(c == '$' ? tok::variable : tok::alias, l + 1);
		}
		size_t l = ident_len(p_);
		if (l > 0)
		{
			std::string w = s_.substr(p_, l);
			std::string lw = w;
			std::transform(lw.begin(), lw.end(), lw.begin(), ::tolower);
			if (lw == "and")
				return emit(tok::land, l);
			if (lw == "or")
				return emit(tok::lor, l);
			if (lw == "xor")
				return emit(tok::lxor, l);
			if (lw == "not")
				return emit(tok::lnot, l);
			if (lw == "true" || lw == "false")
			{
				t = emit(tok::number, l);
				t.num = lw == "true" ? 1.0 : 0.0;
				return t;
			}
			return emit(tok::identifier, l);
		}
		fail(t.line, t.col, std::string("invalid character: ") + c);
	}

private:
	void adv(size_t n)
	{
		for (size_t i = 0; i < n; i++)
		{
			if (s_[p_] == '\n')
			{
				line_++;
				col_ = 1;
			}
			else
				col_++;
			p_++;
		}
	}

	void skip()
	{
		while (p_ < s_.size())
		{
			char c = s_[p_];
			if (c == ' ' || c == '\t' || c == '\r' || c == '\n' || c == '\f' || c == '\v')
				adv(1);
			else if (c == '/' && p_ + 1 < s_.size() && s_[p_ + 1] == '/')
			{
				while (p_ < s_.size() && s_[p_] != '\n')
					adv(1);
			}
			else if (c == '/' && p_ + 1 < s_.size() && s_[p_ + 1] == '*')
			{
				int l = line_, cc = col_;
				adv(2);
				while (p_ + 1 < s_.size() && !(s_[p_] == '*' && s_[p_ + 1] == '/'))
					adv(1);
				if (p_ + 1 >= s_.size())
					fail(l, cc, "unterminated comment");
				adv(2);
			}
			else if ((unsigned char)c == 0xEF && p_ + 2 < s_.size() && (unsigned char)s_[p_ + 1] == 0xBB
					 && (unsigned char)s_[p_ + 2] == 0xBF)
				adv(3); // UTF-8 byte-order mark
			else
				break;
		}
	}

	const std::string& s_;
	std::string fname_;
	size_t p_ = 0;
	int line_ = 1, col_ = 1;
};

const char* tok_name(tok k)
{
	switch (k)
	{
		case tok::end: return "end of file";
		case tok::minus: return "\"-\"";
		case tok::plus: return "\"+\"";
		case tok::star: return "\"*\"";
		case tok::slash: return "\"/\"";
		case tok::lparen: return "\"(\"";
		case tok::rparen: return "\")\"";
		case tok::assign: return "\"=\"";
		case tok::ternary: return "\"?\"";
		case tok::colon: return "\":\"";
		case tok::land: return "\"&&\"";
		case tok::lor: return "\"||\"";
		case tok::lxor: return "\"^\"";
		case tok::lnot: return "\"!\"";
		case tok::eq: return "\"==\"";
		case tok::ne: return "\"!=\"";
		case tok::lt: return "\"<\"";
		case tok::le: return "\"<=\"";
		case tok::gt: return "\">\"";
		case tok::ge: return "\">=\"";
		case tok::lbrace: return "\"{\"";
		case tok::rbrace: return "\"}\"";
		case tok::dot: return "\".\"";
		case tok::semicolon: return "\";\"";
		case tok::lbracket: return "\"[\"";
		case tok::rbracket: return "\"]\"";
		case tok::comma: return "\",\"";
		case tok::number: return "NUMBER";
		case tok::fnumber: return "FLOAT";
		case tok::identifier: return "IDENTIFIER";
		case tok::variable: return "VARIABLE";
		case tok::alias: return "ALIAS";
	}
	return "?";
}

// Binding powers from src/parser/parser.yy:126-134 (higher binds tighter).
int infix_power(tok k)
{
	switch (k)
	{
		case tok::ternary: return 1;
		case tok::lxor: return 2;
		case tok::lor: return 3;
		case tok::land: return 4;
		case tok::eq:
		case tok::ne: return 5;
		case tok::lt:
		case tok::le:
		case tok::gt:
		case tok::ge: return 6;
		case tok::plus:
		case tok::minus: return 7;
		case tok::star:
		case tok::slash: return 8;
		default: return 0;
	}
}
constexpr int unary_power = 9;

op binary_op(tok k)
{
	switch (k)
	{
		case tok::lxor: return op::lxor;
		case tok::lor: return op::lor;
		case tok::land: return op::land;
		case tok::eq: return op::eq;
		case tok::ne: return op::ne;
		case tok::lt: return op::lt;
		case tok::le: return op::le;
		case tok::gt: return op::gt;
		case tok::ge: return op::ge;
		case tok::plus: return op::add;
		case tok::minus: return op::sub;
		case tok::star: return op::mul;
		case tok::slash: return op::div;
		default: return op::lit;
	}
}

class parser
{
public:
	parser(model& m, const std::string& text, const std::string& fname) : m_(m), lex_(text, fname)
	{
		cur_ = lex_.next();
	}

	void bnd_program()
	{
		while (cur_.kind != tok::end)
		{
			token kw = expect(tok::identifier);
			std::string lkw = kw.text;
			std::transform(lkw.begin(), lkw.end(), lkw.begin(), ::tolower);
			if (lkw != "node")
				lex_.fail(kw.line, kw.col, "expected 'node' keyword");
			token name = expect(tok::identifier);
			expect(tok::lbrace);
			node_t n;
			n.name = name.text;
			do
			{
				token an = expect(tok::identifier);
				expect(tok::assign);
				expr_ptr e = expression(0);
				expect(tok::semicolon);
				n.attrs.emplace_back(an.text, std::move(e));
			} while (cur_.kind != tok::rbrace);
			expect(tok::rbrace);
			register_node(std::move(n), name);
		}
	}

	void cfg_program()
	{
		while (cur_.kind != tok::end)
		{
			if (cur_.kind == tok::variable)
			{
				token v = take();
				expect(tok::assign);
				expr_ptr e = expression(0);
				expect(tok::semicolon);
				m_.variables[v.text] = eval(*e, v);
			}
			else if (cur_.kind == tok::lbracket)
			{
				take();
				token n = expect(tok::identifier);
				expect(tok::rbracket);
				expect(tok::dot);
				token a = expect(tok::identifier);
				if (a.text != "istate")
					lex_.fail(a.line, a.col, "expected 'istate' keyword");
				expect(tok::assign);
				expr_ptr el = expression(0);
				expect(tok::lbracket);
				token vl = expect(tok::number);
				expect(tok::rbracket);
				expect(tok::comma);
				expr_ptr er = expression(0);
				expect(tok::lbracket);
				token vr = expect(tok::number);
				expect(tok::rbracket);
				expect(tok::semicolon);
				if (!((vl.num == 0 && vr.num == 1) || (vl.num == 1 && vr.num == 0)))
					lex_.fail(vl.line, vl.col, "numbers in [] must be 0 and 1");
				int idx = m_.node_index(n.text);
				if (idx < 0)
					throw error(MBG_E_MODEL, "Unknown node " + n.text);
				float l = eval(*el, n), r = eval(*er, n);
				// src/parser/driver.cpp:79-96
				m_.nodes[idx].istate = vl.num == 1 ? l / (l + r) : r / (l + r);
			}
			else
			{
				token id = expect(tok::identifier);
				if (cur_.kind == tok::dot)
				{
					take();
					token a = expect(tok::identifier);
					expect(tok::assign);
					expr_ptr e = expression(0);
					expect(tok::semicolon);
					int idx = m_.node_index(id.text);
					if (idx < 0)
						throw error(MBG_E_MODEL, "Unknown node " + id.text);
					// src/parser/driver.cpp:65-77
					if (a.text == "istate")
						m_.nodes[idx].istate = eval(*e, a);
					else if (a.text == "mutant")
						m_.nodes[idx].mutant = mutant_value(*e, a);
					else
						m_.nodes[idx].attrs.emplace_back(a.text, std::move(e));
				}
				else
				{
					expect(tok::assign);
					expr_ptr e = expression(0);
					expect(tok::semicolon);
					float v = eval(*e, id);
					m_.constants[id.text] = v;
					double ex = v;
					if (literal_value(*e, ex))
						m_.constants_exact[id.text] = ex;
					else
						m_.constants_exact[id.text] = v;
				}
			}
		}
	}

private:
	// ON / OFF (any case) or a constant expression equal to 1 / 0
	int mutant_value(const expr& e, const token& where)
	{
		if (e.kind == op::node)
		{
			std::string v = e.name;
			for (char& ch : v)
				ch = (char)std::toupper((unsigned char)ch);
			if (v == "ON")
				return 1;
			if (v == "OFF")
				return -1;
		}
		else
		{
			const float v = eval(e, where);
			if (v == 1.0f)
				return 1;
			if (v == 0.0f)
				return -1;
		}
		lex_.fail(where.line, where.col, "mutant must be ON or OFF (1 or 0)");
	}

	static bool literal_value(const expr& e, double& out)
	{
		if (e.kind == op::lit)
		{
			out = e.exact;
			return true;
		}
		if (e.kind == op::paren || e.kind == op::pos)
			return literal_value(*e.a, out);
		if (e.kind == op::neg && literal_value(*e.a, out))
		{
			out = -out;
			return true;
		}
		return false;
	}

	float eval(const expr& e, const token& where)
	{
		try
		{
			return m_.evaluate(e);
		}
		catch (const error& err)
		{
			lex_.fail(where.line, where.col, err.what());
		}
	}

	void register_node(node_t n, const token& where)
	{
		// src/parser/driver.cpp:49-63
		if (m_.node_index(n.name) >= 0)
			lex_.fail(where.line, where.col, "Node " + n.name + " already exists");
		if (!n.attr("rate_up"))
			lex_.fail(where.line, where.col, "Node " + n.name + " does not have rate_up attribute");
		if (!n.attr("rate_down"))
			lex_.fail(where.line, where.col, "Node " + n.name + " does not have rate_down attribute");
		m_.nodes.push_back(std::move(n));
	}

	token take()
	{
		token t = cur_;
		cur_ = lex_.next();
		return t;
	}

	token expect(tok k)
	{
		if (cur_.kind != k)
			lex_.fail(cur_.line, cur_.col,
					  std::string("syntax error, unexpected ") + tok_name(cur_.kind) + ", expecting " + tok_name(k));
		return take();
	}

	expr_ptr primary()
	{
		token t = take();
		switch (t.kind)
		{
			case tok::number:
			case tok::fnumber: {
				auto e = expr::make(op::lit);
				e->value = (float)t.num;
				e->exact = t.num;
				return e;
			}
			case tok::identifier: {
				auto e = expr::make(op::node);
				e->name = t.text;
				return e;
			}
			case tok::variable: {
				auto e = expr::make(op::var);
				e->name = t.text;
				return e;
			}
			case tok::alias: {
				auto e = expr::make(op::alias);
				e->name = t.text;
				return e;
			}
			case tok::lparen: {
				auto e = expr::make(op::paren);
				e->a = expression(0);
				expect(tok::rparen);
				return e;
			}
			case tok::minus:
			case tok::plus:
			case tok::lnot: {
				auto e = expr::make(t.kind == tok::minus ? op::neg : t.kind == tok::plus ? op::pos : op::lnot);
				e->a = expression(unary_power);
				return e;
			}
			default:
				lex_.fail(t.line, t.col, std::string("syntax error, unexpected ") + tok_name(t.kind));
		}
	}

	// Precedence climbing; every binary operator is left-associative and the
	// ternary is right-associative (src/parser/parser.yy:126-134).
	expr_ptr expression(int min_power)
	{
		expr_ptr lhs = primary();
		for (;;)
		{
			int p = infix_power(cur_.kind);
			if (p == 0 || p <= min_power)
				break;
			token t = take();
			if (t.kind == tok::ternary)
			{
				auto e = expr::make(op::ternary);
				e->a = std::move(lhs);
				e->b = expression(0);
				expect(tok::colon);
				e->c = expression(p - 1);
				lhs = std::move(e);
			}
			else
			{
				auto e = expr::make(binary_op(t.kind));
				e->a = std::move(lhs);
				e->b = expression(p);
				lhs = std::move(e);
			}
		}
		return lhs;
	}

	model& m_;
	lexer lex_;
	token cur_;
};

} // namespace

model::model()
{
	// src/parser/driver.cpp:9-13
	constants["discrete_time"] = 0;
	constants["sample_count"] = 1'000'000;
	constants["max_time"] = 10;
	constants["time_tick"] = 0.1f;
	constants["seed_pseudorandom"] = 0;
	constants_exact["sample_count"] = 1'000'000;
	constants_exact["seed_pseudorandom"] = 0;
}

int model::node_index(const std::string& n) const
{
	for (size_t i = 0; i < nodes.size(); i++)
		if (nodes[i].name == n)
			return (int)i;
	return -1;
}

bool model::is_internal(size_t i) const
{
	const expr* e = nodes[i].attr("is_internal");
	return e && evaluate(*e) != 0;
}

int model::n_observed() const
{
	int c = 0;
	for (size_t i = 0; i < nodes.size(); i++)
		c += !is_internal(i);
	return c;
}

std::vector<int> model::observed_nodes() const
{
	std::vector<int> r;
	for (size_t i = 0; i < nodes.size(); i++)
		if (!is_internal(i))
			r.push_back((int)i);
	return r;
}

bool model::discrete_time() const { return constants.at("discrete_time") != 0; }
float model::max_time() const { return constants.at("max_time"); }
float model::time_tick() const { return constants.at("time_tick"); }

int64_t model::sample_count() const
{
	double v = constants_exact.count("sample_count") ? constants_exact.at("sample_count") : constants.at("sample_count");
	return (int64_t)v;
}

uint64_t model::seed() const
{
	double v = constants_exact.count("seed_pseudorandom") ? constants_exact.at("seed_pseudorandom")
														  : constants.at("seed_pseudorandom");
	if (v < 0)
		return (uint64_t)(int64_t)v;
	return (uint64_t)v;
}

int model::windows() const
{
	float w = std::ceil(max_time() / time_tick());
	return (int)w;
}

float model::evaluate(const expr& e) const
{
	switch (e.kind)
	{
		case op::lit: return e.value;
		case op::paren:
		case op::pos: return evaluate(*e.a);
		case op::neg: return -evaluate(*e.a);
		case op::lnot: return !evaluate(*e.a);
		case op::node: throw error(MBG_E_MODEL, "identifier " + e.name + " in expression which needs to be evaluated");
		case op::alias: throw error(MBG_E_MODEL, "alias " + e.name + " in expression which needs to be evaluated");
		case op::var: {
			auto it = variables.find(e.name);
			if (it == variables.end())
				throw error(MBG_E_MODEL, "undefined variable " + e.name);
			return it->second;
		}
		case op::add: return evaluate(*e.a) + evaluate(*e.b);
		case op::sub: return evaluate(*e.a) - evaluate(*e.b);
		case op::mul: return evaluate(*e.a) * evaluate(*e.b);
		case op::div: return evaluate(*e.a) / evaluate(*e.b);
		case op::land: return evaluate(*e.a) && evaluate(*e.b);
		case op::lor: return evaluate(*e.a) || evaluate(*e.b);
		case op::lxor: return (evaluate(*e.a) != 0) != (evaluate(*e.b) != 0);
		case op::eq: return evaluate(*e.a) == evaluate(*e.b);
		case op::ne: return evaluate(*e.a) != evaluate(*e.b);
		case op::lt: return evaluate(*e.a) < evaluate(*e.b);
		case op::le: return evaluate(*e.a) <= evaluate(*e.b);
		case op::gt: return evaluate(*e.a) > evaluate(*e.b);
		case op::ge: return evaluate(*e.a) >= evaluate(*e.b);
		case op::ternary: return evaluate(*e.a) ? evaluate(*e.b) : evaluate(*e.c);
	}
	throw error(MBG_E_INTERNAL, "bad expression");
}

namespace
{
float interp(const model& m, const node_t& cur, const expr& e, const uint32_t* s, int depth)
{
	if (depth > 64)
		throw error(MBG_E_MODEL, "alias nesting too deep");
	auto f = [&](const expr& x) { return interp(m, cur, x, s, depth); };
	switch (e.kind)
	{
		case op::node: {
			int i = m.node_index(e.name);
			if (i < 0)
				throw error(MBG_E_MODEL, "unknown node name: " + e.name);
			return (float)((s[i / 32] >> (i % 32)) & 1u);
		}
		case op::alias: {
			const expr* t = cur.attr(e.name.substr(1));
			if (!t)
				throw error(MBG_E_MODEL, "Node " + cur.name + " does not have attribute " + e.name.substr(1));
			return interp(m, cur, *t, s, depth + 1);
		}
		case op::lit: return e.value;
		case op::var: return m.variables.at(e.name);
		case op::paren:
		case op::pos: return f(*e.a);
		case op::neg: return -f(*e.a);
		case op::lnot: return !f(*e.a);
		case op::add: return f(*e.a) + f(*e.b);
		case op::sub: return f(*e.a) - f(*e.b);
		case op::mul: return f(*e.a) * f(*e.b);
		case op::div: return f(*e.a) / f(*e.b);
		case op::land: return f(*e.a) && f(*e.b);
		case op::lor: return f(*e.a) || f(*e.b);
		case op::lxor: return (f(*e.a) != 0) != (f(*e.b) != 0);
		case op::eq: return f(*e.a) == f(*e.b);
		case op::ne: return f(*e.a) != f(*e.b);
		case op::lt: return f(*e.a) < f(*e.b);
		case op::le: return f(*e.a) <= f(*e.b);
		case op::gt: return f(*e.a) > f(*e.b);
		case op::ge: return f(*e.a) >= f(*e.b);
		case op::ternary: return f(*e.a) ? f(*e.b) : f(*e.c);
	}
	return 0.f;
}
} // namespace

float model::eval_in_state(const node_t& cur, const expr& e, const uint32_t* state) const
{
	return interp(*this, cur, e, state, 0);
}

void parse_bnd(model& m, const std::string& text, const std::string& fname)
{
	parser p(m, text, fname);
	p.bnd_program();
}

void parse_cfg(model& m, const std::string& text, const std::string& fname)
{
	parser p(m, text, fname);
	p.cfg_program();
}

std::string read_file(const std::string& path)
{
	std::ifstream f(path, std::ios::binary);
	if (!f)
		throw error(MBG_E_IO, "cannot open " + path);
	std::ostringstream ss;
	ss << f.rdbuf();
	std::string s = ss.str();
	// UTF-16 files are not models; give a clear message instead of a lexer error.
	if (s.size() >= 2 && (unsigned char)s[0] == 0xFF && (unsigned char)s[1] == 0xFE)
		throw error(MBG_E_IO, path + ": UTF-16 input is not supported");
	return s;
}

namespace
{
// Mutant nodes of the .cfg: the blocked transition's rate becomes the
// literal 0 and the initial state the forced value, whatever else the
// .cfg said about the node.
void apply_mutants(model& m)
{
	for (node_t& n : m.nodes)
	{
		if (!n.mutant)
			continue;
		const char* blocked = n.mutant > 0 ? "rate_down" : "rate_up";
		for (auto& a : n.attrs)
			if (a.first == blocked)
			{
				a.second = expr::make(op::lit);
				a.second->value = 0.0f;
				a.second->exact = 0.0;
			}
		n.istate = n.mutant > 0 ? 1.0f : 0.0f;
	}
}
} // namespace

std::unique_ptr<model> load_model_strings(const std::string& bnd, const std::string& cfg)
{
	auto m = std::make_unique<model>();
	parse_bnd(*m, bnd, "<bnd>");
	parse_cfg(*m, cfg, "<cfg>");
	apply_mutants(*m);
	validate(*m);
	return m;
}

std::unique_ptr<model> load_model_files(const std::string& bnd_path, const std::string& cfg_path)
{
	auto m = std::make_unique<model>();
	parse_bnd(*m, read_file(bnd_path), bnd_path);
	parse_cfg(*m, read_file(cfg_path), cfg_path);
	apply_mutants(*m);
	validate(*m);
	return m;
}

namespace
{
void validate_expr(const model& m, const node_t& cur, const expr& e, std::set<std::string>& alias_stack)
{
	switch (e.kind)
	{
		case op::lit: return;
		case op::node:
			if (m.node_index(e.name) < 0)
				throw error(MBG_E_MODEL, "unknown node name: " + e.name + " (in node " + cur.name + ")");
			return;
		case op::var:
			if (!m.variables.count(e.name))
				throw error(MBG_E_MODEL, "undefined variable " + e.name + " (in node " + cur.name + ")");
			return;
		case op::alias: {
			std::string an = e.name.substr(1);
			const expr* t = cur.attr(an);
			if (!t)
				throw error(MBG_E_MODEL, "Node " + cur.name + " does not have attribute " + an);
			if (alias_stack.count(an))
				throw error(MBG_E_MODEL, "cyclic alias @" + an + " in node " + cur.name);
			alias_stack.insert(an);
			validate_expr(m, cur, *t, alias_stack);
			alias_stack.erase(an);
			return;
		}
		default: break;
	}
	if (e.a)
		validate_expr(m, cur, *e.a, alias_stack);
	if (e.b)
		validate_expr(m, cur, *e.b, alias_stack);
	if (e.c)
		validate_expr(m, cur, *e.c, alias_stack);
}
} // namespace

void validate(const model& m)
{
	if (m.nodes.empty())
		throw error(MBG_E_MODEL, "model has no nodes");
	for (auto& n : m.nodes)
	{
		for (const char* a : { "rate_up", "rate_down" })
		{
			std::set<std::string> st { a };
			validate_expr(m, n, *n.attr(a), st);
		}
		if (const expr* e = n.attr("is_internal"))
			(void)m.evaluate(*e);
	}
	float tick = m.time_tick(), mt = m.max_time();
	if (!(tick > 0) || !(mt > 0) || !std::isfinite(tick) || !std::isfinite(mt))
		throw error(MBG_E_MODEL, "max_time and time_tick must be positive");
	if (m.sample_count() < 0)
		throw error(MBG_E_MODEL, "sample_count must be non-negative");
}

} // namespace mbg
