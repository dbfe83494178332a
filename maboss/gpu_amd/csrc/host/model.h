// Model front-end: expression AST, node table and the .bnd/.cfg parser.
//
// Mirrors the reference's driver IR (src/parser/driver.h:39-55,
// src/parser/parse_types.h:16-29, src/parser/expressions.h:11-120) with a
// hand-written lexer and precedence-climbing parser in place of the
// flex/bison pair (src/parser/scanner.ll, src/parser/parser.yy).
#pragma once

#include <cstdint>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace mbg
{

// Error raised by the parser, the generator and the runtime; the C-ABI turns
// it into a status code plus mbg_last_error().
struct error : std::runtime_error
{
	int code;
	error(int code, const std::string& msg) : std::runtime_error(msg), code(code) {}
};

enum status_code
{
	MBG_OK_ = 0,
	MBG_E_ARG = 1,
	MBG_E_IO = 2,
	MBG_E_PARSE = 3,
	MBG_E_MODEL = 4,
	MBG_E_LIMIT = 5,
	MBG_E_COMPILE = 6,
	MBG_E_HIP = 7,
	MBG_E_OVERFLOW = 8,
	MBG_E_INTERNAL = 9
};

// Operators, same set as src/parser/expressions.h:11-27.
enum class op
{
	lit,
	node,
	var,
	alias,
	paren,
	neg,
	pos,
	lnot,
	add,
	sub,
	mul,
	div,
	land,
	lor,
	lxor,
	eq,
	ne,
	lt,
	le,
	gt,
	ge,
	ternary
};

struct expr;
using expr_ptr = std::unique_ptr<expr>;

struct expr
{
	op kind = op::lit;
	float value = 0.f;      // op::lit
	double exact = 0.0;     // op::lit: value of the literal text in double
	std::string name;       // op::node / op::var ("$x") / op::alias ("@x")
	expr_ptr a, b, c;       // operands

	static expr_ptr make(op k) { auto e = std::make_unique<expr>(); e->kind = k; return e; }
};

struct node_t
{
	std::string name;
	std::vector<std::pair<std::string, expr_ptr>> attrs;
	float istate = 0.5f; // src/parser/parse_types.cpp:3
	// `X.mutant = ON;` / `X.mutant = OFF;` in the .cfg (no counterpart in the
	// reference, whose README.md:33 lists mutants as a TODO): +1 forced on,
	// -1 forced off, 0 free.  Applied after the .cfg is read: forced on sets
	// rate_down to 0 and istate to 1, forced off rate_up to 0 and istate to 0.
	int mutant = 0;

	// First attribute with this name, like node_t::get_attr
	// (src/parser/parse_types.cpp:5-12); nullptr when absent.
	const expr* attr(const std::string& n) const
	{
		for (auto& a : attrs)
			if (a.first == n)
				return a.second.get();
		return nullptr;
	}
};

struct model
{
	std::vector<node_t> nodes;
	std::map<std::string, float> variables; // "$name" -> value
	std::map<std::string, float> constants; // name -> value (float, as the reference)
	std::map<std::string, double> constants_exact; // literal text value (seed, sample_count)

	model();

	int node_index(const std::string& n) const;
	bool is_internal(size_t i) const; // src/parser/parse_types.cpp:20-22
	int n_nodes() const { return (int)nodes.size(); }
	int n_words() const { return ((int)nodes.size() + 31) / 32; }
	int n_observed() const;
	std::vector<int> observed_nodes() const;

	bool discrete_time() const;
	float max_time() const;
	float time_tick() const;
	int64_t sample_count() const;
	uint64_t seed() const;
	int windows() const; // ceil(max_time / tick) in float, src/statistics/window_average_small.cpp:26

	// Constant evaluation with the reference's float semantics
	// (src/parser/expressions.cpp:10-23, :50-81, :156-159).
	float evaluate(const expr& e) const;

	// Value of e in a state (node references read the state's bits, @aliases
	// resolve against node cur), with the same float semantics.
	float eval_in_state(const node_t& cur, const expr& e, const uint32_t* state) const;
};

// Parse a .bnd text then a .cfg text into a model (src/parser/driver.cpp:16-31).
void parse_bnd(model& m, const std::string& text, const std::string& fname);
void parse_cfg(model& m, const std::string& text, const std::string& fname);
std::unique_ptr<model> load_model_files(const std::string& bnd_path, const std::string& cfg_path);
std::unique_ptr<model> load_model_strings(const std::string& bnd, const std::string& cfg);

// Checks every node can be compiled: rate_up/rate_down present, identifiers
// name nodes, $variables are defined, @aliases resolve without cycles.
void validate(const model& m);

std::string read_file(const std::string& path);

} // namespace mbg
