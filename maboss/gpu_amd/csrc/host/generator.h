// Per-model HIP source generator (the reference's src/generator.{h,cpp}).
#pragma once

#include <string>
#include <vector>

#include "model.h"

namespace mbg
{

enum class entropy_formula
{
	// H = log2(S) - T/S with S = sum of observed rates and T = sum of
	// r*log2(r), in node order; 0 when at most one observed rate is nonzero.
	// Same quantity as the reference's -sum p*log2(p), one log per step.
	log_sum = 0,
	// The reference's per-node form -sum (r/S)*log2(r/S)
	// (src/generator.cpp:146-175).
	reference = 1
};

struct gen_options
{
	entropy_formula entropy = entropy_formula::log_sum;
	// Waves per SIMD the lane kernels are compiled for (the register budget:
	// 512 / waves VGPRs per lane); 0 lets the compiler choose.
	int lane_waves = 0;
	// Waves per SIMD of the gradient kernel; 0: by parameter count.
	int grad_waves = 0;
};

// What the generator decided about each node; exported for the oracle-side
// tests and for the gradient form.
struct node_info
{
	bool fast_form = false;  // rate_up = @logic ? U : 0 and rate_down = @logic ? 0 : D
	int up_param = -1;       // parameter index when U is a bare $variable
	int down_param = -1;     // parameter index when D is a bare $variable
	int literals = 0;        // literal count of the node's formulas
};

struct generated
{
	std::string source;                   // HIP source for hipRTC
	std::vector<std::string> param_names; // $variables used by rates, sorted
	std::vector<node_info> nodes;
	int total_literals = 0;
};

generated generate(const model& m, const gen_options& opt);

// Gradient form: per-model code of the likelihood-ratio kernel (appended to
// generated::source) and, per node, the parameter its up (2i) and down
// (2i+1) jump uses (0xffff: none).  Every rate must be a bare $variable in
// the fast form, or use no $variable; MBG_E_LIMIT otherwise.
struct grad_code
{
	std::string source;
	std::vector<unsigned short> jparam;
	int n_params = 0;
};
grad_code generate_grad(const model& m, const generated& g);

// Value of rate * log2(rate) as the device computes it (shared float code).
float rate_log2_rate(float r);

// Hexadecimal float literal with an f suffix, exact for every float.
std::string float_literal(float v);

} // namespace mbg
