// C-ABI of the engine (declared in include/maboss_gpu.h).  Every entry point
// catches exceptions and turns them into a status code plus a thread-local
// message; nothing here calls exit() (the reference exits on any CUDA error,
// src/utils.cpp:3-30).
#include <algorithm>
#include <cstring>
#include <fstream>
#include <functional>
#include <iostream>

#include "../../../../include/maboss_gpu.h"
#include "engine.h"

static_assert(mbg::k_noalloc_scratch_bytes == MBG_NOALLOC_SCRATCH_BYTES, "scratch size in the header");

struct mbg_model
{
	std::shared_ptr<mbg::model> m;
	std::vector<std::string> params; // rate parameters, as the generator orders them
};

struct mbg_program
{
	std::unique_ptr<mbg::program> p;
};

struct mbg_result
{
	std::unique_ptr<mbg::result> r;
};

namespace
{

thread_local std::string g_last_error;

template <typename F>
int guarded(F&& f)
{
	try
	{
		g_last_error.clear();
		f();
		return MBG_OK;
	}
	catch (const mbg::error& e)
	{
		g_last_error = e.what();
		return e.code;
	}
	catch (const std::bad_alloc&)
	{
		g_last_error = "host out of memory";
		return MBG_ERR_LIMIT;
	}
	catch (const std::exception& e)
	{
		g_last_error = e.what();
		return MBG_ERR_INTERNAL;
	}
}

void refresh_params(mbg_model* mm)
{
	mbg::gen_options o;
	mm->params = mbg::generate(*mm->m, o).param_names;
}

mbg::gen_options to_gen(const mbg_compile_opts* o)
{
	mbg::gen_options g;
	if (o && o->entropy_formula == 1)
		g.entropy = mbg::entropy_formula::reference;
	if (o && o->lane_occupancy > 0)
		g.lane_waves = std::min(o->lane_occupancy, 8);
	if (o && o->grad_occupancy > 0)
		g.grad_waves = std::min(o->grad_occupancy, 8);
	return g;
}

} // namespace

extern "C" {

const char* mbg_last_error(void) { return g_last_error.c_str(); }

const char* mbg_version(void) { return "maboss-gpu-amd 0.1 (gfx950)"; }

int mbg_model_load(const char* bnd_path, const char* cfg_path, mbg_model** out)
{
	return guarded([&] {
		if (!bnd_path || !cfg_path || !out)
			throw mbg::error(mbg::MBG_E_ARG, "null argument");
		auto mm = std::make_unique<mbg_model>();
		mm->m = std::shared_ptr<mbg::model>(mbg::load_model_files(bnd_path, cfg_path).release());
		refresh_params(mm.get());
		*out = mm.release();
	});
}

int mbg_model_load_strings(const char* bnd_text, const char* cfg_text, mbg_model** out)
{
	return guarded([&] {
		if (!bnd_text || !cfg_text || !out)
			throw mbg::error(mbg::MBG_E_ARG, "null argument");
		auto mm = std::make_unique<mbg_model>();
		mm->m = std::shared_ptr<mbg::model>(mbg::load_model_strings(bnd_text, cfg_text).release());
		refresh_params(mm.get());
		*out = mm.release();
	});
}

void mbg_model_free(mbg_model* m) { delete m; }

int mbg_model_get_info(const mbg_model* mm, mbg_model_info* out)
{
	return guarded([&] {
		if (!mm || !out)
			throw mbg::error(mbg::MBG_E_ARG, "null argument");
		const mbg::model& m = *mm->m;
		std::memset(out, 0, sizeof(*out));
		out->n_nodes = m.n_nodes();
		out->n_words = m.n_words();
		out->n_observed = m.n_observed();
		out->windows = m.windows();
		out->discrete_time = m.discrete_time();
		out->max_time = m.max_time();
		out->time_tick = m.time_tick();
		out->sample_count = m.sample_count();
		out->seed = m.seed();
		out->n_params = (int32_t)mm->params.size();
		mbg::gen_options o;
		out->total_literals = mbg::generate(m, o).total_literals;
	});
}

const char* mbg_model_node_name(const mbg_model* mm, int32_t i)
{
	if (!mm || i < 0 || i >= mm->m->n_nodes())
		return nullptr;
	return mm->m->nodes[i].name.c_str();
}

int mbg_model_node_internal(const mbg_model* mm, int32_t i)
{
	if (!mm || i < 0 || i >= mm->m->n_nodes())
		return -1;
	try
	{
		return mm->m->is_internal((size_t)i) ? 1 : 0;
	}
	catch (...)
	{
		return -1;
	}
}

int mbg_model_set_constant(mbg_model* mm, const char* name, double value)
{
	return guarded([&] {
		if (!mm || !name)
			throw mbg::error(mbg::MBG_E_ARG, "null argument");
		mm->m->constants[name] = (float)value;
		mm->m->constants_exact[name] = value;
		mbg::validate(*mm->m);
	});
}

int mbg_model_set_variable(mbg_model* mm, const char* name, double value)
{
	return guarded([&] {
		if (!mm || !name)
			throw mbg::error(mbg::MBG_E_ARG, "null argument");
		std::string n = name[0] == '$' ? name : std::string("$") + name;
		mm->m->variables[n] = (float)value;
		refresh_params(mm);
	});
}

const char* mbg_model_param_name(const mbg_model* mm, int32_t k)
{
	if (!mm || k < 0 || k >= (int32_t)mm->params.size())
		return nullptr;
	return mm->params[k].c_str();
}

float mbg_model_param_value(const mbg_model* mm, int32_t k)
{
	if (!mm || k < 0 || k >= (int32_t)mm->params.size())
		return 0.f;
	return mm->m->variables.at(mm->params[k]);
}

int mbg_model_eval_rate(const mbg_model* mm, int32_t node, const uint32_t* s, float* rate)
{
	return guarded([&] {
		if (!mm || !s || !rate || node < 0 || node >= mm->m->n_nodes())
			throw mbg::error(mbg::MBG_E_ARG, "bad argument");
		const mbg::model& m = *mm->m;
		const mbg::node_t& n = m.nodes[node];
		const bool on = (s[node / 32] >> (node % 32)) & 1u;
		*rate = m.eval_in_state(n, *n.attr(on ? "rate_down" : "rate_up"), s);
	});
}

int mbg_generate_code(const mbg_model* mm, const mbg_compile_opts* opts, char** source)
{
	return guarded([&] {
		if (!mm || !source)
			throw mbg::error(mbg::MBG_E_ARG, "null argument");
		mbg::generated g = mbg::generate(*mm->m, to_gen(opts));
		std::string s = g.source;
		if (opts && opts->with_gradient)
			s += mbg::generate_grad(*mm->m, g).source + "\n#include \"mbg_grad_kernel.h\"\n";
		char* c = (char*)std::malloc(s.size() + 1);
		if (!c)
			throw std::bad_alloc();
		std::memcpy(c, s.c_str(), s.size() + 1);
		*source = c;
	});
}

void mbg_free_string(char* s) { std::free(s); }

int mbg_program_compile(const mbg_model* mm, const mbg_compile_opts* opts, mbg_program** out)
{
	return guarded([&] {
		if (!mm || !out)
			throw mbg::error(mbg::MBG_E_ARG, "null argument");
		// the program keeps its own copy of the model
		auto copy = std::make_shared<mbg::model>();
		{
			// re-parse is not possible (text not kept); deep-copy the IR
			std::function<mbg::expr_ptr(const mbg::expr*)> clone = [&](const mbg::expr* e) -> mbg::expr_ptr {
				if (!e)
					return nullptr;
				auto c = mbg::expr::make(e->kind);
				c->value = e->value;
				c->exact = e->exact;
				c->name = e->name;
				c->a = clone(e->a.get());
				c->b = clone(e->b.get());
				c->c = clone(e->c.get());
				return c;
			};
			copy->variables = mm->m->variables;
			copy->constants = mm->m->constants;
			copy->constants_exact = mm->m->constants_exact;
			for (auto& n : mm->m->nodes)
			{
				mbg::node_t nn;
				nn.name = n.name;
				nn.istate = n.istate;
				for (auto& a : n.attrs)
					nn.attrs.emplace_back(a.first, clone(a.second.get()));
				copy->nodes.push_back(std::move(nn));
			}
		}
		auto pp = std::make_unique<mbg_program>();
		pp->p = mbg::compile_program(copy, to_gen(opts), opts && opts->no_cache);
		*out = pp.release();
	});
}

void mbg_program_free(mbg_program* p) { delete p; }

double mbg_program_compile_seconds(const mbg_program* p) { return p ? p->p->compile_seconds : 0.0; }

void mbg_sim_params_default(mbg_sim_params* p)
{
	if (!p)
		return;
	std::memset(p, 0, sizeof(*p));
	p->n_trajectories = -1;
	p->seed = -1;
	p->window_mode = -1;
	p->final_mode = -1;
	p->kernel = -1;
}

static mbg::sim_params to_sim_params(mbg_program* p, const mbg_sim_params* params)
{
	mbg::sim_params sp;
	if (params)
	{
		sp.n_traj = params->n_trajectories;
		sp.first_id = params->first_id;
		sp.seed = params->seed;
		sp.variant = params->variant;
		sp.max_steps = params->max_steps;
		sp.window_mode = params->window_mode;
		sp.final_mode = params->final_mode;
		sp.window_capacity = params->window_capacity;
		sp.final_capacity = params->final_capacity;
		sp.fixed_capacity = params->fixed_capacity;
		sp.block_size = params->block_size;
		sp.blocks_per_cu = params->blocks_per_cu;
		sp.normalization = params->normalization;
		sp.device_only = (params->flags & MBG_SIM_DEVICE_ONLY) != 0;
		sp.errors = (params->flags & MBG_SIM_ERRORS) != 0;
		sp.moment_states = params->moment_states;
		sp.refill_lanes = params->refill_lanes;
		sp.kernel = params->kernel;
		if (params->rates)
			sp.rates.assign(params->rates, params->rates + p->p->gen.param_names.size());
	}
	if (sp.block_size < 0 || sp.block_size > 1024 || (sp.block_size % 64) != 0)
		throw mbg::error(mbg::MBG_E_ARG, "block_size must be a multiple of 64 up to 1024");
	return sp;
}

static int simulate_entry(mbg_program* p, const mbg_sim_params* params, void* stream, mbg_result** out, bool async)
{
	return guarded([&] {
		if (!p || !out)
			throw mbg::error(mbg::MBG_E_ARG, "null argument");
		const mbg::sim_params sp = to_sim_params(p, params);
		if (sp.errors && sp.device_only)
			throw mbg::error(mbg::MBG_E_ARG, "MBG_SIM_ERRORS needs the statistics on the host");
		if (sp.moment_states != 0 && sp.moment_states != 1 && sp.moment_states != 8)
			throw mbg::error(mbg::MBG_E_ARG, "moment_states must be 0, 1 or 8");
		auto rr = std::make_unique<mbg_result>();
		// with error columns the run completes first: the replay reads nothing
		// of it but its trajectory count, and the result carries both
		rr->r = async && !sp.errors ? mbg::simulate_async(*p->p, sp, (hipStream_t)stream)
									: mbg::simulate(*p->p, sp, (hipStream_t)stream);
		if (sp.errors)
			mbg::run_moments(*p->p, sp, (hipStream_t)stream, *rr->r);
		*out = rr.release();
	});
}

int mbg_simulate(mbg_program* p, const mbg_sim_params* params, void* stream, mbg_result** out)
{
	return simulate_entry(p, params, stream, out, false);
}

int mbg_simulate_async(mbg_program* p, const mbg_sim_params* params, void* stream, mbg_result** out)
{
	return simulate_entry(p, params, stream, out, true);
}

int mbg_window_probs_device(mbg_program* p, const mbg_sim_params* params, const float* rates, double* probs,
							void* stream)
{
	return guarded([&] {
		if (!p || !probs)
			throw mbg::error(mbg::MBG_E_ARG, "null argument");
		if (params && params->rates)
			throw mbg::error(mbg::MBG_E_ARG, "pass the rates as the device pointer argument");
		const mbg::sim_params sp = to_sim_params(p, params);
		mbg::window_probs_device(*p->p, sp, rates, probs, (hipStream_t)stream);
	});
}

void mbg_batch_params_default(mbg_batch_params* p)
{
	if (!p)
		return;
	std::memset(p, 0, sizeof(*p));
	p->seed = -1;
	p->use_graph = 1;
}

int mbg_simulate_batched(mbg_program* p, const mbg_batch_params* params, int32_t n_variants,
						 const mbg_variant* variants, void* stream, mbg_result** results, mbg_batch_info* info)
{
	return guarded([&] {
		if (!p || n_variants < 0 || (n_variants && (!variants || !results)))
			throw mbg::error(mbg::MBG_E_ARG, "bad argument");
		mbg::batch_params bp;
		if (params)
		{
			bp.seed = params->seed;
			bp.max_steps = params->max_steps;
			bp.block = params->block_size;
			bp.per_cu = params->blocks_per_cu;
			bp.tpl = params->traj_per_lane;
			bp.lds_cells = params->lds_cells;
			bp.graph = params->use_graph != 0;
			bp.device_only = (params->flags & MBG_SIM_DEVICE_ONLY) != 0;
		}
		const int np = (int)p->p->gen.param_names.size(), N = p->p->mdl->n_nodes();
		std::vector<mbg::variant_spec> vs(n_variants);
		for (int v = 0; v < n_variants; v++)
		{
			vs[v].n_traj = variants[v].n_trajectories;
			vs[v].max_time = variants[v].max_time;
			if (variants[v].rates)
				vs[v].rates.assign(variants[v].rates, variants[v].rates + np);
			if (variants[v].mutations)
				vs[v].mut.assign(variants[v].mutations, variants[v].mutations + N);
		}
		mbg::batch_timing bt;
		auto rs = mbg::simulate_batched(*p->p, bp, vs, (hipStream_t)stream, &bt);
		for (int v = 0; v < n_variants; v++)
		{
			auto rr = std::make_unique<mbg_result>();
			rr->r = std::move(rs[v]);
			results[v] = rr.release();
		}
		if (info)
		{
			std::memset(info, 0, sizeof(*info));
			info->kernel_ms = bt.kernel_ms;
			info->total_ms = bt.total_ms;
			info->transitions = bt.transitions;
			info->block_size = bt.block;
			info->blocks_per_cu = bt.per_cu;
			info->traj_per_lane = bt.tpl;
			info->used_graph = bt.graph ? 1 : 0;
			info->lds_cells = bt.lds_cells;
			info->tuned_cached = bt.tuned_cached ? 1 : 0;
		}
	});
}

void mbg_grad_params_default(mbg_grad_params* p)
{
	if (!p)
		return;
	std::memset(p, 0, sizeof(*p));
	p->n_trajectories = -1;
	p->seed = -1;
}

int mbg_grad_check(mbg_program* p)
{
	return guarded([&] {
		if (!p)
			throw mbg::error(mbg::MBG_E_ARG, "null program");
		(void)p->p->grad();
		// load what a gradient launch needs, so later launches allocate nothing
		(void)p->p->grad_function();
		p->p->load_util();
		(void)p->p->grad_tables();
	});
}

int mbg_backward(mbg_program* p, const mbg_grad_params* params, const float* rates, const double* grad_probs,
				 double* grad_params, void* stream, mbg_grad_info* info)
{
	return guarded([&] {
		if (!p || !grad_probs || !grad_params)
			throw mbg::error(mbg::MBG_E_ARG, "null argument");
		mbg::grad_params gp;
		if (params)
		{
			gp.n_traj = params->n_trajectories;
			gp.first_id = params->first_id;
			gp.seed = params->seed;
			gp.variant = params->variant;
			gp.max_steps = params->max_steps;
			gp.block = params->block_size;
			gp.per_cu = params->blocks_per_cu;
			gp.fold_lanes = params->fold_lanes;
		}
		mbg::grad_timing gt;
		mbg::simulate_grad(*p->p, gp, rates, grad_probs, grad_params, (hipStream_t)stream, info ? &gt : nullptr);
		if (info)
		{
			std::memset(info, 0, sizeof(*info));
			info->kernel_ms = gt.kernel_ms;
			info->transitions = gt.transitions;
			info->block_size = gt.block;
			info->grid = gt.grid;
		}
	});
}

int mbg_backward_noalloc(mbg_program* p, const mbg_grad_params* params, const float* rates,
						 const double* grad_probs, double* grad_params, void* stream)
{
	return guarded([&] {
		if (!p || !grad_probs || !grad_params)
			throw mbg::error(mbg::MBG_E_ARG, "null argument");
		mbg::grad_params gp;
		if (params)
		{
			gp.n_traj = params->n_trajectories;
			gp.first_id = params->first_id;
			gp.seed = params->seed;
			gp.variant = params->variant;
			gp.max_steps = params->max_steps;
			gp.block = params->block_size;
			gp.per_cu = params->blocks_per_cu;
			gp.fold_lanes = params->fold_lanes;
		}
		void* s = mbg::noalloc_scratch(p->p->device);
		if (!s)
			throw mbg::error(mbg::MBG_E_ARG, "the no-allocation scratch is not set up on the program's device");
		mbg::simulate_grad(*p->p, gp, rates, grad_probs, grad_params, (hipStream_t)stream, nullptr, s);
	});
}

void mbg_result_free(mbg_result* r) { delete r; }

int mbg_result_wait(mbg_result* rr)
{
	return guarded([&] {
		if (!rr)
			throw mbg::error(mbg::MBG_E_ARG, "null argument");
		rr->r->wait();
	});
}

int mbg_result_get_info(const mbg_result* rr, mbg_result_info* out)
{
	return guarded([&] {
		if (!rr || !out)
			throw mbg::error(mbg::MBG_E_ARG, "null argument");
		const mbg::result& r = *rr->r;
		r.wait();
		std::memset(out, 0, sizeof(*out));
		out->windows = r.windows;
		out->n_observed = r.n_obs;
		out->n_words = r.n_words;
		out->discrete_time = r.discrete;
		out->time_tick = r.tick;
		out->max_time = r.max_time;
		out->normalization = r.norm;
		out->n_trajectories = r.n_traj;
		out->transitions = r.transitions;
		out->finished = r.finished;
		out->truncated = r.truncated;
		out->n_fixed_points = r.n_fixed();
		out->n_final_states = r.fin.n;
		out->n_window_cells = r.n_cells();
		out->kernel_ms = r.kernel_ms;
		out->total_ms = r.total_ms;
		out->simulate_ms = r.simulate_ms;
		out->window_pairs = r.window_pairs;
		out->moment_spills = r.moment_spills;
		out->window_spills = r.window_spills;
		out->window_record_bytes = r.window_record_bytes;
		out->window_escapes = r.window_escapes;
		out->has_errors = r.has_moments ? 1 : 0;
		out->window_mode = r.window_mode;
		out->final_mode = r.final_mode;
		out->window_capacity = r.window_capacity;
		out->kernel = r.kernel;
	});
}

int64_t mbg_result_window_size(const mbg_result* rr, int32_t w)
{
	if (!rr || w < 0 || w >= rr->r->windows)
		return -1;
	try
	{
		rr->r->wait();
	}
	catch (const std::exception& e)
	{
		g_last_error = e.what();
		return -1;
	}
	return (int64_t)(rr->r->woff[w + 1] - rr->r->woff[w]);
}

int mbg_result_window(const mbg_result* rr, int32_t w, uint64_t* states, uint64_t* raw)
{
	return guarded([&] {
		if (!rr || w < 0 || w >= rr->r->windows)
			throw mbg::error(mbg::MBG_E_ARG, "bad window");
		const mbg::result& r = *rr->r;
		r.wait();
		const uint64_t om = r.obs_mask();
		size_t i = 0;
		for (uint64_t c = r.woff[w]; c < r.woff[w + 1]; c++, i++)
		{
			if (states)
				states[i] = r.win.k(c) & om;
			if (raw)
				raw[i] = r.win.v(c);
		}
	});
}

int mbg_result_window_probs(const mbg_result* rr, int32_t w, uint64_t* states, double* probs)
{
	return guarded([&] {
		if (!rr || w < 0 || w >= rr->r->windows)
			throw mbg::error(mbg::MBG_E_ARG, "bad window");
		const mbg::result& r = *rr->r;
		r.wait();
		const uint64_t om = r.obs_mask();
		size_t i = 0;
		for (uint64_t c = r.woff[w]; c < r.woff[w + 1]; c++, i++)
		{
			if (states)
				states[i] = r.win.k(c) & om;
			if (probs)
				probs[i] = r.prob(r.win.v(c));
		}
	});
}

int mbg_result_entropies(const mbg_result* rr, double* th, double* h)
{
	return guarded([&] {
		if (!rr)
			throw mbg::error(mbg::MBG_E_ARG, "null argument");
		rr->r->wait();
		for (int w = 0; w < rr->r->windows; w++)
		{
			if (th)
				th[w] = rr->r->th(w);
			if (h)
				h[w] = rr->r->h(w);
		}
	});
}

int mbg_result_save(const mbg_result* rr, const char* path)
{
	return guarded([&] {
		if (!rr || !path)
			throw mbg::error(mbg::MBG_E_ARG, "null argument");
		mbg::save_result(*rr->r, path);
	});
}

int mbg_result_load(const mbg_program* p, const char* path, mbg_result** out)
{
	return guarded([&] {
		if (!p || !path || !out)
			throw mbg::error(mbg::MBG_E_ARG, "null argument");
		auto rr = std::make_unique<mbg_result>();
		rr->r = mbg::load_result(*p->p, path);
		*out = rr.release();
	});
}

int mbg_result_merge(mbg_result* into, const mbg_result* other)
{
	return guarded([&] {
		if (!into || !other || into == other)
			throw mbg::error(mbg::MBG_E_ARG, "bad argument");
		mbg::merge_into(*into->r, *other->r);
	});
}

int mbg_simulate_checkpointed(mbg_program* p, const mbg_sim_params* params, uint64_t chunk, const char* path,
							  void* stream, mbg_result** out)
{
	return guarded([&] {
		if (!p || !path || !out)
			throw mbg::error(mbg::MBG_E_ARG, "null argument");
		const mbg::sim_params sp = to_sim_params(p, params);
		auto rr = std::make_unique<mbg_result>();
		rr->r = mbg::simulate_checkpointed(*p->p, sp, chunk, path, (hipStream_t)stream);
		*out = rr.release();
	});
}

int mbg_result_error_th(const mbg_result* rr, double* err_th)
{
	return guarded([&] {
		if (!rr || !err_th)
			throw mbg::error(mbg::MBG_E_ARG, "null argument");
		rr->r->wait();
		for (int w = 0; w < rr->r->windows; w++)
			err_th[w] = rr->r->error_th(w);
	});
}

int mbg_result_window_errors(const mbg_result* rr, int32_t w, double* err)
{
	return guarded([&] {
		if (!rr || !err || w < 0 || w >= rr->r->windows)
			throw mbg::error(mbg::MBG_E_ARG, "bad window");
		const mbg::result& r = *rr->r;
		r.wait();
		const uint64_t om = r.obs_mask();
		size_t i = 0;
		for (uint64_t c = r.woff[w]; c < r.woff[w + 1]; c++, i++)
			err[i] = r.error_prob(w, r.win.k(c) & om, r.win.v(c));
	});
}

int64_t mbg_result_final_size(const mbg_result* rr)
{
	if (!rr)
		return -1;
	try
	{
		rr->r->wait();
	}
	catch (const std::exception& e)
	{
		g_last_error = e.what();
		return -1;
	}
	return (int64_t)rr->r->fin.n;
}

int mbg_result_final(const mbg_result* rr, uint64_t* states, uint64_t* counts)
{
	return guarded([&] {
		if (!rr)
			throw mbg::error(mbg::MBG_E_ARG, "null argument");
		const mbg::result& r = *rr->r;
		r.wait();
		for (uint64_t i = 0; i < r.fin.n; i++)
		{
			if (states)
				states[i] = r.fin.k(i);
			if (counts)
				counts[i] = r.fin.v(i);
		}
	});
}

int64_t mbg_result_fixed_size(const mbg_result* rr)
{
	if (!rr)
		return -1;
	try
	{
		rr->r->wait();
	}
	catch (const std::exception& e)
	{
		g_last_error = e.what();
		return -1;
	}
	return (int64_t)rr->r->n_fixed();
}

int mbg_result_fixed(const mbg_result* rr, uint32_t* states, uint64_t* counts)
{
	return guarded([&] {
		if (!rr)
			throw mbg::error(mbg::MBG_E_ARG, "null argument");
		const mbg::result& r = *rr->r;
		r.wait();
		const size_t W = (size_t)r.n_words;
		for (uint64_t i = 0; i < r.n_fixed(); i++)
		{
			if (states)
				std::memcpy(states + i * W, r.fixed_state(i), W * 4);
			if (counts)
				counts[i] = r.fp_count[i];
		}
	});
}

int mbg_result_write_csv(const mbg_result* rr, const char* prefix, int with_finals)
{
	return guarded([&] {
		if (!rr || !prefix)
			throw mbg::error(mbg::MBG_E_ARG, "null argument");
		rr->r->wait();
		rr->r->write_csv(prefix, with_finals != 0);
	});
}

int mbg_result_print(const mbg_result* rr, const char* path)
{
	return guarded([&] {
		if (!rr)
			throw mbg::error(mbg::MBG_E_ARG, "null argument");
		rr->r->wait();
		if (!path)
		{
			rr->r->print(std::cout);
			return;
		}
		std::ofstream f(path);
		if (!f)
			throw mbg::error(mbg::MBG_E_IO, std::string("cannot write ") + path);
		rr->r->print(f);
	});
}

int mbg_result_write_raw(const mbg_result* rr, const char* path)
{
	return guarded([&] {
		if (!rr || !path)
			throw mbg::error(mbg::MBG_E_ARG, "null argument");
		std::ofstream f(path);
		if (!f)
			throw mbg::error(mbg::MBG_E_IO, std::string("cannot write ") + path);
		rr->r->wait();
		rr->r->write_raw(f);
	});
}

} // extern "C"
