// Kernel compiler: the generated model code and the kernel source (embedded
// at build time) are compiled together by hipRTC into one code object per
// (model, flavour), loaded as a module.  Replaces
// kernel_compiler::compile_simulation (src/kernel_compiler.cpp:61-164), which
// used NVRTC (-dlto -rdc) + nvJitLink against embedded LTO fatbins.
//
// One translation unit lets the generated rate code inline into the kernel
// loop and needs no bitcode from this build, so any hipRTC / code-object
// manager the process has loaded (the system ROCm's, or one a framework
// bundles) can compile it.  Code objects are cached on disk, keyed by a hash
// of the complete source, the compile options and the compiler and runtime
// versions; a cached file that does not load is compiled again.
#include <hip/hiprtc.h>


#include <cstdio>
#include <iterator>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <sys/stat.h>
#include <unistd.h>

#include "engine.h"
#include "timer.h"

extern "C" {
extern const char mbg_hdr_numeric[];
extern const char mbg_hdr_abi[];
extern const char mbg_hdr_kernel_args[];
extern const char mbg_hdr_stats_device[];
extern const char mbg_hdr_lane_core[];
extern const char mbg_hdr_grad_kernel[];
extern const char mbg_hdr_moments_kernel[];
extern const char mbg_src_lane[];
extern const unsigned char mbg_co_wave[];
extern const unsigned char mbg_co_util[];
}

namespace mbg
{

void hip_check(hipError_t e, const char* what)
{
	if (e != hipSuccess)
		throw error(MBG_E_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

namespace
{

const char* const k_compile_flags[] = { "--offload-arch=gfx950",
										"-O3",
										"-std=c++17",
										"-ffp-contract=off",
										"-fno-gpu-flush-denormals-to-zero",
										"-fhip-fp32-correctly-rounded-divide-sqrt",
										"-Xclang",
										"-target-feature",
										"-Xclang",
										"-packed-fp32-ops" };

uint64_t fnv1a(const void* data, size_t n, uint64_t h = 1469598103934665603ull)
{
	const unsigned char* p = (const unsigned char*)data;
	for (size_t i = 0; i < n; i++)
	{
		h ^= p[i];
		h *= 1099511628211ull;
	}
	return h;
}

std::string cache_dir()
{
	if (const char* d = std::getenv("MBG_CACHE_DIR"))
		return d;
	if (const char* h = std::getenv("HOME"))
		return std::string(h) + "/.cache/maboss_gpu_amd";
	return "/tmp/maboss_gpu_amd_cache";
}

void mkdirs(const std::string& path)
{
	std::string cur;
	std::stringstream ss(path);
	std::string part;
	if (!path.empty() && path[0] == '/')
		cur = "/";
	while (std::getline(ss, part, '/'))
	{
		if (part.empty())
			continue;
		cur += part + "/";
		::mkdir(cur.c_str(), 0755);
	}
}

bool read_all(const std::string& path, std::vector<char>& out)
{
	std::ifstream f(path, std::ios::binary);
	if (!f)
		return false;
	out.assign(std::istreambuf_iterator<char>(f), std::istreambuf_iterator<char>());
	return !out.empty();
}

void write_atomic(const std::string& path, const void* data, size_t n)
{
	std::string tmp = path + ".tmp" + std::to_string(::getpid());
	{
		std::ofstream f(tmp, std::ios::binary);
		if (!f)
			return;
		f.write((const char*)data, (std::streamsize)n);
		if (!f)
			return;
	}
	std::rename(tmp.c_str(), path.c_str());
}

void rtc_check(hiprtcResult r, const char* what)
{
	if (r != HIPRTC_SUCCESS)
		throw error(MBG_E_COMPILE, std::string(what) + ": " + hiprtcGetErrorString(r));
}

// Compile one translation unit (generated model code + kernel source, with
// the device headers supplied by name) straight to a code object.
std::vector<char> compile_to_code_object(const std::string& source)
{
	const char* headers[] = { mbg_hdr_numeric,	  mbg_hdr_abi,		 mbg_hdr_kernel_args,	mbg_hdr_stats_device,
							  mbg_hdr_lane_core, mbg_hdr_grad_kernel, mbg_hdr_moments_kernel };
	const char* names[] = { "mbg_numeric.h",	 "mbg_abi.h",		  "mbg_kernel_args.h",		"mbg_stats_device.h",
							"mbg_lane_core.h", "mbg_grad_kernel.h", "mbg_moments_kernel.h" };
	std::vector<const char*> flags(std::begin(k_compile_flags), std::end(k_compile_flags));
	hiprtcProgram prog;
	rtc_check(hiprtcCreateProgram(&prog, source.c_str(), "mbg_unit.hip", 7, headers, names), "hiprtcCreateProgram");
	hiprtcResult r = hiprtcCompileProgram(prog, (int)flags.size(), flags.data());
	size_t log_size = 0;
	hiprtcGetProgramLogSize(prog, &log_size);
	std::string log(log_size, '\0');
	if (log_size)
		hiprtcGetProgramLog(prog, &log[0]);
	if (r != HIPRTC_SUCCESS)
	{
		hiprtcDestroyProgram(&prog);
		throw error(MBG_E_COMPILE, std::string("hipRTC compile failed: ") + hiprtcGetErrorString(r) + "\n" + log);
	}
	size_t n = 0;
	rtc_check(hiprtcGetCodeSize(prog, &n), "hiprtcGetCodeSize");
	std::vector<char> co(n);
	rtc_check(hiprtcGetCode(prog, co.data()), "hiprtcGetCode");
	hiprtcDestroyProgram(&prog);
	return co;
}

// Hash of a translation unit for the code-object cache: the full source, the
// compile flags, the embedded headers, and the versions of hipRTC and of the
// HIP runtime (a code object from another compiler is never reused).
uint64_t unit_hash(const std::string& src, const std::vector<const char*>& headers)
{
	uint64_t h = fnv1a(src.data(), src.size());
	for (const char* f : k_compile_flags)
		h = fnv1a(f, std::strlen(f), h);
	for (const char* hd : headers)
		h = fnv1a(hd, std::strlen(hd), h);
	int v[3] = { 0, 0, 0 };
	(void)hiprtcVersion(&v[0], &v[1]);
	(void)hipRuntimeGetVersion(&v[2]);
	return fnv1a(v, sizeof(v), h);
}

// The module of a translation unit: loaded from the cache file `name` when it
// is there (a file that does not load, e.g. truncated, is removed and the
// unit compiled again), else compiled and stored.
hipModule_t load_unit(const std::string& src, const std::string& name, bool no_cache, const char* what)
{
	const std::string dir = cache_dir();
	const std::string path = dir + "/" + name;
	std::vector<char> co;
	hipModule_t mod = nullptr;
	if (!no_cache && read_all(path, co))
	{
		if (hipModuleLoadData(&mod, co.data()) == hipSuccess)
			return mod;
		(void)hipGetLastError();
		std::remove(path.c_str());
		mod = nullptr;
	}
	co = compile_to_code_object(src);
	if (!no_cache)
	{
		mkdirs(dir);
		write_atomic(path, co.data(), co.size());
	}
	hip_check(hipModuleLoadData(&mod, co.data()), what);
	return mod;
}

} // namespace

program::~program()
{
	// the copies in flight read the slots' buffers
	drain_copies();
	if (copy_stream)
		(void)hipStreamDestroy(copy_stream);
	if (run_done)
		(void)hipEventDestroy(run_done);
	for (auto& by_win : mods)
		for (auto& by_fin : by_win)
			for (hipModule_t m : by_fin)
				if (m)
					(void)hipModuleUnload(m);
	if (mod_wave)
		(void)hipModuleUnload(mod_wave);
	if (batch_graph)
		(void)hipGraphExecDestroy(batch_graph);
	if (mod_util)
		(void)hipModuleUnload(mod_util);
	if (mod_grad)
		(void)hipModuleUnload(mod_grad);
	for (hipModule_t m : mod_mom)
		if (m)
			(void)hipModuleUnload(m);
}

const grad_code& program::grad()
{
	if (!gcode)
		gcode = std::make_unique<grad_code>(generate_grad(*mdl, gen));
	return *gcode;
}

hipFunction_t program::grad_function()
{
	if (fn_grad)
		return fn_grad;
	const grad_code& gc = grad();
	scope_timer tm("compiler> gradient kernel");
	const double t0 = wall_seconds();
	// Occupancy: the T accumulators are doubles, 2 registers per parameter,
	// next to about 170 for the rest of the kernel; 2 waves per SIMD (256
	// registers) up to 64 parameters, 1 wave (512 registers) above.  Measured
	// on grad_30 (60 parameters): 2 waves 3.6 ms (a few spilled registers),
	// 1 wave 5.3 ms, 3 and 4 waves 10-13 ms (most of T spilled).
	const int waves = opts.grad_waves > 0 ? opts.grad_waves : gc.n_params <= 64 ? 2 : 1;
	const std::string src = "#define MBG_GRAD_WAVES " + std::to_string(waves) + "\n" + gen.source + gc.source
							+ "\n#include \"mbg_grad_kernel.h\"\n";
	const uint64_t h = unit_hash(src, { mbg_hdr_numeric, mbg_hdr_abi, mbg_hdr_kernel_args, mbg_hdr_stats_device,
										mbg_hdr_lane_core, mbg_hdr_grad_kernel });
	char name[64];
	std::snprintf(name, sizeof(name), "%016llx_g.co", (unsigned long long)h);
	mod_grad = load_unit(src, name, no_cache, "hipModuleLoadData(simulate_grad)");
	hip_check(hipModuleGetFunction(&fn_grad, mod_grad, "simulate_grad"), "hipModuleGetFunction(simulate_grad)");
	compile_seconds += wall_seconds() - t0;
	return fn_grad;
}

hipFunction_t program::moments_function(int states_per_lane)
{
	const int slot = states_per_lane == 1 ? 1 : 0;
	if (fn_mom[slot])
		return fn_mom[slot];
	scope_timer tm("compiler> moments kernel");
	const double t0 = wall_seconds();
	const std::string src = "#define MBG_MOM_K " + std::to_string(slot ? 1 : 8) + "\n"
							+ gen.source + "\n#include \"mbg_moments_kernel.h\"\n";
	const uint64_t h = unit_hash(src, { mbg_hdr_numeric, mbg_hdr_abi, mbg_hdr_kernel_args, mbg_hdr_stats_device,
										mbg_hdr_lane_core, mbg_hdr_moments_kernel });
	char name[64];
	std::snprintf(name, sizeof(name), "%016llx_m.co", (unsigned long long)h);
	mod_mom[slot] = load_unit(src, name, no_cache, "hipModuleLoadData(simulate_moments)");
	hip_check(hipModuleGetFunction(&fn_mom[slot], mod_mom[slot], "simulate_moments"),
			  "hipModuleGetFunction(simulate_moments)");
	compile_seconds += wall_seconds() - t0;
	return fn_mom[slot];
}

void program::load_util()
{
	if (mod_util)
		return;
	hip_check(hipModuleLoadData(&mod_util, mbg_co_util), "hipModuleLoadData(mbg_util)");
	hip_check(hipModuleGetFunction(&fn_fill_slots, mod_util, "mbg_fill_slots"), "hipModuleGetFunction(mbg_fill_slots)");
	hip_check(hipModuleGetFunction(&fn_expand_tails, mod_util, "mbg_expand_tails"),
			  "hipModuleGetFunction(mbg_expand_tails)");
	hip_check(hipModuleGetFunction(&fn_fill_u64, mod_util, "mbg_fill_u64"), "hipModuleGetFunction(mbg_fill_u64)");
	hip_check(hipModuleGetFunction(&fn_table_compact, mod_util, "mbg_table_compact"),
			  "hipModuleGetFunction(mbg_table_compact)");
	hip_check(hipModuleGetFunction(&fn_fp_compact, mod_util, "mbg_fp_compact"), "hipModuleGetFunction(mbg_fp_compact)");
	hip_check(hipModuleGetFunction(&fn_rate_view, mod_util, "mbg_rate_view"), "hipModuleGetFunction(mbg_rate_view)");
	hip_check(hipModuleGetFunction(&fn_cells_to_probs, mod_util, "mbg_cells_to_probs"),
			  "hipModuleGetFunction(mbg_cells_to_probs)");
	hip_check(hipModuleGetFunction(&fn_pack_cells, mod_util, "mbg_pack_cells"), "hipModuleGetFunction(mbg_pack_cells)");
}

hipFunction_t program::wave_function()
{
	if (!wtab)
		throw error(MBG_E_LIMIT, "the wave kernel cannot run this model: " + wave_unsupported);
	if (!mod_wave)
		hip_check(hipModuleLoadData(&mod_wave, mbg_co_wave), "hipModuleLoadData(simulate_wave)");
	int n = 1;
	while (n < wtab->npl)
		n <<= 1;
	hipFunction_t fn = nullptr;
	const std::string name = "simulate_wave_n" + std::to_string(n);
	hip_check(hipModuleGetFunction(&fn, mod_wave, name.c_str()), "hipModuleGetFunction(simulate_wave)");
	return fn;
}

hipFunction_t program::function(int flavour, int win_mode, int fin_mode)
{
	if (flavour < 0 || flavour > 2)
		throw error(MBG_E_INTERNAL, "unknown kernel flavour");
	const bool param = flavour == 1;
	if (param)
		win_mode = fin_mode = 0;
	if (win_mode < 0 || win_mode > MBG_HIST_SORT || fin_mode < 0 || fin_mode > MBG_HIST_LDS)
		throw error(MBG_E_INTERNAL, "unknown storage mode");
	hipFunction_t& fn = fns[flavour][win_mode][fin_mode];
	if (fn)
		return fn;
	hipModule_t& mod = mods[flavour][win_mode][fin_mode];

	scope_timer tm("compiler> whole_compilation");
	const double t0 = wall_seconds();
	// Occupancy (waves per SIMD; the compiler alone picks fewer waves with
	// more registers each): the most waves whose registers fit without
	// spilling more than 4 words, starting from 4 up to 40 nodes, 3 up to 160
	// and 2 above.  A spilled register is reloaded from scratch memory, and
	// that load waits for every atomic the wave still has in flight
	// (window-cell inserts), so a kernel that spills more goes down one wave.  Measured on
	// MI355X: at 30 nodes 4 waves beat the compiler's choice by 15 %; on the
	// 100-node model 2 waves without spills beat 3 with 11 spilled registers
	// by 5 %.
	const int n_nodes = mdl->n_nodes();
	int waves = 4;
	if (opts.lane_waves > 0)
		waves = opts.lane_waves;
	else if (n_nodes > 160)
		waves = 2;
	else if (n_nodes > 40)
		waves = 3;
	for (;;)
	{
		std::string src = "#define MBG_KW " + std::to_string(W) + "\n#define MBG_PARAM " + std::to_string(flavour)
						  + "\n#define MBG_LANE_WAVES " + std::to_string(waves) + "\n";
		if (!param)
			src += "#define MBG_WIN_MODE " + std::to_string(win_mode) + "\n#define MBG_FIN_MODE "
				   + std::to_string(fin_mode) + "\n";
		src += gen.source + "\n" + mbg_src_lane;
		const uint64_t h =
			unit_hash(src, { mbg_hdr_numeric, mbg_hdr_abi, mbg_hdr_kernel_args, mbg_hdr_stats_device, mbg_hdr_lane_core });
		char name[64];
		std::snprintf(name, sizeof(name), "%016llx_%d.co", (unsigned long long)h, flavour);
		mod = load_unit(src, name, no_cache, "hipModuleLoadData");
		if (param)
		{
			hip_check(hipModuleGetFunction(&fn, mod, "simulate_batched"), "hipModuleGetFunction(simulate_batched)");
			hip_check(hipModuleGetFunction(&fn_fill, mod, "mbg_fill_u64"), "hipModuleGetFunction(mbg_fill_u64)");
		}
		else
			hip_check(hipModuleGetFunction(&fn, mod, "simulate"), "hipModuleGetFunction(simulate)");
		int scratch = 0;
		hip_check(hipFuncGetAttribute(&scratch, HIP_FUNC_ATTRIBUTE_LOCAL_SIZE_BYTES, fn), "hipFuncGetAttribute");
		// a few spilled registers (up to 4 words) are kept: their reloads are rare
		if (scratch <= 16 || waves <= 2 || opts.lane_waves > 0)
			break;
		(void)hipModuleUnload(mod);
		mod = nullptr;
		fn = nullptr;
		--waves;
	}
	compile_seconds += wall_seconds() - t0;
	return fn;
}

std::string code_cache_dir() { return cache_dir(); }

std::unique_ptr<program> compile_program(std::shared_ptr<const model> m, const gen_options& opt, bool no_cache)
{
	auto p = std::make_unique<program>();
	p->W = m->n_words();
	try
	{
		p->wtab = std::make_unique<wave_tables>(build_wave_tables(*m));
	}
	catch (const error& e)
	{
		p->wave_unsupported = e.what();
	}
	if (p->W > max_lane_words && !p->wtab)
		throw error(MBG_E_LIMIT, "models with more than " + std::to_string(32 * max_lane_words)
									 + " nodes need the wave kernel, which cannot run this one: " + p->wave_unsupported);
	{
		scope_timer tm("generator> generate");
		p->gen = generate(*m, opt);
	}
	p->mdl = std::move(m);
	p->opts = opt;
	p->no_cache = no_cache;
	hip_check(hipGetDevice(&p->device), "hipGetDevice");
	hipDeviceProp_t prop;
	hip_check(hipGetDeviceProperties(&prop, p->device), "hipGetDeviceProperties");
	if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
		throw error(MBG_E_HIP, std::string("this build targets gfx950; current device is ") + prop.gcnArchName);
	p->num_cus = prop.multiProcessorCount;
	// the no-allocation entry point's scratch on this device (once per device)
	noalloc_setup(p->device);
	if (p->W <= auto_lane_words || !p->wtab)
	{
		int wm = 0, fm = 0;
		default_storage_modes(*p->mdl, wm, fm);
		p->function(0, wm, fm);
	}
	else
		p->wave_function();
	return p;
}

} // namespace mbg
