// Kernel compiler: hipRTC compile of the generated source, link against the
// precompiled kernel bitcode, load as a module.  Replaces
// kernel_compiler::compile_simulation (src/kernel_compiler.cpp:61-164), which
// used NVRTC (-dlto -rdc) + nvJitLink against embedded LTO fatbins.
//
// The link is whole-program: hiprtcLink merges the two LLVM bitcode modules
// and optimises them together at -O3, so the generated rate code is inlined
// into the kernel loop.  Linked code objects are cached on disk, keyed by a
// hash of the generated source, the compile options and the kernel bitcode.
#include <hip/hiprtc.h>

#include <atomic>

#include <cstdio>
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
extern const char mbg_src_lane[];
extern const unsigned char mbg_co_wave[];
extern const unsigned char mbg_co_util[];
}

namespace mbg
{

// Kernel bitcode embedded at build time (see csrc/Makefile), one module per
// (state words, parametric) pair.
struct embedded_bc
{
	const unsigned char* begin;
	const unsigned char* end;
};
embedded_bc kernel_bitcode(int words, bool param);

void hip_check(hipError_t e, const char* what)
{
	if (e != hipSuccess)
		throw error(MBG_E_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

namespace
{

const char* const k_compile_flags[] = { "--offload-arch=gfx950",
										"-fgpu-rdc",
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

std::vector<char> compile_to_bitcode(const std::string& source)
{
	const char* headers[] = { mbg_hdr_numeric, mbg_hdr_abi };
	const char* names[] = { "mbg_numeric.h", "mbg_abi.h" };
	hiprtcProgram prog;
	rtc_check(hiprtcCreateProgram(&prog, source.c_str(), "mbg_model.hip", 2, headers, names), "hiprtcCreateProgram");
	hiprtcResult r = hiprtcCompileProgram(prog, (int)(sizeof(k_compile_flags) / sizeof(k_compile_flags[0])),
										  k_compile_flags);
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
	size_t bc_size = 0;
	rtc_check(hiprtcGetBitcodeSize(prog, &bc_size), "hiprtcGetBitcodeSize");
	std::vector<char> bc(bc_size);
	rtc_check(hiprtcGetBitcode(prog, bc.data()), "hiprtcGetBitcode");
	hiprtcDestroyProgram(&prog);
	return bc;
}

// The whole kernel in one translation unit: compile straight to a code
// object (gradient kernel; lane kernels when the bitcode link is refused).
std::vector<char> compile_to_code_object(const std::string& source)
{
	const char* headers[] = { mbg_hdr_numeric,		mbg_hdr_abi,	   mbg_hdr_kernel_args,
							  mbg_hdr_stats_device, mbg_hdr_lane_core, mbg_hdr_grad_kernel };
	const char* names[] = { "mbg_numeric.h",	  "mbg_abi.h",		"mbg_kernel_args.h",
							"mbg_stats_device.h", "mbg_lane_core.h", "mbg_grad_kernel.h" };
	std::vector<const char*> flags;
	for (const char* f : k_compile_flags)
		if (std::strcmp(f, "-fgpu-rdc") != 0)
			flags.push_back(f);
	hiprtcProgram prog;
	rtc_check(hiprtcCreateProgram(&prog, source.c_str(), "mbg_grad.hip", 6, headers, names), "hiprtcCreateProgram");
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

std::vector<char> link_code_object(std::vector<char>& model_bc, const embedded_bc& kbc)
{
	hiprtcLinkState ls;
	const char* isa_opts[] = { "-O3" };
	hiprtcJIT_option lopts[] = { HIPRTC_JIT_IR_TO_ISA_OPT_EXT, HIPRTC_JIT_IR_TO_ISA_OPT_COUNT_EXT };
	void* lvals[] = { (void*)isa_opts, (void*)(uintptr_t)1 };
	rtc_check(hiprtcLinkCreate(2, lopts, lvals, &ls), "hiprtcLinkCreate");
	hiprtcResult r = hiprtcLinkAddData(ls, HIPRTC_JIT_INPUT_LLVM_BITCODE, model_bc.data(), model_bc.size(),
									   "mbg_model", 0, nullptr, nullptr);
	if (r == HIPRTC_SUCCESS)
		r = hiprtcLinkAddData(ls, HIPRTC_JIT_INPUT_LLVM_BITCODE, (void*)kbc.begin, (size_t)(kbc.end - kbc.begin),
							  "mbg_kernels", 0, nullptr, nullptr);
	void* bin = nullptr;
	size_t bin_size = 0;
	if (r == HIPRTC_SUCCESS)
		r = hiprtcLinkComplete(ls, &bin, &bin_size);
	std::vector<char> out;
	if (r == HIPRTC_SUCCESS)
		out.assign((const char*)bin, (const char*)bin + bin_size);
	hiprtcLinkDestroy(ls);
	if (r != HIPRTC_SUCCESS)
		throw error(MBG_E_COMPILE, std::string("hipRTC link failed: ") + hiprtcGetErrorString(r));
	return out;
}

} // namespace

program::~program()
{
	if (mod_plain)
		(void)hipModuleUnload(mod_plain);
	if (mod_param)
		(void)hipModuleUnload(mod_param);
	if (mod_wave)
		(void)hipModuleUnload(mod_wave);
	if (batch_graph)
		(void)hipGraphExecDestroy(batch_graph);
	if (mod_util)
		(void)hipModuleUnload(mod_util);
	if (mod_grad)
		(void)hipModuleUnload(mod_grad);
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
	const std::string src = gen.source + gc.source + "\n#include \"mbg_grad_kernel.h\"\n";
	uint64_t h = fnv1a(src.data(), src.size());
	for (const char* f : k_compile_flags)
		h = fnv1a(f, std::strlen(f), h);
	for (const char* hd : { mbg_hdr_numeric, mbg_hdr_abi, mbg_hdr_kernel_args, mbg_hdr_stats_device,
							mbg_hdr_lane_core, mbg_hdr_grad_kernel })
		h = fnv1a(hd, std::strlen(hd), h);
	char name[64];
	std::snprintf(name, sizeof(name), "%016llx_g.co", (unsigned long long)h);
	const std::string dir = cache_dir();
	const std::string path = dir + "/" + name;
	std::vector<char> co;
	if (no_cache || !read_all(path, co))
	{
		co = compile_to_code_object(src);
		if (!no_cache)
		{
			mkdirs(dir);
			write_atomic(path, co.data(), co.size());
		}
	}
	hip_check(hipModuleLoadData(&mod_grad, co.data()), "hipModuleLoadData(simulate_grad)");
	hip_check(hipModuleGetFunction(&fn_grad, mod_grad, "simulate_grad"), "hipModuleGetFunction(simulate_grad)");
	compile_seconds += wall_seconds() - t0;
	return fn_grad;
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

hipFunction_t program::function(bool param)
{
	hipFunction_t& fn = param ? fn_param : fn_plain;
	if (fn)
		return fn;
	hipModule_t& mod = param ? mod_param : mod_plain;

	scope_timer tm("compiler> whole_compilation");
	const double t0 = wall_seconds();
	const embedded_bc kbc = kernel_bitcode(W, param);
	if (!kbc.begin || kbc.end <= kbc.begin)
		throw error(MBG_E_INTERNAL, "no precompiled kernel for " + std::to_string(W) + " state words");

	uint64_t h = fnv1a(gen.source.data(), gen.source.size());
	for (const char* f : k_compile_flags)
		h = fnv1a(f, std::strlen(f), h);
	h = fnv1a(kbc.begin, (size_t)(kbc.end - kbc.begin), h);
	h = fnv1a(mbg_hdr_numeric, std::strlen(mbg_hdr_numeric), h);
	h = fnv1a(mbg_hdr_abi, std::strlen(mbg_hdr_abi), h);
	char name[64];
	std::snprintf(name, sizeof(name), "%016llx_%d.co", (unsigned long long)h, param ? 1 : 0);
	const std::string dir = cache_dir();
	const std::string path = dir + "/" + name;

	std::vector<char> co;
	if (no_cache || !read_all(path, co))
	{
		static std::atomic<bool> link_refused { false };
		try
		{
			if (link_refused.load())
				throw error(MBG_E_COMPILE, "bitcode link refused earlier in this process");
			std::vector<char> bc = compile_to_bitcode(gen.source);
			co = link_code_object(bc, kbc);
		}
		catch (const error&)
		{
			link_refused.store(true);
			// The link reads the embedded bitcode with the process's code
			// object manager; when an older one is loaded first (for example
			// the copy bundled with a framework), it cannot read that bitcode.
			// Compiling the kernel source and the model in one translation
			// unit needs no bitcode from this build.
			co = compile_to_code_object("#define MBG_KW " + std::to_string(W) + "\n#define MBG_PARAM "
										+ (param ? "1" : "0") + "\n" + gen.source + "\n" + mbg_src_lane);
		}
		if (!no_cache)
		{
			mkdirs(dir);
			write_atomic(path, co.data(), co.size());
		}
	}
	hip_check(hipModuleLoadData(&mod, co.data()), "hipModuleLoadData");
	if (param)
	{
		hip_check(hipModuleGetFunction(&fn, mod, "simulate_batched"), "hipModuleGetFunction(simulate_batched)");
		hip_check(hipModuleGetFunction(&fn_fill, mod, "mbg_fill_u64"), "hipModuleGetFunction(mbg_fill_u64)");
	}
	else
		hip_check(hipModuleGetFunction(&fn, mod, "simulate"), "hipModuleGetFunction(simulate)");
	compile_seconds += wall_seconds() - t0;
	return fn;
}

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
	if (p->W <= max_lane_words)
		p->function(false);
	else
		p->wave_function();
	return p;
}

} // namespace mbg
