// Compiles a HIP translation unit with hipRTC (no GPU needed) the way the
// engine's JIT does, with the given headers made available by name.
// Usage: rtc_compile source.hip header1.h [header2.h ...]; prints the log
// and exits non-zero on failure.
#include <hip/hiprtc.h>

#include <cstdio>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

static std::string slurp(const char* p)
{
	std::ifstream f(p);
	std::stringstream ss;
	ss << f.rdbuf();
	return ss.str();
}

int main(int argc, char** argv)
{
	if (argc < 2)
		return 2;
	std::string src = slurp(argv[1]);
	std::vector<std::string> texts, names;
	for (int i = 2; i < argc; i++)
	{
		texts.push_back(slurp(argv[i]));
		std::string n = argv[i];
		names.push_back(n.substr(n.find_last_of('/') + 1));
	}
	std::vector<const char*> tp, np;
	for (size_t i = 0; i < texts.size(); i++)
	{
		tp.push_back(texts[i].c_str());
		np.push_back(names[i].c_str());
	}
	const char* flags[] = { "--offload-arch=gfx950", "-O3", "-std=c++17", "-ffp-contract=off",
							"-fno-gpu-flush-denormals-to-zero", "-fhip-fp32-correctly-rounded-divide-sqrt",
							"-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops" };
	hiprtcProgram prog;
	if (hiprtcCreateProgram(&prog, src.c_str(), "tu.hip", (int)tp.size(), tp.data(), np.data()) != HIPRTC_SUCCESS)
		return 3;
	hiprtcResult r = hiprtcCompileProgram(prog, (int)(sizeof(flags) / sizeof(flags[0])), flags);
	size_t n = 0;
	hiprtcGetProgramLogSize(prog, &n);
	std::string log(n, '\0');
	if (n)
		hiprtcGetProgramLog(prog, &log[0]);
	std::fputs(log.c_str(), stderr);
	size_t code = 0;
	if (r == HIPRTC_SUCCESS)
		hiprtcGetCodeSize(prog, &code);
	std::printf("%s %zu\n", r == HIPRTC_SUCCESS ? "OK" : hiprtcGetErrorString(r), code);
	hiprtcDestroyProgram(&prog);
	return r == HIPRTC_SUCCESS ? 0 : 1;
}
