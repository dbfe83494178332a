// MaBoSS.GPU command line: `MaBoSS.GPU [-o prefix] bnd_file cfg_file`
// (src/main.cpp:101-169).  With -o the results go to <prefix>_probtraj.csv
// and <prefix>_fp.csv, otherwise the report is printed to stdout.  Built on
// the C-ABI only.  Extra options (not in the reference): --finals also writes
// <prefix>_finalstates.csv, --raw <path> writes the exact integer dump,
// --errors fills the ErrorTH / ErrorProba columns (MBG_SIM_ERRORS; 0 without,
// as the reference prints them).
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../../../include/maboss_gpu.h"

static int fail(const char* what)
{
	std::fprintf(stderr, "%s: %s\n", what, mbg_last_error());
	return 1;
}

int main(int argc, char** argv)
{
	std::vector<std::string> args;
	bool finals = false, errors = false;
	std::string raw;
	for (int i = 1; i < argc; i++)
	{
		std::string a = argv[i];
		if (a == "--finals")
			finals = true;
		else if (a == "--errors")
			errors = true;
		else if (a == "--raw" && i + 1 < argc)
			raw = argv[++i];
		else
			args.push_back(a);
	}
	if (!(args.size() == 2 || (args.size() == 4 && args[0] == "-o")))
	{
		std::printf("Usage: MaBoSS.GPU [-o prefix] bnd_file cfg_file\n");
		return 1;
	}
	std::string prefix, bnd = args[0], cfg = args[1];
	if (args[0] == "-o")
	{
		prefix = args[1];
		bnd = args[2];
		cfg = args[3];
	}

	mbg_model* m = nullptr;
	if (mbg_model_load(bnd.c_str(), cfg.c_str(), &m) != MBG_OK)
		return fail("parse");
	mbg_program* p = nullptr;
	if (mbg_program_compile(m, nullptr, &p) != MBG_OK)
	{
		mbg_model_free(m);
		return fail("compilation");
	}
	mbg_sim_params sp;
	mbg_sim_params_default(&sp);
	if (errors)
		sp.flags |= MBG_SIM_ERRORS;
	mbg_result* r = nullptr;
	int rc = mbg_simulate(p, &sp, nullptr, &r);
	if (rc != MBG_OK)
	{
		mbg_program_free(p);
		mbg_model_free(m);
		return fail("simulation");
	}
	if (!prefix.empty())
		rc = mbg_result_write_csv(r, prefix.c_str(), finals ? 1 : 0);
	else
		rc = mbg_result_print(r, nullptr);
	if (rc == MBG_OK && !raw.empty())
		rc = mbg_result_write_raw(r, raw.c_str());
	if (rc != MBG_OK)
		fail("output");
	mbg_result_free(r);
	mbg_program_free(p);
	mbg_model_free(m);
	return rc == MBG_OK ? 0 : 1;
}
