// Table of the kernel bitcode modules embedded by embed.S (generated at build
// time); one module per state-word count, plain and parametric flavours.
#include "engine.h"

extern "C" {
#define MBG_DECL(w)                                                                                                    \
	extern const unsigned char mbg_bc_w##w##_p0[], mbg_bc_w##w##_p0_end[], mbg_bc_w##w##_p1[], mbg_bc_w##w##_p1_end[];
MBG_DECL(1)
MBG_DECL(2)
MBG_DECL(3)
MBG_DECL(4)
MBG_DECL(5)
MBG_DECL(6)
MBG_DECL(7)
MBG_DECL(8)
#undef MBG_DECL
}

namespace mbg
{

struct embedded_bc
{
	const unsigned char* begin;
	const unsigned char* end;
};

embedded_bc kernel_bitcode(int words, bool param)
{
#define MBG_CASE(w)                                                                                                    \
	case w: return param ? embedded_bc { mbg_bc_w##w##_p1, mbg_bc_w##w##_p1_end } : embedded_bc { mbg_bc_w##w##_p0, mbg_bc_w##w##_p0_end };
	switch (words)
	{
		MBG_CASE(1)
		MBG_CASE(2)
		MBG_CASE(3)
		MBG_CASE(4)
		MBG_CASE(5)
		MBG_CASE(6)
		MBG_CASE(7)
		MBG_CASE(8)
		default: return embedded_bc { nullptr, nullptr };
	}
#undef MBG_CASE
}

} // namespace mbg
