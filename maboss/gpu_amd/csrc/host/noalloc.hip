// Static device scratch of the no-allocation gradient entry point
// (mbg_backward_noalloc).  No counterpart in the reference.
//
// The array below is the whole device workspace of that entry point (the
// gradient kernel's counters and tables); it lives in this library's code
// object, so every GPU the library runs on holds one copy of it from load
// until the process ends.  In host code `scratch` is only a host-side handle
// of the symbol, not a device address: the address on a device is resolved
// once per device, with that device current, by noalloc_setup() (called when
// a program is compiled there), and only that address is handed to kernels
// and copies.  Every call on one GPU shares the array, so calls that overlap
// on a GPU race; each call clears what must start at zero.
#include <hip/hip_runtime.h>

#include <mutex>

#include "engine.h"

static __device__ __align__(256) unsigned char scratch[mbg::k_noalloc_scratch_bytes];

namespace mbg
{

namespace
{
constexpr int k_max_devices = 64;
std::mutex g_mu;
void* g_scratch[k_max_devices] = {};
} // namespace

void noalloc_setup(int device)
{
	if (device < 0 || device >= k_max_devices)
		throw error(MBG_E_LIMIT, "device index out of range for the no-allocation scratch");
	std::lock_guard<std::mutex> lk(g_mu);
	if (g_scratch[device])
		return;
	int cur = -1;
	hip_check(hipGetDevice(&cur), "hipGetDevice");
	if (cur != device)
		throw error(MBG_E_INTERNAL, "no-allocation scratch set up with another device current");
	void* ptr = nullptr;
	hip_check(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
	if (!ptr)
		throw error(MBG_E_HIP, "hipGetSymbolAddress(scratch) returned no address");
	g_scratch[device] = ptr;
}

void* noalloc_scratch(int device)
{
	if (device < 0 || device >= k_max_devices)
		return nullptr;
	std::lock_guard<std::mutex> lk(g_mu);
	return g_scratch[device];
}

} // namespace mbg
