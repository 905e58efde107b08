// The force readout for callers that cannot allocate device memory (code under stream capture, serving loops,
// harnesses that forbid hipMalloc): aero_read_forces_noalloc runs the kernels of aero_read_forces with their block
// partials and results in a fixed file-scope device array instead of the context's lazily grown buffer.
//
// Each GPU the library runs on holds its own copy of `scratch` until the process ends. Its device address is
// looked up once per device when a context is created there (aero_noalloc_setup) and kept in g_scratch; host code
// only ever passes that address to kernels and copies, never the array's name (which in host code is a stand-in,
// not a device address). Every call on a GPU shares that GPU's array: calls that overlap on one GPU (two host
// threads, or contexts whose streams run concurrently) race, so callers serialize them.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <mutex>
#include <string>

#include "../../include/aerolbm.h"
#include "aero_ctx.h"
#include "lbm_launch.h"

using namespace aero;

namespace {

// Boundary cells the array serves: one block of partials (8 doubles) per 256 boundary cells, then 8 for the
// results. 2^20 boundary cells (256 KiB) cover the 1024x512x512 rocket (about 92 thousand) many times over.
constexpr uint32_t kMaxForceCells = 1u << 20;
constexpr uint32_t kMaxForceBlocks = kMaxForceCells / 256;
constexpr size_t kScratchDoubles = (size_t)8 * kMaxForceBlocks + 8;

__device__ __align__(256) double scratch[kScratchDoubles];

constexpr int kMaxDevices = 64;
double* g_scratch[kMaxDevices] = {};
std::mutex g_scratch_mu;

}  // namespace

// Looks up this device's copy of the array (the device must be current). Called by aero_create.
int aero_noalloc_setup(int dev, std::string* err) {
  if (dev < 0 || dev >= kMaxDevices) {
    *err = "aero_noalloc_setup: device ordinal beyond the supported range";
    return AERO_E_ARG;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch[dev]) return AERO_OK;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p) {
    *err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
    return AERO_E_HIP;
  }
  g_scratch[dev] = static_cast<double*>(p);
  return AERO_OK;
}

extern "C" int aero_read_forces_noalloc(aero_ctx* c, const aero_refs* refs, aero_forces* out) {
  if (!c || !out) return AERO_E_ARG;
  AERO_REQUIRE(c, c->init, AERO_E_STATE, "aero_read_forces_noalloc: call aero_reset first");
  AERO_REQUIRE(c, c->B <= kMaxForceCells, AERO_E_ARG,
               "aero_read_forces_noalloc: more than 2^20 boundary cells (use aero_read_forces)");
  double* ptr = nullptr;
  {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    ptr = c->dev >= 0 && c->dev < kMaxDevices ? g_scratch[c->dev] : nullptr;
  }
  AERO_REQUIRE(c, ptr != nullptr, AERO_E_STATE, "aero_read_forces_noalloc: no scratch address for this device");
  AERO_HIP(c, hipSetDevice(c->dev));
  double* part = ptr;                                 // 8 per block: every block writes all of its own
  double* res = ptr + (size_t)8 * kMaxForceBlocks;    // 7 sums, written whole by the final kernel
  const StepArgs a = c->args(false);
  AERO_HIP(c, launch_forces(a, c->storage == AERO_STORE_FP16, c->blist, c->aa, c->halo[c->hp], part, res, c->stream));
  double h[8];
  AERO_HIP(c, hipMemcpyAsync(h, res, 7 * 8, hipMemcpyDeviceToHost, c->stream));
  AERO_HIP(c, hipStreamSynchronize(c->stream));
  std::memset(out, 0, sizeof(*out));
  out->fx = h[0]; out->fy = h[1]; out->fz = h[2]; out->mx = h[3]; out->my = h[4]; out->mz = h[5];
  out->cut_face_links = (int64_t)std::llround(h[6]);
  if (refs) aero_coefficients(out, refs);
  return AERO_OK;
}
