// Solver context behind the C ABI: device buffers, host mirrors of the geometry and the step state.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/aerolbm.h"
#include "lbm_kernels.h"

// Error reporting of the entry points on a context: the message goes to ctx->err, the code is returned.
#define AERO_HIP(ctx, call)                                                                      \
  do {                                                                                          \
    hipError_t e_ = (call);                                                                     \
    if (e_ != hipSuccess) {                                                                     \
      (ctx)->err = std::string(#call) + ": " + hipGetErrorString(e_);                          \
      return AERO_E_HIP;                                                                        \
    }                                                                                           \
  } while (0)

#define AERO_REQUIRE(ctx, cond, code, msg) \
  do {                                     \
    if (!(cond)) {                         \
      (ctx)->err = (msg);                  \
      return (code);                       \
    }                                      \
  } while (0)

// Device buffers of the GPU voxelizer (BVH, scanline and compaction scratch); freed on destroy.
struct VoxScratch {
  void* buf[12] = {};
  size_t cap[12] = {};
  void* get(int i, size_t bytes) {
    if (cap[i] < bytes) {
      if (buf[i]) hipFree(buf[i]);
      buf[i] = nullptr;
      cap[i] = 0;
      if (hipMalloc(&buf[i], bytes) != hipSuccess) return nullptr;
      cap[i] = bytes;
    }
    return buf[i];
  }
  void release() {
    for (int i = 0; i < 12; ++i) {
      if (buf[i]) hipFree(buf[i]);
      buf[i] = nullptr;
      cap[i] = 0;
    }
  }
};

// A captured pair of plain steps, replayed for runs of plain steps. The pair starting in A-A state E and the one
// starting in state O differ in their kernels and halo buffers, so each context keeps one of each. `key` is the
// first step's kernel arguments and `width` the bulk group width at capture: the graph holds both by value, so
// it is replayed only while a fresh capture would record the same.
struct StepPairGraph {
  hipGraphExec_t exec = nullptr;
  aero::StepArgs key;
  aero::StepArgs key2;  // the second step's arguments (graphs that end with a readout step)
  int width = 0;
};

struct aero_ctx {
  int dev = 0;
  hipStream_t stream = nullptr;
  uint32_t nx = 0, ny = 0, nz = 0, N = 0;
  int storage = AERO_STORE_FP32;
  aero::Geo geo;
  aero::Phys ph;
  aero_params p;

  // device memory
  void* f = nullptr;               // 19 x N populations (float or half)
  uint8_t* flags = nullptr;        // N
  uint32_t* rowfluid = nullptr;    // ny nz: 1 where the whole row is plain fluid
  int32_t* bmap = nullptr;         // N
  int32_t* blist = nullptr;        // B
  float* bq = nullptr;             // B x 18
  float* minq = nullptr;           // B
  uint32_t* bmask = nullptr;       // B
  float* bown = nullptr;           // 19 x B
  uint32_t B = 0;
  float* halo[2] = {nullptr, nullptr};
  size_t halo_n = 0;
  double* partials = nullptr;
  uint32_t nblk = 0;
  int vw = 1;        // cells per thread of the bulk step kernel (1, 2 or 4; 0: no bulk kernel)
  int32_t* spec = nullptr;         // cells outside bulk groups of spec_v that the split step visits:
  uint32_t nspec = 0;              // nspec in all, the first nface of them plain fluid cells
  uint32_t nface = 0;
  hipStream_t side = nullptr;      // the other listed cells run beside the bulk kernel
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  uint32_t nbulk = 0;              // cells the bulk kernel updates
  int spec_v = 0;                  // group width the list was built for (0: none)
  const aero_mlp* mlp_src = nullptr;  // learned closure (null: Smagorinsky)
  const void* mlp = nullptr;          // its device weights
  aero::StepState* st = nullptr;
  float* mac = nullptr;            // 4 x N (readout)
  float* snap = nullptr;           // 19 x N (readout / staging)
  double* fpart = nullptr;
  uint32_t fpart_cap = 0;
  double* fout = nullptr;
  VoxScratch vox;
  // derived-field readouts (allocated on first use)
  uint32_t* slice_buf = nullptr;   // RGBA8 words of the largest slice plane
  float* qbuf = nullptr;           // 5 x N: Q, |omega|, omega (3 per cell)
  double* qpart = nullptr;         // 2 per Q-criterion block, then the 2 maxima
  double* probe_buf = nullptr;     // per probe: 3 in, 13 out, and the surface flags
  size_t probe_cap = 0;
  double* mass_buf = nullptr;      // mass statistics: 4 per block, then the 5 results
  double* uprof = nullptr;         // ny inflow profile factors (StepArgs::uprof)
  std::vector<double> uprof_h;     // their host copy, the source of the upload

  // host mirrors and state
  std::vector<uint8_t> h_flags;
  int32_t counts[3] = {0, 0, 0};   // solid, boundary, fluid
  bool geom = false, init = false;
  bool readout_valid = false;       // macro fields of the last step are in mac
  bool pops_valid = false;          // and its populations in snap
  bool mac_solids = false;          // mac holds (0, 0, 0, 1) for every solid cell of the current geometry
  int aa = 0;        // A-A state of the buffer: 0 = E (even steps done), 1 = O
  int hp = 0;        // index of the most recently written halo
  long long steps = 0;
  double vox_ms = 0.0;
  std::string err;
  StepPairGraph pair_graph[2];     // indexed by the A-A state a pair starts from
  StepPairGraph ro_graph[2][2];    // a plain step then a readout step: [A-A state][with the population snapshot]

  // Kernel arguments of a step; `macros` / `populations`: it writes the macro fields / the snapshot.
  aero::StepArgs args(bool macros, bool populations = false) const;
  int upload_geometry_common();
  int prepare_split();
  // group width of the split step: the learned closure's bulk kernel takes one cell per thread; 0 selects the
  // one-cell-per-thread kernel for every cell
  int split_width() const { return vw == 0 ? 0 : (mlp ? 1 : vw); }
  // Readout buffers: the macro fields, and the population snapshot when `populations` is set. A split readout step
  // visits only cells with work, so the macro fields of solid cells are written once per geometry.
  int ensure_readout(bool populations);
  void invalidate_readout() { readout_valid = pops_valid = false; }
  // Waits for the context's streams: host-synchronous copies into its buffers (null stream) must not overtake
  // stream-ordered work that is still reading or writing them.
  int quiesce();
  void drop_graphs();
};

void aero_coefficients(aero_forces* o, const aero_refs* r);
// Per-device set-up of aero_read_forces_noalloc's fixed scratch array (with the device current).
int aero_noalloc_setup(int dev, std::string* err);
const void* aero_mlp_device_weights(const aero_mlp* m);
int aero_mlp_device(const aero_mlp* m);
