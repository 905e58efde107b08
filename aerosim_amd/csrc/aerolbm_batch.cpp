// Many independent tunnels stepped together, captured in a HIP graph. Each tunnel is an ordinary aero_ctx
// (geometry, parameters, reset and readouts go through the single-tunnel entry points); the batch only
// replaces the stepping. A batched step takes one of two forms, chosen by the launch shape:
// - grid 0: per-kind launches, one block per tile: bulk tiles (runs of consecutive cells; bulk groups take the
//   vector path), list tiles (each tunnel's listed cells: face rows and cells next to the body; the cells next to
//   the body on a second stream), then a finalize launch;
// - grid G > 0: one persistent launch of G blocks over every tile of every tunnel, split by XCD, drawn from
//   per-XCD work counters and finalized by its last block (k_step_batch_one).
// The shape (threads per block, cells per thread, grid) is chosen by aero_batch_tune from timed candidates.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/aerolbm.h"
#include "aero_ctx.h"
#include "lbm_launch.h"

using namespace aero;

struct aero_batch {
  std::vector<aero_ctx*> t;
  int dev = 0;
  hipStream_t stream = nullptr;     // the batch's own stream, ordered against every tunnel's stream by events
  hipEvent_t ev_in = nullptr, ev_out = nullptr;
  BatchShape shape{256, 1, 0};      // until aero_batch_set_shape / aero_batch_tune: per-kind launches, 1 cell per
                                    // thread, 512 threads for FP32 storage and 256 for FP16 (set at creation),
                                    // the shapes measured fastest on the 64-tunnel sweep
  bool fp16 = false;
  // device tables
  StepArgs* targs[2] = {nullptr, nullptr};  // per tunnel, the arguments of a step starting from A-A state 0 / 1
  BatchTile* tiles = nullptr;       // bulk tiles, then plain-fluid list tiles, then other list tiles
  uint32_t* tile_begin = nullptr;   // per tunnel: first slot of its partials (and one past the last)
  StepState** states = nullptr;
  double* part = nullptr;
  uint32_t nbulk = 0, nfluid = 0, nother = 0, cap_tiles = 0;
  hipStream_t side = nullptr;       // other list tiles beside the bulk kernel
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  // What the device tables were built from (the shape and every tunnel's state-0 arguments) and, per A-A state a
  // pair starts from, its captured graph, valid while the tables are.
  bool tab_valid = false;
  BatchShape tab_shape{0, 0, 0};
  std::vector<StepArgs> tab_key;
  hipGraphExec_t pair_exec[2] = {nullptr, nullptr};
  // one-launch form (shape grid > 0)
  BatchTileP* ptiles = nullptr;
  uint32_t cap_ptiles = 0;
  uint32_t* xcd_begin = nullptr;    // 9 entries
  uint32_t* sync = nullptr;         // ticket, XCD work counters, density totals (batch_sync_bytes)
  std::string err;
  int num_cu = 256;
};

#define B_HIP(b, call)                                                                     \
  do {                                                                                    \
    hipError_t e_ = (call);                                                               \
    if (e_ != hipSuccess) {                                                               \
      (b)->err = std::string(#call) + ": " + hipGetErrorString(e_);                       \
      return AERO_E_HIP;                                                                  \
    }                                                                                     \
  } while (0)

namespace {
std::string g_batch_create_error;  // why the last aero_batch_create failed (read with a null handle)

int check_lockstep(aero_batch* b) {
  const aero_ctx* c0 = b->t[0];
  for (aero_ctx* c : b->t) {
    if (!c->init) { b->err = "aero_batch: every tunnel needs aero_reset before batched stepping"; return AERO_E_STATE; }
    if (c->mlp) { b->err = "aero_batch: batched stepping runs the Smagorinsky closure only"; return AERO_E_STATE; }
    if (c->aa != c0->aa || c->hp != c0->hp) {
      b->err = "aero_batch: tunnels are not in step (reset them together)";
      return AERO_E_STATE;
    }
  }
  return AERO_OK;
}

void drop_pair_graphs(aero_batch* b) {
  for (hipGraphExec_t& g : b->pair_exec) {
    if (g) {
      (void)hipStreamSynchronize(b->stream);  // no replay may still be in flight
      (void)hipGraphExecDestroy(g);
    }
    g = nullptr;
  }
}

// The arguments of tunnel c's next step if it started from A-A state `state` (the state and the halo parity flip
// together on every step).
StepArgs args_from(aero_ctx* c, int state) {
  const int aa0 = c->aa, hp0 = c->hp;
  c->aa = c->hp = state;
  const StepArgs a = c->args(false);
  c->aa = aa0;
  c->hp = hp0;
  return a;
}

int build_one_tables(aero_batch* b, const std::vector<StepArgs>& key);
int build_kind_tables(aero_batch* b, const std::vector<StepArgs>& key);

// Brings the tile lists and the StepArgs tables up to date with the current shape and tunnel state (each tunnel's
// listed cells are built for the shape's cells per thread). They are rebuilt, and the captured graphs dropped, only
// when the shape or some tunnel's step arguments changed.

int build_tables(aero_batch* b) {
  for (aero_ctx* c : b->t)
    if (c->spec_v != b->shape.cpt) {
      B_HIP(b, hipStreamSynchronize(b->stream));  // earlier batched steps may still read the lists rebuilt below
      break;
    }
  for (aero_ctx* c : b->t) {
    c->vw = b->shape.cpt;
    const int rc = c->prepare_split();
    if (rc != AERO_OK) { b->err = c->err; return rc; }
  }
  std::vector<StepArgs> key;
  key.reserve(b->t.size());
  for (aero_ctx* c : b->t) key.push_back(args_from(c, 0));
  if (b->tab_valid && b->tab_shape.block == b->shape.block && b->tab_shape.cpt == b->shape.cpt &&
      b->tab_shape.grid == b->shape.grid &&
      std::memcmp(b->tab_key.data(), key.data(), key.size() * sizeof(StepArgs)) == 0)
    return AERO_OK;
  drop_pair_graphs(b);
  b->tab_valid = false;
  const int rc = b->shape.grid ? build_one_tables(b, key) : build_kind_tables(b, key);
  if (rc) return rc;
  b->tab_key.swap(key);
  b->tab_shape = b->shape;
  b->tab_valid = true;
  return AERO_OK;
}

// StepArgs of both A-A states and the state pointers of every tunnel (key: the state-0 arguments).
int upload_args(aero_batch* b, const std::vector<StepArgs>& key) {
  std::vector<StepArgs> a1;
  std::vector<StepState*> st;
  for (aero_ctx* c : b->t) {
    a1.push_back(args_from(c, 1));
    st.push_back(c->st);
  }
  B_HIP(b, hipMemcpyAsync(b->targs[0], key.data(), key.size() * sizeof(StepArgs), hipMemcpyHostToDevice, b->stream));
  B_HIP(b, hipMemcpyAsync(b->targs[1], a1.data(), a1.size() * sizeof(StepArgs), hipMemcpyHostToDevice, b->stream));
  B_HIP(b, hipMemcpyAsync(b->states, st.data(), st.size() * sizeof(StepState*), hipMemcpyHostToDevice, b->stream));
  B_HIP(b, hipStreamSynchronize(b->stream));
  return AERO_OK;
}

// Tables of the per-kind launches (shape grid 0).
int build_kind_tables(aero_batch* b, const std::vector<StepArgs>& key) {
  const uint32_t cells_per_tile = (uint32_t)(b->shape.block * b->shape.cpt);
  const uint32_t list_tile = 256;
  std::vector<BatchTile> bulk, fluid, other;
  std::vector<uint32_t> begin;
  uint32_t slot = 0;
  for (size_t i = 0; i < b->t.size(); ++i) {
    aero_ctx* c = b->t[i];
    begin.push_back(slot);
    for (uint32_t s = 0; s < c->N; s += cells_per_tile)
      bulk.push_back(BatchTile{(uint32_t)i, s, std::min<uint32_t>(cells_per_tile, c->N - s), slot++});
    // plain fluid entries and the others never share a tile
    for (uint32_t s = 0; s < c->nface; s += list_tile)
      fluid.push_back(BatchTile{(uint32_t)i, s, std::min<uint32_t>(list_tile, c->nface - s), slot++});
    for (uint32_t s = c->nface; s < c->nspec; s += list_tile)
      other.push_back(BatchTile{(uint32_t)i, s, std::min<uint32_t>(list_tile, c->nspec - s), slot++});
  }
  begin.push_back(slot);
  b->nbulk = (uint32_t)bulk.size();
  b->nfluid = (uint32_t)fluid.size();
  b->nother = (uint32_t)other.size();
  const uint32_t ntiles = b->nbulk + b->nfluid + b->nother;
  B_HIP(b, hipStreamSynchronize(b->stream));  // earlier steps may still read the tables replaced below
  if (ntiles > b->cap_tiles) {
    if (b->tiles) hipFree(b->tiles);
    if (b->part) hipFree(b->part);
    b->tiles = nullptr;
    b->part = nullptr;
    b->cap_tiles = 0;
    B_HIP(b, hipMalloc(&b->tiles, (size_t)ntiles * sizeof(BatchTile)));
    B_HIP(b, hipMalloc(&b->part, (size_t)ntiles * 8));
    b->cap_tiles = ntiles;
  }
  bulk.insert(bulk.end(), fluid.begin(), fluid.end());
  bulk.insert(bulk.end(), other.begin(), other.end());
  B_HIP(b, hipMemcpyAsync(b->tiles, bulk.data(), bulk.size() * sizeof(BatchTile), hipMemcpyHostToDevice, b->stream));
  B_HIP(b, hipMemcpyAsync(b->tile_begin, begin.data(), begin.size() * 4, hipMemcpyHostToDevice, b->stream));
  return upload_args(b, key);
}

// Tables of the one-launch form (shape grid G > 0). Every tunnel contributes bulk tiles of block x cpt cells and
// list tiles of block entries. The tiles, in tunnel order, are cut into 8 contiguous XCD ranges of about equal
// estimated cost; within a range the tiles of cells next to a body come first, so that the heaviest tiles start
// early. The blocks of XCD x then draw the tiles of its range from that XCD's counter.
int build_one_tables(aero_batch* b, const std::vector<StepArgs>& key) {
  const uint32_t bsz = (uint32_t)b->shape.block, G = b->shape.grid;
  if (G == 0 || (G & 7u) != 0) {
    b->err = "aero_batch: the one-launch form needs a grid that is a positive multiple of 8";
    return AERO_E_ARG;
  }
  const uint32_t cells_per_tile = bsz * (uint32_t)b->shape.cpt;
  struct HT {
    uint32_t kind, tun, start, count;
  };
  std::vector<HT> all;
  std::vector<double> cost;
  for (size_t i = 0; i < b->t.size(); ++i) {
    aero_ctx* c = b->t[i];
    const double bulk_share = c->N ? (double)c->nbulk / (double)c->N : 1.0;
    for (uint32_t s = 0; s < c->N; s += cells_per_tile) {
      all.push_back(HT{TILE_BULK, (uint32_t)i, s, std::min<uint32_t>(cells_per_tile, c->N - s)});
      cost.push_back(all.back().count * (0.1 + 0.9 * bulk_share));
    }
    for (uint32_t s = 0; s < c->nface; s += bsz) {
      all.push_back(HT{TILE_FLUID, (uint32_t)i, s, std::min<uint32_t>(bsz, c->nface - s)});
      cost.push_back(1.5 * all.back().count);
    }
    for (uint32_t s = c->nface; s < c->nspec; s += bsz) {
      all.push_back(HT{TILE_OTHER, (uint32_t)i, s, std::min<uint32_t>(bsz, c->nspec - s)});
      cost.push_back(3.0 * all.back().count);
    }
  }
  const size_t nt = all.size();
  if (nt >= (1u << 31)) {
    b->err = "aero_batch: too many tiles for the one-launch form";
    return AERO_E_ARG;
  }
  // XCD ranges of about equal cost
  double total = 0.0;
  for (double w : cost) total += w;
  std::vector<uint32_t> xb(9, (uint32_t)nt);
  xb[0] = 0;
  {
    double acc = 0.0;
    int x = 1;
    for (size_t k = 0; k < nt && x < 8; ++k) {
      acc += cost[k];
      while (x < 8 && acc >= total * x / 8.0) xb[x++] = (uint32_t)(k + 1);
    }
  }
  // within each range: tiles of cells next to a body first, the others in tunnel order
  std::vector<BatchTileP> pt;
  pt.reserve(nt);
  for (int x = 0; x < 8; ++x) {
    for (int pass = 0; pass < 2; ++pass)
      for (uint32_t k = xb[x]; k < xb[x + 1]; ++k)
        if ((all[k].kind == TILE_OTHER) == (pass == 0))
          pt.push_back(BatchTileP{all[k].kind << 30 | all[k].tun, all[k].start, all[k].count, 0u});
  }
  B_HIP(b, hipStreamSynchronize(b->stream));  // earlier steps may still read the tables replaced below
  if (nt > b->cap_ptiles) {
    if (b->ptiles) hipFree(b->ptiles);
    b->ptiles = nullptr;
    b->cap_ptiles = 0;
    B_HIP(b, hipMalloc(&b->ptiles, nt * sizeof(BatchTileP)));
    b->cap_ptiles = (uint32_t)nt;
  }
  B_HIP(b, hipMemcpyAsync(b->ptiles, pt.data(), nt * sizeof(BatchTileP), hipMemcpyHostToDevice, b->stream));
  B_HIP(b, hipMemcpyAsync(b->xcd_begin, xb.data(), 9 * 4, hipMemcpyHostToDevice, b->stream));
  B_HIP(b, hipMemsetAsync(b->sync, 0, batch_sync_bytes((uint32_t)b->t.size()), b->stream));
  return upload_args(b, key);
}

// Enqueues `count` steps (1 or 2) starting from A-A state `state`.
int enqueue(aero_batch* b, int state, int count) {
  for (int s = 0; s < count; ++s) {
    const int st = state ^ (s & 1);
    if (b->shape.grid) {
      const BatchOneArgs p{b->targs[st], b->ptiles, b->xcd_begin, b->states, (uint32_t)b->t.size(), b->sync};
      B_HIP(b, launch_step_batch_one(p, b->fp16, st == 0, b->shape, b->stream));
      continue;
    }
    const SideLaunch side{b->side, b->ev_fork, b->ev_join};
    B_HIP(b, launch_step_batch(b->targs[st], b->tiles, b->nbulk, b->tiles + b->nbulk, b->nfluid,
                               b->tiles + b->nbulk + b->nfluid, b->nother, b->part, b->fp16, st == 0, b->shape,
                               b->stream, &side));
    B_HIP(b, launch_finalize_batch(b->part, b->tile_begin, b->states, (uint32_t)b->t.size(), b->stream));
  }
  return AERO_OK;
}

// The captured pair of steps starting from A-A state `state` (captured on first use after the tables changed).
int pair_graph(aero_batch* b, int state, hipGraphExec_t* out) {
  if (!b->pair_exec[state]) {
    B_HIP(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    const int rc = enqueue(b, state, 2);
    hipGraph_t cap = nullptr;
    const hipError_t ee = hipStreamEndCapture(b->stream, &cap);
    if (rc) {
      if (cap) hipGraphDestroy(cap);
      return rc;
    }
    if (ee != hipSuccess) { b->err = std::string("graph capture: ") + hipGetErrorString(ee); return AERO_E_HIP; }
    const hipError_t ei = hipGraphInstantiate(&b->pair_exec[state], cap, nullptr, nullptr, 0);
    hipGraphDestroy(cap);
    if (ei != hipSuccess) {
      b->pair_exec[state] = nullptr;
      b->err = std::string("graph instantiate: ") + hipGetErrorString(ei);
      return AERO_E_HIP;
    }
  }
  *out = b->pair_exec[state];
  return AERO_OK;
}

// Blocks of a one-launch grid: mult per CU, rounded up to whole rounds of the 8 XCDs.
uint32_t one_launch_grid(int num_cu, int mult) { return ((uint32_t)(num_cu * mult) + 7u) & ~7u; }

void advance(aero_batch* b, int n) {
  for (aero_ctx* c : b->t) {
    if (n & 1) {
      c->aa ^= 1;
      c->hp ^= 1;
    }
    c->steps += n;
    c->invalidate_readout();
  }
}

// Runs n steps: pairs through a captured graph, a trailing odd step directly. The batch's stream first waits for
// the work queued on every tunnel's own stream, and every tunnel's stream then waits for the batch, so calls on
// single tunnels before and after stay ordered with the batched steps.
int run(aero_batch* b, int n, hipEvent_t e0, hipEvent_t e1) {
  int rc = build_tables(b);
  if (rc) return rc;
  const int state = b->t[0]->aa;
  hipGraphExec_t ge = nullptr;
  if (n >= 2) {
    rc = pair_graph(b, state, &ge);
    if (rc) return rc;
  }
  for (aero_ctx* c : b->t) {
    B_HIP(b, hipEventRecord(b->ev_in, c->stream));
    B_HIP(b, hipStreamWaitEvent(b->stream, b->ev_in, 0));
  }
  if (e0) B_HIP(b, hipEventRecord(e0, b->stream));
  for (int p = 0; p < n / 2; ++p) B_HIP(b, hipGraphLaunch(ge, b->stream));
  if (n & 1) {
    rc = enqueue(b, state, 1);
    if (rc) return rc;
  }
  if (e1) B_HIP(b, hipEventRecord(e1, b->stream));
  B_HIP(b, hipEventRecord(b->ev_out, b->stream));
  for (aero_ctx* c : b->t) B_HIP(b, hipStreamWaitEvent(c->stream, b->ev_out, 0));
  advance(b, n);
  return AERO_OK;
}

}  // namespace

extern "C" {

int aero_batch_create(aero_ctx* const* tunnels, int32_t n, aero_batch** out) {
  auto fail = [](const std::string& m, int rc) {
    g_batch_create_error = "aero_batch_create: " + m;
    return rc;
  };
  if (!tunnels || n <= 0 || !out) return fail("null argument or no tunnels", AERO_E_ARG);
  *out = nullptr;
  for (int i = 0; i < n; ++i) {
    const std::string which = "tunnel " + std::to_string(i);
    if (!tunnels[i]) return fail(which + " is null", AERO_E_ARG);
    if (tunnels[i]->dev != tunnels[0]->dev) return fail(which + " is on another device than tunnel 0", AERO_E_ARG);
    if (tunnels[i]->storage != tunnels[0]->storage) return fail(which + " has another storage than tunnel 0", AERO_E_ARG);
    if (!tunnels[i]->geo.bufok)  // a cross-section beyond the bulk kernels' 32-bit buffer offsets
      return fail(which + "'s cross-section is too large for the batched bulk kernel", AERO_E_ARG);
  }
  aero_batch* b = new aero_batch();
  for (int i = 0; i < n; ++i) b->t.push_back(tunnels[i]);
  b->dev = tunnels[0]->dev;
  b->fp16 = tunnels[0]->storage == AERO_STORE_FP16;
  b->shape.block = b->fp16 ? 256 : 512;
  hipSetDevice(b->dev);
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, b->dev) == hipSuccess && prop.multiProcessorCount > 0) b->num_cu = prop.multiProcessorCount;
  bool ok = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) == hipSuccess &&
            hipStreamCreateWithFlags(&b->side, hipStreamNonBlocking) == hipSuccess &&
            hipEventCreateWithFlags(&b->ev_in, hipEventDisableTiming) == hipSuccess &&
            hipEventCreateWithFlags(&b->ev_out, hipEventDisableTiming) == hipSuccess &&
            hipEventCreateWithFlags(&b->ev_fork, hipEventDisableTiming) == hipSuccess &&
            hipEventCreateWithFlags(&b->ev_join, hipEventDisableTiming) == hipSuccess &&
            hipMalloc(&b->targs[0], (size_t)n * sizeof(StepArgs)) == hipSuccess &&
            hipMalloc(&b->targs[1], (size_t)n * sizeof(StepArgs)) == hipSuccess &&
            hipMalloc(&b->tile_begin, (size_t)(n + 1) * 4) == hipSuccess &&
            hipMalloc(&b->xcd_begin, 9 * 4) == hipSuccess &&
            hipMalloc(&b->sync, batch_sync_bytes((uint32_t)n)) == hipSuccess &&
            hipMalloc(&b->states, (size_t)n * sizeof(StepState*)) == hipSuccess;
  if (!ok) {
    aero_batch_destroy(b);
    return fail("stream, event or device allocation failed", AERO_E_NOMEM);
  }
  *out = b;
  return AERO_OK;
}

int aero_batch_destroy(aero_batch* b) {
  if (!b) return AERO_OK;
  hipSetDevice(b->dev);
  if (b->stream) hipStreamSynchronize(b->stream);
  if (b->side) hipStreamSynchronize(b->side);
  drop_pair_graphs(b);
  for (int i = 0; i < 2; ++i)
    if (b->targs[i]) hipFree(b->targs[i]);
  if (b->tiles) hipFree(b->tiles);
  if (b->tile_begin) hipFree(b->tile_begin);
  if (b->states) hipFree(b->states);
  if (b->part) hipFree(b->part);
  if (b->ptiles) hipFree(b->ptiles);
  if (b->xcd_begin) hipFree(b->xcd_begin);
  if (b->sync) hipFree(b->sync);
  if (b->ev_fork) hipEventDestroy(b->ev_fork);
  if (b->ev_join) hipEventDestroy(b->ev_join);
  if (b->ev_in) hipEventDestroy(b->ev_in);
  if (b->ev_out) hipEventDestroy(b->ev_out);
  if (b->side) hipStreamDestroy(b->side);
  if (b->stream) hipStreamDestroy(b->stream);
  delete b;
  return AERO_OK;
}

const char* aero_batch_last_error(const aero_batch* b) { return b ? b->err.c_str() : g_batch_create_error.c_str(); }

int aero_batch_set_shape(aero_batch* b, int32_t block, int32_t cpt, int32_t grid) {
  if (!b) return AERO_E_ARG;
  bool ok = (block == 128 || block == 256 || block == 512) && (cpt == 1 || cpt == 2 || cpt == 4) && grid >= 0 &&
            grid % 8 == 0;
  for (aero_ctx* c : b->t) ok = ok && c->nx % (uint32_t)cpt == 0;
  if (!ok) {
    b->err = "aero_batch_set_shape: block in {128,256,512}, cells per thread in {1,2,4} dividing every nx, grid 0 or a "
             "positive multiple of 8";
    return AERO_E_ARG;
  }
  b->shape = BatchShape{block, cpt, (uint32_t)grid};
  return AERO_OK;
}

int aero_batch_get_shape(aero_batch* b, int32_t* out) {
  if (!b || !out) return AERO_E_ARG;
  out[0] = b->shape.block;
  out[1] = b->shape.cpt;
  out[2] = (int32_t)b->shape.grid;
  return AERO_OK;
}

int aero_batch_step(aero_batch* b, int32_t n) {
  if (!b || n < 0) return AERO_E_ARG;
  B_HIP(b, hipSetDevice(b->dev));
  int rc = check_lockstep(b);
  if (rc) return rc;
  if (n == 0) return AERO_OK;
  return run(b, n, nullptr, nullptr);
}

int aero_batch_time_steps(aero_batch* b, int32_t n, double* out) {
  if (!b || n <= 0 || !out) return AERO_E_ARG;
  B_HIP(b, hipSetDevice(b->dev));
  int rc = check_lockstep(b);
  if (rc) return rc;
  hipEvent_t e0, e1;
  B_HIP(b, hipEventCreate(&e0));
  B_HIP(b, hipEventCreate(&e1));
  rc = run(b, n, e0, e1);
  if (rc) return rc;
  B_HIP(b, hipEventSynchronize(e1));
  float ms = 0.f;
  B_HIP(b, hipEventElapsedTime(&ms, e0, e1));
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  out[0] = ms;
  out[1] = (double)ms / n;
  return AERO_OK;
}

// Times every candidate launch shape for `steps` steps (after one warm-up pair) and keeps the fastest.
// The flow state advances while tuning: reset the tunnels afterwards. out: block, cells per thread, grid
// of the winner; ms_out: its milliseconds per step.
int aero_batch_tune(aero_batch* b, int32_t steps, int32_t* out, double* ms_out) {
  if (!b || steps < 2) return AERO_E_ARG;
  B_HIP(b, hipSetDevice(b->dev));
  int rc = check_lockstep(b);
  if (rc) return rc;
  const int blocks[3] = {128, 256, 512};
  const int cpts[3] = {1, 2, 4};
  const int mults[5] = {2, 4, 8, 16, 0};  // one-launch grids of mult x CUs blocks, and 0: per-kind launches
  BatchShape best = b->shape;
  double best_ms = 1e30;
  for (int ci = 0; ci < 3; ++ci) {
    bool fits = true;
    for (aero_ctx* c : b->t) fits = fits && c->nx % (uint32_t)cpts[ci] == 0;
    if (!fits) continue;
    for (int bi = 0; bi < 3; ++bi)
      for (int mi = 0; mi < 5; ++mi) {
        b->shape = BatchShape{blocks[bi], cpts[ci], mults[mi] ? one_launch_grid(b->num_cu, mults[mi]) : 0u};
        rc = run(b, 2, nullptr, nullptr);
        if (rc) return rc;
        double o[2];
        rc = aero_batch_time_steps(b, steps & ~1, o);
        if (rc) return rc;
        if (o[1] < best_ms) {
          best_ms = o[1];
          best = b->shape;
        }
      }
  }
  b->shape = best;
  if (out) {
    out[0] = best.block;
    out[1] = best.cpt;
    out[2] = (int32_t)best.grid;
  }
  if (ms_out) *ms_out = best_ms;
  return AERO_OK;
}

}  // extern "C"
