// C ABI of the MI355X D3Q19 wind-tunnel solver (include/aerolbm.h): context lifecycle, geometry upload,
// parameters, stepping and readouts. The context owns all device memory and one HIP stream; every entry
// point reports failures through its return code and aero_last_error().
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/aerolbm.h"
#include "aero_ctx.h"
#include "geometry.h"
#include "lbm_launch.h"

using namespace aero;

static std::string g_create_error;

void aero::set_host_error(const std::string& msg) { g_create_error = msg; }

namespace {

template <typename P>
void dfree(P*& p) {
  if (p) hipFree((void*)p);
  p = nullptr;
}

int set_geo(aero_ctx* c) {
  Geo& g = c->geo;
  g.nx = c->nx; g.ny = c->ny; g.nz = c->nz; g.N = c->N;
  g.dnx = FastDiv::make(c->nx);
  g.dny = FastDiv::make(c->ny);
  {
    // bulk kernels that address the populations through a buffer descriptor use 32-bit offsets spanning
    // 2 lmax + 19 nx slots around a block's cells; larger cross-sections step with the per-cell kernel
    const int64_t esz = c->storage == AERO_STORE_FP16 ? 2 : 4;
    const int64_t lm = (int64_t)19 * c->nx * (1 + (int64_t)c->ny) + 1;
    g.bufok = !bulk_buffer_addressing((int)esz) || (2 * lm + (int64_t)19 * c->nx + 1024) * esz < ((int64_t)1 << 31) - 1
                  ? 1u : 0u;
  }
  g.lrow = (int64_t)19 * c->nx;
  g.lplane = (int64_t)19 * c->nx * c->ny;
  g.lmax = 0;
  for (int k = 0; k < 19; ++k) {
    g.loff[k] = cx(k) + (int64_t)19 * c->nx * (cy(k) + (int64_t)c->ny * cz(k));
    g.lmax = std::max(g.lmax, g.loff[k] < 0 ? -g.loff[k] : g.loff[k]);
  }
  g.hsz[FX0] = g.hsz[FX1] = c->ny * c->nz;
  g.hsz[FY0] = g.hsz[FY1] = c->nx * c->nz;
  g.hsz[FZ0] = g.hsz[FZ1] = c->nx * c->ny;
  uint32_t o = 0;
  for (int f = 0; f < 6; ++f) {
    g.hoff[f] = o;
    o += 5 * g.hsz[f];
  }
  c->halo_n = o;
  return 0;
}

void set_phys(aero_ctx* c) {
  Phys& ph = c->ph;
  const aero_params& p = c->p;
  ph.tau0 = (float)(3.0 * p.nu + 0.5);
  ph.smag = (float)(18.0 * M_SQRT2 * p.smagorinsky_cs * p.smagorinsky_cs);
  ph.u_target = p.u_inlet;
  ph.ramp = p.ramp_steps;
  ph.profile = p.inflow_profile == AERO_PROFILE_LOGLAW ? 2 : (p.ground_noslip ? 1 : 0);
  ph.noslip = p.ground_noslip ? 1 : 0;
  ph.sides = c->nz > 4 ? 1 : 0;
  ph.z0 = p.z0 > 0 ? p.z0 : 0.5;
  ph.zref = p.z_ref > 0 ? p.z_ref : 64.0;
  const int k = p.fp16_shift > 0 ? p.fp16_shift : 15;
  ph.nu = (float)p.nu;
  ph.scale = std::ldexp(1.0f, k);
  ph.inv_scale = std::ldexp(1.0f, -k);
  // inflow profile factor of each row (lbmCpuReference.ts:122-125 and 335-338 for the ground power law; the log
  // law u(y) = u ln((y + 1/2) / z0) / ln(z_ref / z0), zero below z0), applied to the inlet velocity of every step
  // and to the initial field
  c->uprof_h.assign(c->ny, 1.0);
  for (uint32_t iy = 0; iy < c->ny; ++iy) {
    if (ph.profile == 1) {
      const double yFrac = (iy + 0.5) / c->ny;
      c->uprof_h[iy] = std::fmin(1.15, 0.35 + 0.8 * std::pow(yFrac, 0.3));
    } else if (ph.profile == 2) {
      const double r = std::log((iy + 0.5) / ph.z0) / std::log(ph.zref / ph.z0);
      c->uprof_h[iy] = r > 0.0 ? r : 0.0;
    }
  }
}

// Copies the host profile table to the device, ordered after the steps already queued on the context's stream.
hipError_t upload_profile(aero_ctx* c) {
  return hipMemcpyAsync(c->uprof, c->uprof_h.data(), c->uprof_h.size() * sizeof(double), hipMemcpyHostToDevice,
                        c->stream);
}

}  // namespace

StepArgs aero_ctx::args(bool macros, bool populations) const {
  StepArgs a;
  std::memset(&a, 0, sizeof(a));
  a.g = geo;
  a.ph = ph;
  a.f = f;
  a.flags = flags;
  a.rowfluid = rowfluid;
  a.bmap = bmap;
  a.bmask = bmask;
  a.bq = bq;
  a.bown = bown;
  a.B = B;
  a.halo_in = halo[hp];
  a.halo_out = halo[hp ^ 1];
  a.state = st;
  a.partials = partials;
  a.mac = macros ? mac : nullptr;
  a.snap = macros && populations ? snap : nullptr;
  a.mlp = static_cast<const MlpDev*>(mlp);
  // a cell list built for another group width is not handed to the learned closure's split step, which then
  // steps every cell through its per-cell kernel
  const bool list_ok = !mlp || spec_v == 1;
  a.spec = list_ok ? spec : nullptr;
  a.nspec = list_ok ? nspec : 0;
  a.nface = list_ok ? nface : 0;
  a.uprof = uprof;
  return a;
}

// Special-cell list for the current group width, and partials room for the split step.
int aero_ctx::prepare_split() {
  const int vw = split_width();
  if (spec_v == vw) return AERO_OK;
  if (spec) {
    const int rc = quiesce();  // queued steps may still read the list replaced below
    if (rc != AERO_OK) return rc;
  }
  dfree(spec);
  nspec = nface = 0;
  spec_v = 0;
  if (vw == 0) {  // every cell through the per-cell kernel: no list, one partial per block (sized at create)
    nbulk = 0;
    return AERO_OK;
  }
  uint32_t nother = 0;
  AERO_HIP(this, build_special_list(flags, geo, vw, &spec, &nface, &nother, &nbulk, stream));
  nspec = nface + nother;
  const uint32_t need = split_blocks(N, vw, nface, nother);
  if (need > nblk) {
    AERO_HIP(this, hipStreamSynchronize(stream));
    dfree(partials);
    AERO_HIP(this, hipMalloc(&partials, (size_t)partials_len(need) * 8));
    nblk = need;
  }
  spec_v = vw;
  return AERO_OK;
}

int aero_ctx::upload_geometry_common() {
  // bmask from flags and blist; device buffers sized to B.
  dfree(bmask);
  dfree(bown);
  dfree(spec);
  nspec = nface = 0;
  spec_v = 0;
  if (B > 0) {
    AERO_HIP(this, hipMalloc(&bmask, (size_t)B * 4));
    AERO_HIP(this, hipMalloc(&bown, (size_t)B * 19 * 4));
    AERO_HIP(this, launch_links(flags, blist, B, geo, bmask, stream));
  }
  AERO_HIP(this, launch_row_fluid(flags, geo, rowfluid, stream));
  AERO_HIP(this, hipStreamSynchronize(stream));
  geom = true;
  init = false;
  mac_solids = false;
  invalidate_readout();
  return AERO_OK;
}

extern "C" {

const char* aero_last_error(const aero_ctx* ctx) { return ctx ? ctx->err.c_str() : g_create_error.c_str(); }

int aero_create(const aero_grid* grid, aero_ctx** out) {
  if (!grid || !out) { g_create_error = "aero_create: null argument"; return AERO_E_ARG; }
  *out = nullptr;
  if (grid->nx < 4 || grid->ny < 2 || grid->nz < 1) { g_create_error = "aero_create: grid must be at least 4x2x1"; return AERO_E_ARG; }
  const uint64_t N = (uint64_t)grid->nx * grid->ny * grid->nz;
  if (N >= (1ull << 31)) { g_create_error = "aero_create: grid exceeds 2^31 cells"; return AERO_E_ARG; }
  if (grid->storage != AERO_STORE_FP32 && grid->storage != AERO_STORE_FP16) {
    g_create_error = "aero_create: unknown storage";
    return AERO_E_ARG;
  }
  int ndev = 0;
  hipError_t e = hipGetDeviceCount(&ndev);
  if (e != hipSuccess || ndev == 0) {
    g_create_error = std::string("aero_create: no HIP device: ") + hipGetErrorString(e);
    return AERO_E_HIP;
  }
  if (grid->device < 0 || grid->device >= ndev) { g_create_error = "aero_create: bad device ordinal"; return AERO_E_ARG; }
  aero_ctx* c = new aero_ctx();
  c->dev = grid->device;
  c->nx = grid->nx; c->ny = grid->ny; c->nz = grid->nz; c->N = (uint32_t)N;
  c->storage = grid->storage;
  set_geo(c);
  aero_params p;
  std::memset(&p, 0, sizeof(p));
  p.u_inlet = 0.08; p.nu = 0.0005; p.smagorinsky_cs = 0.12; p.z0 = 0.5; p.z_ref = 64.0; p.u_init = -1.0;
  c->p = p;
  set_phys(c);
  auto fail = [&](const char* what, hipError_t err) {
    g_create_error = std::string("aero_create: ") + what + ": " + hipGetErrorString(err);
    aero_destroy(c);
    return AERO_E_HIP;
  };
  if ((e = hipSetDevice(c->dev)) != hipSuccess) return fail("hipSetDevice", e);
  {
    std::string msg;
    if (aero_noalloc_setup(c->dev, &msg) != AERO_OK) {
      g_create_error = "aero_create: " + msg;
      aero_destroy(c);
      return AERO_E_HIP;
    }
  }
  if ((e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking)) != hipSuccess) return fail("stream", e);
  if ((e = hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking)) != hipSuccess) return fail("stream", e);
  if ((e = hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming)) != hipSuccess) return fail("event", e);
  if ((e = hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming)) != hipSuccess) return fail("event", e);
  const size_t esz = c->storage == AERO_STORE_FP16 ? 2 : 4;
  if ((e = hipMalloc(&c->f, (size_t)19 * N * esz)) != hipSuccess) return fail("populations", e);
  if ((e = hipMalloc(&c->flags, N)) != hipSuccess) return fail("flags", e);
  if ((e = hipMalloc(&c->rowfluid, (size_t)grid->ny * grid->nz * 4)) != hipSuccess) return fail("rows", e);
  if ((e = hipMalloc(&c->bmap, N * 4)) != hipSuccess) return fail("bmap", e);
  for (int i = 0; i < 2; ++i)
    if ((e = hipMalloc(&c->halo[i], c->halo_n * 4)) != hipSuccess) return fail("halo", e);
  c->nblk = step_blocks(c->N);
  // cells per thread of the bulk kernel: pairs pay for the 2-byte FP16 populations, FP32 streams best one
  // cell per thread (fewest registers, highest occupancy)
  c->vw = (c->storage == AERO_STORE_FP16 && c->nx % 2 == 0) ? 2 : 1;
  if ((e = hipMalloc(&c->partials, (size_t)partials_len(c->nblk) * 8)) != hipSuccess) return fail("partials", e);
  if ((e = hipMalloc(&c->st, sizeof(StepState))) != hipSuccess) return fail("state", e);
  if ((e = hipMalloc(&c->fout, 8 * 8)) != hipSuccess) return fail("forces", e);
  if ((e = hipMalloc(&c->uprof, (size_t)c->ny * sizeof(double))) != hipSuccess) return fail("profile", e);
  if ((e = upload_profile(c)) != hipSuccess) return fail("profile", e);
  if ((e = hipMemsetAsync(c->halo[0], 0, c->halo_n * 4, c->stream)) != hipSuccess) return fail("memset", e);
  if ((e = hipMemsetAsync(c->halo[1], 0, c->halo_n * 4, c->stream)) != hipSuccess) return fail("memset", e);
  if ((e = hipStreamSynchronize(c->stream)) != hipSuccess) return fail("sync", e);
  *out = c;
  return AERO_OK;
}

int aero_destroy(aero_ctx* c) {
  if (!c) return AERO_OK;
  hipSetDevice(c->dev);
  if (c->stream) hipStreamSynchronize(c->stream);
  if (c->side) hipStreamSynchronize(c->side);
  c->drop_graphs();
  dfree(c->f); dfree(c->flags); dfree(c->rowfluid); dfree(c->bmap); dfree(c->blist); dfree(c->bq); dfree(c->bmask);
  dfree(c->bown); dfree(c->minq); dfree(c->halo[0]); dfree(c->halo[1]); dfree(c->partials); dfree(c->st);
  dfree(c->mac); dfree(c->snap); dfree(c->fpart); dfree(c->fout); dfree(c->spec);
  dfree(c->slice_buf); dfree(c->qbuf); dfree(c->qpart); dfree(c->probe_buf); dfree(c->mass_buf); dfree(c->uprof);
  c->vox.release();
  if (c->ev_fork) hipEventDestroy(c->ev_fork);
  if (c->ev_join) hipEventDestroy(c->ev_join);
  if (c->side) hipStreamDestroy(c->side);
  if (c->stream) hipStreamDestroy(c->stream);
  delete c;
  return AERO_OK;
}

int aero_set_voxels(aero_ctx* c, const uint8_t* flags, const int32_t* bmap, const int32_t* blist, const float* q,
                    int32_t nb) {
  if (!c) return AERO_E_ARG;
  AERO_REQUIRE(c, flags && bmap && (nb == 0 || (blist && q)) && nb >= 0, AERO_E_ARG, "aero_set_voxels: null array");
  AERO_HIP(c, hipSetDevice(c->dev));
  {
    const int rc = c->quiesce();  // the copies below go through the null stream
    if (rc != AERO_OK) return rc;
  }
  const uint32_t N = c->N;
  // validate on the host: flag values, boundary list consistency
  int32_t solid = 0, bnd = 0;
  for (uint32_t i = 0; i < N; ++i) {
    const uint8_t fl = flags[i];
    AERO_REQUIRE(c, fl <= 2, AERO_E_ARG, "aero_set_voxels: flag value out of range");
    if (fl == CELL_SOLID) ++solid;
    if (fl == CELL_BOUNDARY) {
      ++bnd;
      AERO_REQUIRE(c, bmap[i] >= 0 && bmap[i] < nb, AERO_E_ARG, "aero_set_voxels: boundary cell without index");
    }
  }
  AERO_REQUIRE(c, bnd == nb, AERO_E_ARG, "aero_set_voxels: boundary count does not match nb");
  for (int32_t b = 0; b < nb; ++b) {
    AERO_REQUIRE(c, blist[b] >= 0 && (uint32_t)blist[b] < N && flags[blist[b]] == CELL_BOUNDARY &&
                        bmap[blist[b]] == b,
                 AERO_E_ARG, "aero_set_voxels: boundary list inconsistent with bmap");
  }
  dfree(c->blist); dfree(c->bq); dfree(c->minq);
  c->B = (uint32_t)nb;
  AERO_HIP(c, hipMemcpy(c->flags, flags, N, hipMemcpyHostToDevice));
  AERO_HIP(c, hipMemcpy(c->bmap, bmap, (size_t)N * 4, hipMemcpyHostToDevice));
  c->h_flags.assign(flags, flags + N);
  std::vector<float> minq(nb, 1.0f);
  for (int32_t b = 0; b < nb; ++b)
    for (int d = 0; d < 18; ++d) {
      const float v = q[(size_t)b * 18 + d];
      if (v > 0.0f && v < minq[b]) minq[b] = v;
    }
  if (nb > 0) {
    AERO_HIP(c, hipMalloc(&c->blist, (size_t)nb * 4));
    AERO_HIP(c, hipMalloc(&c->bq, (size_t)nb * 18 * 4));
    AERO_HIP(c, hipMalloc(&c->minq, (size_t)nb * 4));
    AERO_HIP(c, hipMemcpy(c->blist, blist, (size_t)nb * 4, hipMemcpyHostToDevice));
    AERO_HIP(c, hipMemcpy(c->bq, q, (size_t)nb * 18 * 4, hipMemcpyHostToDevice));
    AERO_HIP(c, hipMemcpy(c->minq, minq.data(), (size_t)nb * 4, hipMemcpyHostToDevice));
  }
  c->counts[0] = solid;
  c->counts[1] = nb;
  c->counts[2] = (int32_t)(N - solid - nb);
  return c->upload_geometry_common();
}

int aero_read_voxels(aero_ctx* c, uint8_t* flags, int32_t* bmap, int32_t* blist, float* q, float* minq,
                     int32_t* counts) {
  if (!c) return AERO_E_ARG;
  AERO_REQUIRE(c, c->geom, AERO_E_STATE, "aero_read_voxels: no geometry");
  AERO_HIP(c, hipSetDevice(c->dev));
  AERO_HIP(c, hipStreamSynchronize(c->stream));
  if (counts) std::memcpy(counts, c->counts, sizeof(c->counts));
  if (flags) AERO_HIP(c, hipMemcpy(flags, c->flags, c->N, hipMemcpyDeviceToHost));
  if (bmap) AERO_HIP(c, hipMemcpy(bmap, c->bmap, (size_t)c->N * 4, hipMemcpyDeviceToHost));
  if (c->B > 0) {
    if (blist) AERO_HIP(c, hipMemcpy(blist, c->blist, (size_t)c->B * 4, hipMemcpyDeviceToHost));
    if (q) AERO_HIP(c, hipMemcpy(q, c->bq, (size_t)c->B * 18 * 4, hipMemcpyDeviceToHost));
    if (minq) AERO_HIP(c, hipMemcpy(minq, c->minq, (size_t)c->B * 4, hipMemcpyDeviceToHost));
  }
  return AERO_OK;
}

int aero_voxelize_ms(aero_ctx* c, double* ms) {
  if (!c || !ms) return AERO_E_ARG;
  *ms = c->vox_ms;
  return AERO_OK;
}

int aero_set_params(aero_ctx* c, const aero_params* p) {
  if (!c || !p) return AERO_E_ARG;
  AERO_REQUIRE(c, std::isfinite(p->u_inlet) && std::isfinite(p->nu) && p->nu >= 0 && std::isfinite(p->smagorinsky_cs),
               AERO_E_ARG, "aero_set_params: non-finite or negative parameter");
  AERO_REQUIRE(c, p->ramp_steps >= 0, AERO_E_ARG, "aero_set_params: negative ramp");
  AERO_REQUIRE(c, p->closure == AERO_CLOSURE_SMAGORINSKY || p->closure == AERO_CLOSURE_MLP, AERO_E_ARG,
               "aero_set_params: unknown closure");
  AERO_REQUIRE(c, p->closure == AERO_CLOSURE_SMAGORINSKY || c->mlp_src, AERO_E_STATE,
               "aero_set_params: the MLP closure needs aero_set_closure first");
  // The codec stores half((f - w) 2^k). Values at or above 65504 2^-k become infinite: k <= 16 keeps 1.0 in
  // range, about four times the largest equilibrium |f - w| the moment clamps allow (rest population,
  // rho = 1.8). Below the normal range, f - w is a multiple of FP32 f's ulp (>= 2^-29, since f is near
  // w >= 1/36), which half subnormals hold exactly for k >= 5; the lower limit 10 leaves margin.
  AERO_REQUIRE(c, p->fp16_shift == 0 || (p->fp16_shift >= 10 && p->fp16_shift <= 16), AERO_E_ARG,
               "aero_set_params: fp16_shift must be 0 (default 15) or 10..16");
  AERO_REQUIRE(c, p->inflow_profile == AERO_PROFILE_UNIFORM || p->inflow_profile == AERO_PROFILE_LOGLAW, AERO_E_ARG,
               "aero_set_params: unknown inflow profile");
  const bool codec_change = c->storage == AERO_STORE_FP16 && p->fp16_shift != c->p.fp16_shift;
  c->p = *p;
  c->mlp = p->closure == AERO_CLOSURE_MLP ? aero_mlp_device_weights(c->mlp_src) : nullptr;
  set_phys(c);
  AERO_HIP(c, hipSetDevice(c->dev));
  AERO_HIP(c, upload_profile(c));
  if (codec_change) c->init = false;  // stored halves were encoded with the old scale
  return AERO_OK;
}

int aero_reset(aero_ctx* c, uint64_t seed, double noise_amp) {
  if (!c) return AERO_E_ARG;
  AERO_REQUIRE(c, c->geom, AERO_E_STATE, "aero_reset: no geometry");
  AERO_REQUIRE(c, noise_amp >= 0 && std::isfinite(noise_amp), AERO_E_ARG, "aero_reset: bad noise amplitude");
  AERO_HIP(c, hipSetDevice(c->dev));
  StepState s;
  std::memset(&s, 0, sizeof(s));
  s.corr = 1.0;
  s.nfluid = (long long)c->N - c->counts[0];
  AERO_HIP(c, hipMemcpyAsync(c->st, &s, sizeof(s), hipMemcpyHostToDevice, c->stream));
  c->hp = 1;  // k_init writes halo[hp ^ 1] = halo[0]
  StepArgs a = c->args(false);
  const double u0 = c->p.u_init >= 0 ? c->p.u_init : c->p.u_inlet;
  AERO_HIP(c, launch_init(a, c->storage == AERO_STORE_FP16, u0, seed, noise_amp, c->stream));
  c->hp = 0;
  c->aa = 0;
  c->steps = 0;
  AERO_HIP(c, hipStreamSynchronize(c->stream));
  c->init = true;
  c->invalidate_readout();
  return AERO_OK;
}

namespace {
// Enqueues one step (split kernels, reduction) from the context's current A-A state and advances it; ro: it writes
// the macro fields, pop: also the population snapshot.
int enqueue_step(aero_ctx* c, bool fp16, bool ro, bool pop = false) {
  StepArgs a = c->args(ro, pop);
  uint32_t nb = 0;
  const SideLaunch side{c->side, c->ev_fork, c->ev_join};
  AERO_HIP(c, launch_step(a, fp16, c->split_width(), c->aa == 0, ro, c->stream, &nb, nullptr, &side));
  AERO_HIP(c, launch_finalize(c->partials, nb, c->st, c->stream));
  c->aa ^= 1;
  c->hp ^= 1;
  ++c->steps;
  return AERO_OK;
}

// Captures a pair of steps starting from A-A state `state` (with their side-stream fork and join) into a graph:
// two plain steps, or with `ro` a plain step and then a readout step (`pop`: with the population snapshot).
// Capture records the launches without running them, and the context's step state is restored after.
// *exec stays null when capture is unavailable.
int capture_pair(aero_ctx* c, bool fp16, int state, hipGraphExec_t* exec, bool ro = false, bool pop = false) {
  *exec = nullptr;
  if (hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal) != hipSuccess) {
    (void)hipGetLastError();
    return AERO_OK;
  }
  const int aa0 = c->aa, hp0 = c->hp;
  const long long s0 = c->steps;
  c->aa = c->hp = state;  // the two flip together on every step
  int rc = enqueue_step(c, fp16, false);
  if (rc == AERO_OK) rc = enqueue_step(c, fp16, ro, pop);
  c->aa = aa0;
  c->hp = hp0;
  c->steps = s0;
  hipGraph_t g = nullptr;
  const hipError_t ec = hipStreamEndCapture(c->stream, &g);
  if (rc != AERO_OK || ec != hipSuccess || !g) {
    if (g) (void)hipGraphDestroy(g);
    (void)hipGetLastError();
    return rc;  // capture failed: nothing was enqueued
  }
  const hipError_t e = hipGraphInstantiate(exec, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) {
    *exec = nullptr;
    (void)hipGetLastError();
  }
  return AERO_OK;
}

// Makes sure both step-pair graphs match the context's current arguments, capturing again what does not (for
// instance after new parameters, geometry, closure or group width). Both are captured together, so a first run of
// an odd number of steps leaves the graph for the other A-A state ready too.
int refresh_pair_graphs(aero_ctx* c, bool fp16) {
  const int width = c->split_width();
  for (int state = 0; state < 2; ++state) {
    StepPairGraph& pg = c->pair_graph[state];
    const int aa0 = c->aa, hp0 = c->hp;
    c->aa = c->hp = state;
    const StepArgs key = c->args(false);
    c->aa = aa0;
    c->hp = hp0;
    if (pg.exec && pg.width == width && std::memcmp(&pg.key, &key, sizeof(key)) == 0) continue;
    if (pg.exec) {
      AERO_HIP(c, hipStreamSynchronize(c->stream));  // no replay of the old graph may still be in flight
      (void)hipGraphExecDestroy(pg.exec);
      pg.exec = nullptr;
    }
    const int rc = capture_pair(c, fp16, state, &pg.exec);
    if (rc != AERO_OK) return rc;
    pg.key = key;
    pg.width = width;
  }
  return AERO_OK;
}

// Runs a plain step and then a readout step from the current A-A state through a captured graph (the last two steps
// of a frame's substeps), capturing it first where its arguments changed. *done is 0 when no graph is available.
int readout_pair_graph(aero_ctx* c, bool fp16, bool pop, int32_t* done) {
  *done = 0;
  const int state = c->aa;
  StepPairGraph& pg = c->ro_graph[state][pop ? 1 : 0];
  const int width = c->split_width();
  const int aa0 = c->aa, hp0 = c->hp;
  const StepArgs key = c->args(false);
  c->aa ^= 1;
  c->hp ^= 1;
  const StepArgs key2 = c->args(true, pop);
  c->aa = aa0;
  c->hp = hp0;
  if (!(pg.exec && pg.width == width && std::memcmp(&pg.key, &key, sizeof(key)) == 0 &&
        std::memcmp(&pg.key2, &key2, sizeof(key2)) == 0)) {
    if (pg.exec) {
      AERO_HIP(c, hipStreamSynchronize(c->stream));  // no replay of the old graph may still be in flight
      (void)hipGraphExecDestroy(pg.exec);
      pg.exec = nullptr;
    }
    const int rc = capture_pair(c, fp16, state, &pg.exec, true, pop);
    if (rc != AERO_OK) return rc;
    pg.key = key;
    pg.key2 = key2;
    pg.width = width;
  }
  if (!pg.exec) return AERO_OK;
  const hipError_t e = hipGraphLaunch(pg.exec, c->stream);
  if (e != hipSuccess) {
    c->err = std::string("graph launch: ") + hipGetErrorString(e);
    return AERO_E_HIP;
  }
  *done = 2;
  c->steps += 2;
  return AERO_OK;
}

// Runs an even number m of plain steps by replaying the captured pair for the current A-A state m/2 times: the
// A-A state, the halo parity and every kernel argument repeat with period two. *done is 0 when no graph is
// available, and the caller then launches the steps directly.
int step_pairs_graph(aero_ctx* c, bool fp16, int32_t m, int32_t* done) {
  *done = 0;
  const int rc = refresh_pair_graphs(c, fp16);
  if (rc != AERO_OK) return rc;
  hipGraphExec_t ge = c->pair_graph[c->aa].exec;
  if (!ge) return AERO_OK;
  hipError_t e = hipSuccess;
  for (int32_t p = 0; p < m / 2 && e == hipSuccess; ++p) e = hipGraphLaunch(ge, c->stream);
  if (e != hipSuccess) {
    c->err = std::string("graph launch: ") + hipGetErrorString(e);
    return AERO_E_HIP;
  }
  *done = 2 * (m / 2);
  c->steps += *done;
  return AERO_OK;
}
}  // namespace

int aero_step(aero_ctx* c, int32_t n, int32_t flags) {
  if (!c) return AERO_E_ARG;
  AERO_REQUIRE(c, n >= 0, AERO_E_ARG, "aero_step: negative count");
  AERO_REQUIRE(c, c->init, AERO_E_STATE, "aero_step: call aero_reset after setting geometry");
  AERO_HIP(c, hipSetDevice(c->dev));
  AERO_REQUIRE(c, (flags & ~(AERO_STEP_READOUT | AERO_STEP_MACROS)) == 0, AERO_E_ARG, "aero_step: unknown flags");
  const bool pop = (flags & AERO_STEP_READOUT) != 0;
  const bool want = pop || (flags & AERO_STEP_MACROS) != 0;
  if (want) {
    const int rc = c->ensure_readout(pop);
    if (rc != AERO_OK) return rc;
  }
  if (n > 0) {
    const int rc = c->prepare_split();
    if (rc != AERO_OK) return rc;
  }
  if (want && n > 0 && !c->mac_solids) {
    AERO_HIP(c, launch_solid_macros(c->flags, c->N, c->mac, c->stream));
    c->mac_solids = true;
  }
  const bool fp16 = c->storage == AERO_STORE_FP16;
  const int32_t plain = want ? n - 1 : n;
  int32_t done = 0;
  // runs of plain steps replay a captured pair; with a readout and an odd number of plain steps before it, the
  // last plain step and the readout step replay a captured pair of their own
  const bool ro_pair = want && (plain & 1);
  const int32_t pairs = ro_pair ? plain - 1 : plain;
  if (pairs >= 2) {
    const int rc = step_pairs_graph(c, fp16, pairs, &done);
    if (rc != AERO_OK) return rc;
  }
  if (ro_pair && done == pairs) {
    int32_t d = 0;
    const int rc = readout_pair_graph(c, fp16, pop, &d);
    if (rc != AERO_OK) return rc;
    done += d;
  }
  for (int32_t s = done; s < n; ++s) {
    const int rc = enqueue_step(c, fp16, want && s == n - 1, pop);
    if (rc != AERO_OK) return rc;
  }
  if (n > 0) {
    c->readout_valid = want;
    c->pops_valid = pop;
  }
  return AERO_OK;
}

int aero_sync(aero_ctx* c) {
  if (!c) return AERO_E_ARG;
  AERO_HIP(c, hipSetDevice(c->dev));
  AERO_HIP(c, hipStreamSynchronize(c->stream));
  return AERO_OK;
}

int aero_step_index(aero_ctx* c, int64_t* out) {
  if (!c || !out) return AERO_E_ARG;
  *out = c->steps;
  return AERO_OK;
}

int aero_set_step_index(aero_ctx* c, int64_t step) {
  if (!c) return AERO_E_ARG;
  AERO_REQUIRE(c, step >= 0, AERO_E_ARG, "aero_set_step_index: negative step");
  AERO_REQUIRE(c, c->init, AERO_E_STATE, "aero_set_step_index: call aero_reset or aero_write_populations first");
  AERO_HIP(c, hipSetDevice(c->dev));
  StepState s;
  AERO_HIP(c, hipMemcpyAsync(&s, c->st, sizeof(s), hipMemcpyDeviceToHost, c->stream));
  AERO_HIP(c, hipStreamSynchronize(c->stream));
  s.step = step;
  AERO_HIP(c, hipMemcpyAsync(c->st, &s, sizeof(s), hipMemcpyHostToDevice, c->stream));
  AERO_HIP(c, hipStreamSynchronize(c->stream));
  c->steps = step;
  return AERO_OK;
}

int aero_set_closure(aero_ctx* c, const aero_mlp* m) {
  if (!c) return AERO_E_ARG;
  if (m) AERO_REQUIRE(c, aero_mlp_device(m) == c->dev, AERO_E_ARG, "aero_set_closure: device mismatch");
  c->mlp_src = m;
  c->p.closure = m ? AERO_CLOSURE_MLP : AERO_CLOSURE_SMAGORINSKY;
  c->mlp = m ? aero_mlp_device_weights(m) : nullptr;
  return AERO_OK;
}

int aero_set_cells_per_thread(aero_ctx* c, int32_t v) {
  if (!c) return AERO_E_ARG;
  AERO_REQUIRE(c, v == 0 || ((v == 1 || v == 2 || v == 4) && c->nx % v == 0), AERO_E_ARG,
               "aero_set_cells_per_thread: 0 (no bulk kernel), or 1, 2 or 4, dividing nx");
  c->vw = v;
  return AERO_OK;
}

int aero_split_info(aero_ctx* c, int64_t* out) {
  if (!c || !out) return AERO_E_ARG;
  AERO_REQUIRE(c, c->geom, AERO_E_STATE, "aero_split_info: no geometry");
  AERO_HIP(c, hipSetDevice(c->dev));
  const int rc = c->prepare_split();
  if (rc != AERO_OK) return rc;
  out[0] = c->nbulk;
  out[1] = c->split_width() ? c->nspec : c->N;
  out[2] = c->split_width();
  out[3] = c->nface;
  return AERO_OK;
}

int aero_anchor_count(aero_ctx* c, int64_t* out) {
  if (!c || !out) return AERO_E_ARG;
  AERO_HIP(c, hipSetDevice(c->dev));
  StepState s;
  AERO_HIP(c, hipMemcpyAsync(&s, c->st, sizeof(s), hipMemcpyDeviceToHost, c->stream));
  AERO_HIP(c, hipStreamSynchronize(c->stream));
  *out = s.anchors;
  return AERO_OK;
}

int aero_time_steps(aero_ctx* c, int32_t n, double* out) {
  if (!c || !out) return AERO_E_ARG;
  AERO_REQUIRE(c, n > 0, AERO_E_ARG, "aero_time_steps: n must be positive");
  AERO_REQUIRE(c, c->init, AERO_E_STATE, "aero_time_steps: call aero_reset first");
  AERO_HIP(c, hipSetDevice(c->dev));
  {
    const int rc = c->prepare_split();
    if (rc != AERO_OK) return rc;
  }
  const bool fp16 = c->storage == AERO_STORE_FP16;
  std::vector<hipEvent_t> ev(2 * (size_t)n + 2), mid(n);
  for (auto& e : ev) AERO_HIP(c, hipEventCreate(&e));
  for (auto& e : mid) AERO_HIP(c, hipEventCreate(&e));
  AERO_HIP(c, hipEventRecord(ev[0], c->stream));
  for (int32_t s = 0; s < n; ++s) {
    StepArgs a = c->args(false);
    AERO_HIP(c, hipEventRecord(ev[1 + 2 * s], c->stream));
    uint32_t nb = 0;
    const SideLaunch side{c->side, c->ev_fork, c->ev_join};
    AERO_HIP(c, launch_step(a, fp16, c->split_width(), c->aa == 0, false, c->stream, &nb, mid[s], &side));
    AERO_HIP(c, hipEventRecord(ev[2 + 2 * s], c->stream));
    AERO_HIP(c, launch_finalize(c->partials, nb, c->st, c->stream));
    c->aa ^= 1;
    c->hp ^= 1;
    ++c->steps;
  }
  AERO_HIP(c, hipEventRecord(ev[2 * n + 1], c->stream));
  AERO_HIP(c, hipEventSynchronize(ev[2 * n + 1]));
  float tot = 0.f;
  AERO_HIP(c, hipEventElapsedTime(&tot, ev[0], ev[2 * n + 1]));
  double ksum = 0.0, bsum = 0.0;
  for (int32_t s = 0; s < n; ++s) {
    float t = 0.f, tb = 0.f;
    AERO_HIP(c, hipEventElapsedTime(&t, ev[1 + 2 * s], ev[2 + 2 * s]));
    AERO_HIP(c, hipEventElapsedTime(&tb, ev[1 + 2 * s], mid[s]));
    ksum += t;
    bsum += tb;
  }
  for (auto& e : ev) (void)hipEventDestroy(e);
  for (auto& e : mid) (void)hipEventDestroy(e);
  out[0] = tot;
  out[1] = ksum / n;
  out[2] = (double)tot / n;
  out[3] = bsum / n;
  out[4] = (ksum - bsum) / n;
  c->invalidate_readout();
  return AERO_OK;
}

int aero_read_forces(aero_ctx* c, const aero_refs* refs, aero_forces* out) {
  if (!c || !out) return AERO_E_ARG;
  AERO_REQUIRE(c, c->init, AERO_E_STATE, "aero_read_forces: call aero_reset first");
  AERO_HIP(c, hipSetDevice(c->dev));
  const uint32_t nblk = forces_blocks(c->B);
  if (nblk > c->fpart_cap) {
    dfree(c->fpart);
    AERO_HIP(c, hipMalloc(&c->fpart, (size_t)nblk * 8 * 8));
    c->fpart_cap = nblk;
  }
  StepArgs a = c->args(false);
  AERO_HIP(c, launch_forces(a, c->storage == AERO_STORE_FP16, c->blist, c->aa, c->halo[c->hp], c->fpart, c->fout,
                            c->stream));
  double h[8];
  AERO_HIP(c, hipMemcpyAsync(h, c->fout, 7 * 8, hipMemcpyDeviceToHost, c->stream));
  AERO_HIP(c, hipStreamSynchronize(c->stream));
  std::memset(out, 0, sizeof(*out));
  out->fx = h[0]; out->fy = h[1]; out->fz = h[2]; out->mx = h[3]; out->my = h[4]; out->mz = h[5];
  out->cut_face_links = (int64_t)std::llround(h[6]);
  if (refs) aero_coefficients(out, refs);
  return AERO_OK;
}

int aero_read_macro(aero_ctx* c, float* ux, float* uy, float* uz, float* rho) {
  if (!c) return AERO_E_ARG;
  AERO_REQUIRE(c, c->readout_valid, AERO_E_STATE,
               "aero_read_macro: last aero_step had neither AERO_STEP_READOUT nor AERO_STEP_MACROS");
  AERO_HIP(c, hipSetDevice(c->dev));
  AERO_HIP(c, hipStreamSynchronize(c->stream));
  const size_t N = c->N;
  float* dst[4] = {ux, uy, uz, rho};
  for (int j = 0; j < 4; ++j)
    if (dst[j]) AERO_HIP(c, hipMemcpy(dst[j], c->mac + j * N, N * 4, hipMemcpyDeviceToHost));
  if (rho) {
    StepState s;
    AERO_HIP(c, hipMemcpy(&s, c->st, sizeof(s), hipMemcpyDeviceToHost));
    if (s.corr != 1.0)
      for (size_t i = 0; i < N; ++i)
        if (c->h_flags[i] != CELL_SOLID) rho[i] = (float)((double)rho[i] * s.corr);
  }
  return AERO_OK;
}

int aero_read_populations(aero_ctx* c, float* f) {
  if (!c || !f) return AERO_E_ARG;
  AERO_REQUIRE(c, c->pops_valid, AERO_E_STATE, "aero_read_populations: last aero_step had no AERO_STEP_READOUT");
  AERO_HIP(c, hipSetDevice(c->dev));
  AERO_HIP(c, hipStreamSynchronize(c->stream));
  const size_t N = c->N;
  AERO_HIP(c, hipMemcpy(f, c->snap, 19 * N * 4, hipMemcpyDeviceToHost));
  StepState s;
  AERO_HIP(c, hipMemcpy(&s, c->st, sizeof(s), hipMemcpyDeviceToHost));
  for (size_t i = 0; i < N; ++i) {
    if (c->h_flags[i] == CELL_SOLID) {
      for (int k = 0; k < 19; ++k) f[k * N + i] = w32(k);
    } else if (s.corr != 1.0) {
      for (int k = 0; k < 19; ++k) f[k * N + i] = (float)((double)f[k * N + i] * s.corr);
    }
  }
  return AERO_OK;
}

int aero_write_populations(aero_ctx* c, const float* f) {
  if (!c || !f) return AERO_E_ARG;
  AERO_REQUIRE(c, c->geom, AERO_E_STATE, "aero_write_populations: no geometry");
  AERO_HIP(c, hipSetDevice(c->dev));
  {
    const int rc = c->quiesce();  // pending steps read and write the state the copies below replace
    if (rc != AERO_OK) return rc;
  }
  {
    const int rc = c->ensure_readout(true);  // the snapshot buffer stages the upload
    if (rc != AERO_OK) return rc;
  }
  const size_t N = c->N;
  AERO_HIP(c, hipMemcpy(c->snap, f, 19 * N * 4, hipMemcpyHostToDevice));
  if (!c->init) {
    StepState s;
    std::memset(&s, 0, sizeof(s));
    s.nfluid = (long long)c->N - c->counts[0];
    s.corr = 1.0;
    AERO_HIP(c, hipMemcpy(c->st, &s, sizeof(s), hipMemcpyHostToDevice));
    c->steps = 0;
  } else {
    // keep the step counter; the loaded values already include any anchor factor
    StepState s;
    AERO_HIP(c, hipMemcpy(&s, c->st, sizeof(s), hipMemcpyDeviceToHost));
    s.corr = 1.0;
    AERO_HIP(c, hipMemcpy(c->st, &s, sizeof(s), hipMemcpyHostToDevice));
  }
  c->hp = 1;  // k_load writes halo[hp ^ 1] = halo[0]
  StepArgs a = c->args(false);
  AERO_HIP(c, launch_load(a, c->storage == AERO_STORE_FP16, c->snap, c->stream));
  c->hp = 0;
  c->aa = 0;
  AERO_HIP(c, hipStreamSynchronize(c->stream));
  c->init = true;
  c->invalidate_readout();
  return AERO_OK;
}

}  // extern "C"

int aero_ctx::ensure_readout(bool populations) {
  if (!mac) {
    hipError_t e = hipMalloc(&mac, (size_t)4 * N * 4);
    if (e != hipSuccess) { err = std::string("readout buffer: ") + hipGetErrorString(e); mac = nullptr; return AERO_E_NOMEM; }
    mac_solids = false;
  }
  if (populations && !snap) {
    hipError_t e = hipMalloc(&snap, (size_t)19 * N * 4);
    if (e != hipSuccess) { err = std::string("snapshot buffer: ") + hipGetErrorString(e); snap = nullptr; return AERO_E_NOMEM; }
  }
  return AERO_OK;
}

int aero_ctx::quiesce() {
  AERO_HIP(this, hipSetDevice(dev));
  AERO_HIP(this, hipStreamSynchronize(side));
  AERO_HIP(this, hipStreamSynchronize(stream));
  return AERO_OK;
}

void aero_ctx::drop_graphs() {
  StepPairGraph* all[6] = {&pair_graph[0], &pair_graph[1], &ro_graph[0][0], &ro_graph[0][1], &ro_graph[1][0],
                           &ro_graph[1][1]};
  for (StepPairGraph* pg : all) {
    if (pg->exec) {
      if (stream) (void)hipStreamSynchronize(stream);
      (void)hipGraphExecDestroy(pg->exec);
    }
    pg->exec = nullptr;
    pg->width = 0;
  }
}

// computeNormalizedForceCoefficients (forceIntegrator.ts:240-315).
void aero_coefficients(aero_forces* o, const aero_refs* r) {
  const double dx = r->dx;
  const double cellArea = dx * dx;
  const double qLat = std::max(1e-9, 0.5 * 1.0 * r->u_ref * r->u_ref);
  const double aRefLat = std::max(1e-4, r->frontal_area / cellArea);
  const double lRefLat = std::max(1.0, r->char_length / dx);
  o->cd_raw = o->fx / (qLat * aRefLat);
  o->cl = o->fy / (qLat * aRefLat);
  o->cm = o->mz / (qLat * aRefLat * lRefLat);
  const double tunnelArea = std::max(1e-4, r->tunnel_width * r->tunnel_height);
  const double blockage = std::min(0.40, r->frontal_area / tunnelArea);
  const double maskell = 1.0 + 2.5 * std::max(0.0, o->cd_raw) * blockage;
  o->cd = o->cd_raw / maskell;
  o->xcp = std::fabs(o->fy) > 1e-5
               ? std::max(-r->char_length * 0.5, std::min(r->char_length * 0.5, -(o->mz / o->fy) * dx))
               : 0.0;
}
